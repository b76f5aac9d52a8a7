// Circuit description and finalization, mirroring the reference's Circuit
// (src/core/circuit.hpp:157-418) for the device families on the Newton hot
// path: R, C, V, I, level-1 diode, MOS1 and BSIM4v7.
//
// finalize() reproduces the reference's numbering and ordering exactly,
// because they fix the MNA unknowns and the floating-point summation order
// of every matrix slot:
//   * load order: ngspice family rank (circuit.cpp:34-59), then newest
//     instance first for devices built through the typed API (:79-97);
//   * branch unknowns assigned in load order after all nodes (:213-223);
//   * pattern = every device stamp + a diagonal for each structurally
//     isolated node (:225-257);
//   * state slots assigned in load order (:272-280).
// It then lowers every device into "stamp sites": one value per stamp call
// in the device's own code order.  A slot's value is the sum of its sites
// in load order, which is exactly NumericMatrix::add's accumulation.
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace nsa {

enum DevKind : int32_t {
    DK_R = 0, DK_C = 1, DK_V = 2, DK_I = 3, DK_D = 4, DK_M1 = 5, DK_B4 = 6, DK_COUNT = 7
};
enum Wave : int32_t { WAVE_DC = 0, WAVE_PULSE = 1, WAVE_PWL = 2 };

struct ModelCard {
    std::string name, type;  // type: "d", "nmos", "pmos"
    int level = 1;
    std::map<std::string, double> p;  // lower-case parameter names
    bool has(const std::string& k) const { return p.count(k) != 0; }
    double get(const std::string& k, double dflt) const {
        auto it = p.find(k);
        return it == p.end() ? dflt : it->second;
    }
};

// Parses ".model <name> <type> [(] k=v ... [)]" with SPICE number suffixes.
ModelCard parse_model_card(const std::string& text);
double parse_spice_number(const std::string& s);

struct Device {
    DevKind kind;
    std::string name;
    int32_t insertion = 0;
    int32_t node[4] = {-1, -1, -1, -1};  // MNA index, -1 = ground
    double value = 0.0;                  // R ohms, C farads, V/I dc
    int32_t wave = WAVE_DC;
    double pulse[7] = {0, 0, 0, -1, -1, -1, -1};  // v1 v2 td tr tf pw per
    std::vector<double> pwl_t, pwl_v;
    int32_t model = -1;
    double w = 0.0, l = 0.0;
    bool w_given = false, l_given = false;
    // finalize() outputs
    int32_t branch = -1;
    int32_t state_base = -1;
    int32_t family_index = -1;     // index within its family (load order)
    std::vector<int32_t> mat_slot;  // per local matrix stamp, -1 for ground
    std::vector<int32_t> rhs_row;   // per local rhs stamp, -1 for ground
};

int32_t load_rank(DevKind k);
int32_t state_count(DevKind k);

class Circuit {
public:
    // Node by name; "0", "gnd", "GND" are ground (-1).
    int32_t node(const std::string& name);
    int32_t find_node(const std::string& name) const;

    int32_t R(const std::string& name, int32_t a, int32_t b, double ohms);
    int32_t C(const std::string& name, int32_t a, int32_t b, double farads);
    int32_t V(const std::string& name, int32_t p, int32_t n, double dc);
    int32_t V_pulse(const std::string& name, int32_t p, int32_t n, const double pulse[7]);
    int32_t V_pwl(const std::string& name, int32_t p, int32_t n, const std::vector<double>& t,
                  const std::vector<double>& v);
    int32_t I(const std::string& name, int32_t p, int32_t n, double dc);
    int32_t D(const std::string& name, int32_t a, int32_t k, const std::string& model);
    int32_t M(const std::string& name, int32_t d, int32_t g, int32_t s, int32_t b,
              const std::string& model, double w, double l);
    int32_t model(const std::string& text);
    int32_t find_model(const std::string& name) const;

    void finalize();
    bool finalized() const { return finalized_; }

    int32_t num_nodes() const { return next_node_; }
    int32_t num_vars() const { return num_vars_; }
    int32_t num_states() const { return num_states_; }
    int32_t nnz() const { return (int32_t)row_idx_.size(); }
    bool is_linear() const;
    const std::vector<int32_t>& col_ptr() const { return col_ptr_; }
    const std::vector<int32_t>& row_idx() const { return row_idx_; }
    const std::vector<int32_t>& load_order() const { return load_order_; }
    const std::vector<int32_t>& dead_nodes() const { return dead_nodes_; }
    int32_t offset(int32_t r, int32_t c) const;  // -1 if ground or absent
    std::vector<Device>& devices() { return devs_; }
    const std::vector<Device>& devices() const { return devs_; }
    const std::vector<ModelCard>& models() const { return models_; }
    // Devices of one family in load order (indices into devices()).
    const std::vector<int32_t>& family(DevKind k) const { return fam_[k]; }

private:
    int32_t add(Device d);
    std::map<std::string, int32_t> node_map_;
    std::vector<std::string> node_names_;
    std::vector<Device> devs_;
    std::vector<ModelCard> models_;
    std::vector<int32_t> load_order_, dead_nodes_;
    std::vector<int32_t> col_ptr_, row_idx_;
    std::vector<int32_t> fam_[DK_COUNT];
    int32_t next_node_ = 0, num_vars_ = 0, num_states_ = 0;
    bool finalized_ = false;
};

}  // namespace nsa
