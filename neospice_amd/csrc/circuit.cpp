#include "circuit.hpp"

#include "bsim4.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <sstream>

#include "common.hpp"
#include "devices.hpp"

namespace nsa {

namespace {
std::string lower(std::string s) {
    for (auto& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}
}  // namespace

double parse_spice_number(const std::string& tok) {
    const std::string s = lower(tok);
    size_t i = 0;
    double v = std::stod(s, &i);
    const std::string suf = s.substr(i);
    if (suf.empty()) return v;
    if (suf.rfind("meg", 0) == 0) return v * 1e6;
    if (suf.rfind("mil", 0) == 0) return v * 25.4e-6;
    switch (suf[0]) {
        case 't': return v * 1e12;
        case 'g': return v * 1e9;
        case 'k': return v * 1e3;
        case 'm': return v * 1e-3;
        case 'u': return v * 1e-6;
        case 'n': return v * 1e-9;
        case 'p': return v * 1e-12;
        case 'f': return v * 1e-15;
        case 'a': return v * 1e-18;
        default: return v;  // unit letters (v, a, s, ...) carry no scale
    }
}

ModelCard parse_model_card(const std::string& text) {
    std::string t = text;
    for (auto& c : t) if (c == '(' || c == ')' || c == ',') c = ' ';
    // tokenise, keeping k=v pairs together
    std::vector<std::string> tok;
    {
        std::string cur;
        for (size_t i = 0; i < t.size(); ++i) {
            char c = t[i];
            if (std::isspace((unsigned char)c)) {
                if (!cur.empty()) { tok.push_back(cur); cur.clear(); }
            } else if (c == '=') {
                if (!cur.empty()) { tok.push_back(cur); cur.clear(); }
                tok.push_back("=");
            } else cur.push_back(c);
        }
        if (!cur.empty()) tok.push_back(cur);
    }
    size_t i = 0;
    if (i < tok.size() && lower(tok[i]) == ".model") ++i;
    if (i + 2 > tok.size()) throw Error(NSA_ERR_ARGUMENT, "model card: expected name and type");
    ModelCard m;
    m.name = lower(tok[i++]);
    m.type = lower(tok[i++]);
    while (i < tok.size()) {
        if (i + 1 < tok.size() && tok[i + 1] == "=") {
            if (i + 2 >= tok.size()) throw Error(NSA_ERR_ARGUMENT, "model card: dangling '='");
            m.p[lower(tok[i])] = parse_spice_number(tok[i + 2]);
            i += 3;
        } else {
            ++i;  // flag without value
        }
    }
    if (m.has("level")) m.level = (int)m.get("level", 1);
    return m;
}

// ngspice static device table ranks (src/core/circuit.cpp:34-59).
int32_t load_rank(DevKind k) {
    switch (k) {
        case DK_M1: return 3;
        case DK_B4: return 3;   // "M" family, like MOS1
        case DK_C: return 17;
        case DK_D: return 22;
        case DK_I: return 31;
        case DK_R: return 41;
        case DK_V: return 50;
        default: return 1000;
    }
}

int32_t state_count(DevKind k) {
    switch (k) {
        case DK_C: return 2;    // capacitor.hpp:61
        case DK_D: return 5;    // dio_device.hpp:50
        case DK_M1: return 17;  // mos1_device.hpp:64
        case DK_B4: return 29;  // BSIM4v7numStates (bsim4v7_def.hpp:545)
        default: return 0;
    }
}

int32_t Circuit::node(const std::string& name) {
    if (name == "0" || name == "gnd" || name == "GND") return -1;
    auto it = node_map_.find(name);
    if (it != node_map_.end()) return it->second;
    if (finalized_) throw Error(NSA_ERR_LOGIC, "Cannot add node after circuit is finalized");
    const int32_t idx = next_node_++;
    node_map_[name] = idx;
    node_names_.push_back(name);
    return idx;
}

int32_t Circuit::find_node(const std::string& name) const {
    if (name == "0" || name == "gnd" || name == "GND") return -1;
    auto it = node_map_.find(name);
    if (it == node_map_.end()) throw Error(NSA_ERR_ARGUMENT, "unknown node " + name);
    return it->second;
}

int32_t Circuit::add(Device d) {
    if (finalized_) throw Error(NSA_ERR_LOGIC, "Cannot add device after circuit is finalized");
    for (int i = 0; i < 4; ++i)
        if (d.node[i] < -1 || d.node[i] >= next_node_)
            throw Error(NSA_ERR_ARGUMENT, "device " + d.name + ": node index out of range");
    d.insertion = (int32_t)devs_.size();
    devs_.push_back(std::move(d));
    return (int32_t)devs_.size() - 1;
}

int32_t Circuit::R(const std::string& name, int32_t a, int32_t b, double ohms) {
    Device d{DK_R, name};
    d.node[0] = a; d.node[1] = b;
    // resistor.cpp:8-17: zero / non-finite resistance becomes 1 mOhm.
    d.value = (!std::isfinite(ohms) || ohms == 0.0) ? 1e-3 : ohms;
    return add(d);
}

int32_t Circuit::C(const std::string& name, int32_t a, int32_t b, double farads) {
    Device d{DK_C, name};
    d.node[0] = a; d.node[1] = b;
    d.value = farads;
    return add(d);
}

int32_t Circuit::V(const std::string& name, int32_t p, int32_t n, double dc) {
    Device d{DK_V, name};
    d.node[0] = p; d.node[1] = n;
    d.value = dc;
    return add(d);
}

int32_t Circuit::V_pulse(const std::string& name, int32_t p, int32_t n, const double pulse[7]) {
    Device d{DK_V, name};
    d.node[0] = p; d.node[1] = n;
    d.wave = WAVE_PULSE;
    for (int i = 0; i < 7; ++i) d.pulse[i] = pulse[i];
    return add(d);
}

int32_t Circuit::V_pwl(const std::string& name, int32_t p, int32_t n, const std::vector<double>& t,
                       const std::vector<double>& v) {
    if (t.size() != v.size()) throw Error(NSA_ERR_ARGUMENT, "PWL: time/value length mismatch");
    Device d{DK_V, name};
    d.node[0] = p; d.node[1] = n;
    d.wave = WAVE_PWL;
    d.pwl_t = t;
    d.pwl_v = v;
    return add(d);
}

int32_t Circuit::I(const std::string& name, int32_t p, int32_t n, double dc) {
    Device d{DK_I, name};
    d.node[0] = p; d.node[1] = n;
    d.value = dc;
    return add(d);
}

int32_t Circuit::model(const std::string& text) {
    ModelCard m = parse_model_card(text);
    if (find_model(m.name) >= 0) throw Error(NSA_ERR_ARGUMENT, "Model already exists: " + m.name);
    if (m.type == "d") {
        if (m.get("rs", 0.0) != 0.0)
            throw Error(NSA_ERR_UNSUPPORTED, "diode RS != 0 (internal node) is not supported");
    } else if (m.type == "nmos" || m.type == "pmos") {
        if (m.level != 1 && !is_bsim4_level(m.level))
            throw Error(NSA_ERR_UNSUPPORTED,
                        "only MOSFET LEVEL=1 (MOS1) and LEVEL=14 (BSIM4v7) are supported by this build");
        if (m.level == 1 &&
            (m.get("rd", 0.0) != 0.0 || m.get("rs", 0.0) != 0.0 || m.get("rsh", 0.0) != 0.0))
            throw Error(NSA_ERR_UNSUPPORTED, "MOS1 RD/RS/RSH (internal nodes) are not supported");
    } else {
        throw Error(NSA_ERR_UNSUPPORTED, "unsupported model type " + m.type);
    }
    models_.push_back(m);
    return (int32_t)models_.size() - 1;
}

int32_t Circuit::find_model(const std::string& name) const {
    const std::string k = lower(name);
    for (size_t i = 0; i < models_.size(); ++i)
        if (models_[i].name == k) return (int32_t)i;
    return -1;
}

int32_t Circuit::D(const std::string& name, int32_t a, int32_t k, const std::string& model) {
    const int32_t mi = find_model(model);
    if (mi < 0 || models_[mi].type != "d")
        throw Error(NSA_ERR_ARGUMENT, "D device: model not found or wrong type: " + model);
    if (card_needs_internal_nodes(models_[mi]))
        throw Error(NSA_ERR_UNSUPPORTED, "D device " + name + ": RS != 0 (internal node) not covered");
    Device d{DK_D, name};
    d.node[0] = a; d.node[1] = k;
    d.model = mi;
    return add(d);
}

int32_t Circuit::M(const std::string& name, int32_t dn, int32_t g, int32_t s, int32_t b,
                   const std::string& model, double w, double l) {
    const int32_t mi = find_model(model);
    if (mi < 0 || (models_[mi].type != "nmos" && models_[mi].type != "pmos"))
        throw Error(NSA_ERR_ARGUMENT, "M device: model not found or wrong type: " + model);
    if (is_bsim4_level(models_[mi].level)) {
        Device d{DK_B4, name};
        d.node[0] = dn; d.node[1] = g; d.node[2] = s; d.node[3] = b;
        d.model = mi;
        d.w = w; d.l = l;
        d.w_given = true; d.l_given = true;
        bsim4_param(models_[mi], d, DeviceEnv{});  // refuses cards outside the covered subset
        return add(d);
    }
    if (models_[mi].level != 1)
        throw Error(NSA_ERR_UNSUPPORTED,
                    "M device " + name + ": only LEVEL=1 (MOS1) and LEVEL=14 (BSIM4v7) are covered");
    if (card_needs_internal_nodes(models_[mi]))
        throw Error(NSA_ERR_UNSUPPORTED,
                    "M device " + name + ": RD/RS/RSH != 0 (internal nodes) not covered");
    Device d{DK_M1, name};
    d.node[0] = dn; d.node[1] = g; d.node[2] = s; d.node[3] = b;
    d.model = mi;
    d.w = w; d.l = l;
    d.w_given = true; d.l_given = true;
    return add(d);
}

bool Circuit::is_linear() const {
    for (const auto& d : devs_)
        if (d.kind == DK_D || d.kind == DK_M1 || d.kind == DK_B4) return false;
    return true;
}

int32_t Circuit::offset(int32_t r, int32_t c) const {
    if (r < 0 || c < 0) return -1;
    const int32_t b = col_ptr_[c], e = col_ptr_[c + 1];
    auto it = std::lower_bound(row_idx_.begin() + b, row_idx_.begin() + e, r);
    if (it == row_idx_.begin() + e || *it != r) return -1;
    return (int32_t)(it - row_idx_.begin());
}

namespace {

// Local stamp lists, in each device's own code order.
// Resistor::evaluate (resistor.cpp:33-42): pp, pn, np, nn.
// Capacitor::evaluate (capacitor.cpp:96-102): pp, pn, np, nn; rhs np, nn.
// VSource::evaluate (vsource.cpp:185-213): (p,br), (n,br), (br,p), (br,n); rhs br.
// ISource::evaluate (isource.cpp:157-174): rhs np, nn.
// DIOload (dio_load.cpp:393-406): rhs neg, posPrime; then posPos, negNeg,
//   posPrimePosPrime, posPosPrime, negPosPrime, posPrimePos, posPrimeNeg.
// MOS1load (mos1_load.cpp:846-883): rhs g, b, dPrime, sPrime; then the 22
//   matrix stamps in the order listed there.
// BSIM4v7load (bsim4v7_load.cpp:4883-5024): rhs dPrime, gPrime, bPrime,
//   sPrime; then the GP, DP, SP and BP rows (columns gp/dp/sp/bp within each
//   row's order there).  With no internal nodes the primed nodes are the
//   terminals; the D/S external-node stamps and the GIDL/GISL stamps are
//   identically zero under the covered cards and are not lowered.
void stamp_lists(const Device& d, std::vector<std::pair<int32_t, int32_t>>& mat,
                 std::vector<int32_t>& rhs) {
    mat.clear();
    rhs.clear();
    const int32_t p = d.node[0], n = d.node[1];
    switch (d.kind) {
        case DK_R:
        case DK_C:
            mat = {{p, p}, {p, n}, {n, p}, {n, n}};
            if (d.kind == DK_C) rhs = {p, n};
            break;
        case DK_V:
            mat = {{p, d.branch}, {n, d.branch}, {d.branch, p}, {d.branch, n}};
            rhs = {d.branch};
            break;
        case DK_I:
            rhs = {p, n};
            break;
        case DK_D: {
            const int32_t pp = p;  // RS = 0: posPrime is the anode
            rhs = {n, pp};
            mat = {{p, p}, {n, n}, {pp, pp}, {p, pp}, {n, pp}, {pp, p}, {pp, n}};
            break;
        }
        case DK_M1: {
            const int32_t D = d.node[0], G = d.node[1], S = d.node[2], B = d.node[3];
            const int32_t DP = D, SP = S;  // RD = RS = 0
            rhs = {G, B, DP, SP};
            mat = {{D, D},   {G, G},   {S, S},   {B, B},   {DP, DP}, {SP, SP},
                   {D, DP},  {G, B},   {G, DP},  {G, SP},  {S, SP},  {B, G},
                   {B, DP},  {B, SP},  {DP, D},  {DP, G},  {DP, B},  {DP, SP},
                   {SP, G},  {SP, S},  {SP, B},  {SP, DP}};
            break;
        }
        case DK_B4: {
            const int32_t D = d.node[0], G = d.node[1], S = d.node[2], B = d.node[3];
            rhs = {D, G, B, S};
            mat = {{G, G}, {G, D}, {G, S}, {G, B}, {D, D}, {D, G}, {D, S}, {D, B},
                   {S, D}, {S, G}, {S, S}, {S, B}, {B, D}, {B, G}, {B, S}, {B, B}};
            break;
        }
        default: break;
    }
}

// Pattern reservations (stamp_pattern): the same positions as the stamps
// for R/C/V; the TSTALLOC lists for the UCB devices, which cover the same
// set of positions.
}  // namespace

void Circuit::finalize() {
    if (finalized_) throw Error(NSA_ERR_LOGIC, "Circuit::finalize() called twice");
    // load order
    std::vector<int32_t> idx(devs_.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int32_t)i;
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) {
        const int32_t ra = load_rank(devs_[a].kind), rb = load_rank(devs_[b].kind);
        if (ra != rb) return ra < rb;
        return a > b;  // typed-API devices: newest first within a family
    });
    load_order_ = idx;
    // branch unknowns after all nodes, in load order
    int32_t next = next_node_;
    for (int32_t i : load_order_)
        if (devs_[i].kind == DK_V) devs_[i].branch = next++;
    num_vars_ = next;
    // pattern
    std::vector<std::pair<int32_t, int32_t>> ent;  // (row, col)
    std::vector<std::pair<int32_t, int32_t>> mat;
    std::vector<int32_t> rhs;
    std::vector<char> touched(num_vars_, 0);
    for (int32_t i : load_order_) {
        stamp_lists(devs_[i], mat, rhs);
        for (auto& e : mat)
            if (e.first >= 0 && e.second >= 0) {
                ent.push_back(e);
                touched[e.first] = 1;
                touched[e.second] = 1;
            }
    }
    dead_nodes_.clear();
    for (int32_t i = 0; i < next_node_; ++i)
        if (!touched[i]) { ent.push_back({i, i}); dead_nodes_.push_back(i); }
    std::sort(ent.begin(), ent.end(), [](const auto& a, const auto& b) {
        if (a.second != b.second) return a.second < b.second;
        return a.first < b.first;
    });
    ent.erase(std::unique(ent.begin(), ent.end()), ent.end());
    col_ptr_.assign(num_vars_ + 1, 0);
    row_idx_.resize(ent.size());
    for (size_t k = 0; k < ent.size(); ++k) {
        col_ptr_[ent[k].second + 1]++;
        row_idx_[k] = ent[k].first;
    }
    for (int32_t j = 0; j < num_vars_; ++j) col_ptr_[j + 1] += col_ptr_[j];
    // offsets, states, families
    num_states_ = 0;
    for (auto& f : fam_) f.clear();
    for (int32_t i : load_order_) {
        Device& d = devs_[i];
        stamp_lists(d, mat, rhs);
        d.mat_slot.resize(mat.size());
        for (size_t k = 0; k < mat.size(); ++k) d.mat_slot[k] = offset(mat[k].first, mat[k].second);
        d.rhs_row = rhs;
        const int32_t ns = state_count(d.kind);
        d.state_base = ns ? num_states_ : -1;
        num_states_ += ns;
        d.family_index = (int32_t)fam_[d.kind].size();
        fam_[d.kind].push_back(i);
    }
    finalized_ = true;
}

}  // namespace nsa
