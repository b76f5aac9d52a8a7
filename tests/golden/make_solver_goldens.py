"""Writes solver_known_answers.json: the known-answer cases of the
reference's own solver tests, as data (inputs and expected outputs).

Sources (reference repository):
  tests/unit/test_neo_solver.cpp:15-378   NeoSolver solves / refactorize
  tests/unit/test_amd_lu.cpp:51-450       AmdLuSolver parity cases
  tests/unit/test_amd.cpp:8-70            minimum-degree ordering properties

Each case: n, entries [[row, col, value], ...] (duplicates add, as
NumericMatrix::add does), optional "refactor" entries (values for a
refactorize on the same pattern), the expected solution "x" for the final
values, the right-hand side "b" (given, or A x computed here in the
reference's own accumulation order), a tolerance, and which solvers the
reference test exercises.  Run:  python tests/golden/make_solver_goldens.py
"""
import json
import os


def dense(n, f):
    return [[i, j, f(i, j)] for j in range(n) for i in range(n)]


def tridiag(n, d, o):
    e = []
    for i in range(n):
        e.append([i, i, d])
        if i > 0:
            e.append([i, i - 1, o])
        if i < n - 1:
            e.append([i, i + 1, o])
    return e


def rhs_of(n, entries, x):
    # rhs[r] += value(r, c) * x[c] over the pattern's entries in (col, row)
    # order, as the tests' make_rhs loops over pat.entries()
    acc = {}
    for r, c, v in entries:
        acc[(r, c)] = acc.get((r, c), 0.0) + v
    b = [0.0] * n
    for (r, c) in sorted(acc, key=lambda rc: (rc[1], rc[0])):
        b[r] += acc[(r, c)] * x[c]
    return b


cases = []


def add(name, n, entries, x, tol, solvers, b=None, refactor=None, gmin=0.0, src=""):
    final = refactor if refactor is not None else entries
    cases.append({"name": name, "n": n, "entries": entries, "refactor": refactor,
                  "x": x, "b": b if b is not None else rhs_of(n, final, x), "tol": tol,
                  "gmin": gmin, "solvers": solvers, "source": src})


# test_neo_solver.cpp
add("neo_1x1", 1, [[0, 0, 3.0]], [3.0], 1e-12, ["neo"], b=[9.0], src="test_neo_solver.cpp:15-27")
add("neo_2x2", 2, [[0, 0, 2.0], [1, 0, 1.0], [0, 1, 1.0], [1, 1, 3.0]], [1.6, 1.8], 1e-10,
    ["neo"], b=[5.0, 7.0], src="test_neo_solver.cpp:29-46")
add("neo_5x5", 5, dense(5, lambda i, j: 10.0 if i == j else 1.0), [1, 2, 3, 4, 5], 1e-10, ["neo"],
    src="test_neo_solver.cpp:48-72")
add("neo_refactorize_2x2", 2, [[0, 0, 2.0], [1, 0, 1.0], [0, 1, 1.0], [1, 1, 3.0]], [2.0, 3.0], 1e-10,
    ["neo"], b=[14.0, 17.0], refactor=[[0, 0, 4.0], [1, 0, 1.0], [0, 1, 2.0], [1, 1, 5.0]],
    src="test_neo_solver.cpp:74-99")
add("neo_tridiag_24", 24, tridiag(24, 4.0, -1.0), [float(i + 1) for i in range(24)], 1e-10, ["neo"],
    src="test_neo_solver.cpp:229-262")
add("neo_tridiag_25", 25, tridiag(25, 4.0, -1.0), [float(i + 1) for i in range(25)], 1e-10, ["neo"],
    src="test_neo_solver.cpp:264-297")
band = []
for i in range(100):
    band.append([i, i, 20.0])
for j in range(100):
    for k in range(1, 4):
        if j + k < 100:
            band.append([j + k, j, 1.0])
            band.append([j, j + k, 1.0])
add("neo_band_100", 100, band, [float(i % 7 + 1) for i in range(100)], 1e-9, ["neo"],
    src="test_neo_solver.cpp:299-337")
add("neo_refactorize_tridiag_50", 50, tridiag(50, 4.0, -1.0), [1.0] * 50, 1e-10, ["neo"],
    refactor=tridiag(50, 8.0, -2.0), src="test_neo_solver.cpp:339-378")

# test_amd_lu.cpp (check_parity: AmdLu == x_true and == NeoSolver within tol)
add("amd_1x1", 1, [[0, 0, 3.0]], [7.0], 1e-12, ["neo", "amd"], src="test_amd_lu.cpp:51-58")
add("amd_2x2", 2, [[0, 0, 2.0], [1, 0, 1.0], [0, 1, 1.0], [1, 1, 3.0]], [1.6, 1.8], 1e-10,
    ["neo", "amd"], src="test_amd_lu.cpp:60-70")
add("amd_dense_8", 8, dense(8, lambda i, j: 20.0 if i == j else 1.0), [float(i + 1) for i in range(8)],
    1e-9, ["neo", "amd"], src="test_amd_lu.cpp:72-84")
add("amd_tridiag_60", 60, tridiag(60, 4.0, -1.0), [float(i % 5 + 1) for i in range(60)], 1e-9,
    ["neo", "amd"], src="test_amd_lu.cpp:86-104")
arrow = [[i, i, 100.0 if i == 0 else 5.0] for i in range(16)]
for i in range(1, 16):
    arrow += [[0, i, 1.0], [i, 0, 1.0]]
add("amd_arrowhead_16", 16, arrow, [(i % 3) + 0.5 for i in range(16)], 1e-9, ["neo", "amd"],
    src="test_amd_lu.cpp:106-122")
add("amd_indefinite_2x2", 2, [[0, 1, 1.0], [1, 0, 1.0], [1, 1, 1.0]], [1.0, 1.0], 1e-12, ["neo", "amd"],
    src="test_amd_lu.cpp:189-200")
add("amd_zero_diag_block_4", 4, [[0, 2, 1.0], [1, 2, 1.0], [1, 3, 1.0], [2, 0, 1.0], [2, 1, 1.0],
                                 [2, 2, 5.0], [3, 1, 1.0], [3, 3, 3.0]],
    [2.0, -1.0, 0.5, 1.5], 1e-9, ["neo", "amd"], src="test_amd_lu.cpp:202-219")
add("amd_gmin_2x2", 2, [[0, 0, 0.0], [1, 1, 0.0]], [1.0 / 1e-3, 2.0 / 1e-3], 1e-9, ["neo", "amd"],
    b=[1.0, 2.0], gmin=1e-3, src="test_amd_lu.cpp:221-248")
add("amd_refactorize_tridiag_40", 40, tridiag(40, 4.0, -1.0), [1.0] * 40,
    1e-10, ["amd"], refactor=tridiag(40, 8.0, -2.0), src="test_amd_lu.cpp:250-284")
a0 = [[10, 1, 2], [1, 12, 3], [2, 1, 15]]
a1 = [[1e-12, 1, 2], [1, 12, 3], [2, 1, 15]]
add("amd_refactor_fallback_3", 3, [[i, j, float(a0[i][j])] for j in range(3) for i in range(3)],
    [2.0, -1.0, 0.5], 1e-7, ["neo", "amd"],
    refactor=[[i, j, float(a1[i][j])] for j in range(3) for i in range(3)],
    src="test_amd_lu.cpp:356-400 (replay must be rejected by the growth check)")

out = {"cases": cases,
       "amd_order": [
           {"name": "identity_3", "n": 3, "pattern": [[0, 0], [1, 1], [2, 2]], "check": "permutation",
            "source": "test_amd.cpp:8-18"},
           {"name": "tridiag_10", "n": 10,
            "pattern": [[i, i] for i in range(10)] + [[i, i - 1] for i in range(1, 10)] + [[i - 1, i] for i in range(1, 10)],
            "check": "permutation", "source": "test_amd.cpp:20-34"},
           {"name": "arrowhead_20", "n": 20,
            "pattern": [[i, i] for i in range(20)] + [[0, i] for i in range(1, 20)] + [[i, 0] for i in range(1, 20)],
            "check": "hub_last", "source": "test_amd.cpp:36-50"},
           {"name": "band3_50", "n": 50,
            "pattern": [[i, i + d] for i in range(50) for d in range(-3, 4) if 0 <= i + d < 50],
            "check": "permutation", "source": "test_amd.cpp:52-68"},
       ]}
here = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(here, "solver_known_answers.json"), "w") as f:
    json.dump(out, f, indent=1)
print("wrote", len(cases), "solver cases")
