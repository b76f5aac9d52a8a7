"""nsa_lu_solve_noalloc (no device allocation; the staging vector in the
library's static device array) against nsa_lu_solve on the same factored
handles: bit-equal solutions on two calls in a row, the second on a smaller
system, and a clear refusal of sizes beyond the array."""
import numpy as np
import pytest

import neospice_amd as na
from neospice_amd import workloads as W
from neospice_amd._lib import NsaError

pytestmark = pytest.mark.gpu


def mesh_system(m, B, seed):
    """Diagonally dominant matrices on the m x m RC mesh's pattern, one per
    instance, and right-hand sides."""
    ckt, _ = W.rc_mesh(m)
    ckt.finalize()
    n, cp, ri = ckt.n, np.asarray(ckt.col_ptr), np.asarray(ckt.row_idx)
    rng = np.random.default_rng(seed)
    vals = -rng.uniform(0.1, 1.0, (cp[n], B))
    diag = np.zeros(cp[n], bool)
    for j in range(n):
        for p in range(cp[j], cp[j + 1]):
            if ri[p] == j:
                diag[p] = True
    for j in range(n):
        for p in range(cp[j], cp[j + 1]):
            if diag[p]:
                vals[p] = np.abs(vals[cp[j]:cp[j + 1]]).sum(axis=0) + 1.0
    return n, cp, ri, vals, rng.standard_normal((n, B))


def solved_pair(m, B, seed):
    n, cp, ri, vals, rhs = mesh_system(m, B, seed)
    s = na.Solver(n, cp, ri, batch=B, is_linear=True)
    assert not s.numeric(vals).any()
    x_ref = s.solve(rhs.copy())
    x = s.solve_noalloc(rhs.copy())
    return s, x_ref, x


def test_solve_noalloc_matches_solve():
    # first call: 24 x 24 mesh, 64 instances (36928 values staged)
    s1, x_ref1, x1 = solved_pair(24, 64, 1)
    print("call 1: n", s1.n, "batch", s1.B, "name", s1.name, flush=True)
    assert np.array_equal(x1, x_ref1)
    # second call: a smaller system, so the array still holds the first call's tail
    s2, x_ref2, x2 = solved_pair(6, 3, 2)
    print("call 2: n", s2.n, "batch", s2.B, flush=True)
    assert np.array_equal(x2, x_ref2)
    # and the first handle again
    _, _, _, _, rhs = mesh_system(24, 64, 3)
    assert np.array_equal(s1.solve_noalloc(rhs.copy()), s1.solve(rhs.copy()))


def test_solve_noalloc_refuses_oversize():
    ckt, _ = W.rc_mesh(4)
    ckt.finalize()
    n = ckt.n
    B = (16 << 20) // 8 // n + 1
    s = na.Solver(n, ckt.col_ptr, ckt.row_idx, batch=B, is_linear=True)
    with pytest.raises(NsaError) as e:
        s.solve_noalloc(np.zeros((n, B)))
    print("refused:", e.value, flush=True)
    assert e.value.code == -2 and "scratch" in str(e.value)
