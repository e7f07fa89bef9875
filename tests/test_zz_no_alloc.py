"""GPU: lamb_hip_trace_noalloc, the trace that allocates nothing, against
lamb_hip_trace through the C-ABI.

- Two calls in a row, the second on a smaller system (the scratch array then
  holds the first call's partials past the second's atom count), each bit for
  bit equal to lamb_hip_trace on the same matrix.
- A distribution with more atoms than the scratch array holds is refused
  with type 100 (lamb_wrong_args) before any kernel runs.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from lamb_amd import device as D
from lamb_amd import systems as S
from lamb_amd import _native as N
from lamb_amd.matrix import Distribution, LambMatrix

pytestmark = pytest.mark.gpu

WRONG_ARGS = 100
MAX_ATOMS = 1048576  # LAMB_HIP_NOALLOC_MAX_ATOMS


def _traces(ctx, atoms, h):
    dist = Distribution(atoms.blk, atoms.pos, atoms.box)
    dd = D.DeviceDistribution(ctx, dist)
    dm = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, h))
    try:
        want = dm.trace()
        out = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        ctx.check(ctx.lib.lamb_hip_trace_noalloc(dm.h, C.c_void_p(out.data_ptr())),
                  "trace_noalloc")
        torch.cuda.synchronize()
        return want, float(out.item())
    finally:
        dm.close()
        dd.close()


def test_noalloc_trace_equals_trace_twice_shrinking(ctx):
    big = S.water_system(600, seed=121)
    hb = S.decay_matrix(big, 8.0, 0.4, 122, symmetric=True, diag_shift=2.0)
    small = S.water_system(90, seed=123)
    hs = S.decay_matrix(small, 8.0, 0.4, 124, symmetric=True, diag_shift=2.0)
    for name, atoms, h in (("first", big, hb), ("second, smaller", small, hs)):
        print(f"case: trace_noalloc {name}: {atoms.n_atoms} atoms", flush=True)
        want, got = _traces(ctx, atoms, h)
        print(f"  trace {want!r} noalloc {got!r}", flush=True)
        assert np.float64(got).view(np.int64) == np.float64(want).view(np.int64)
        dense = float(np.trace(h.to_dense())) if atoms.n_atoms <= 100 else want
        assert got == pytest.approx(dense, rel=1e-13)


def test_noalloc_trace_refuses_too_many_atoms(ctx):
    n = MAX_ATOMS + 1
    n_subsets = 64  # local indices stay within int16
    print(f"case: trace_noalloc with {n} atoms", flush=True)
    blk = np.ones(n, np.int8)
    subset = (np.arange(n) % n_subsets + 1).astype(np.int32)
    local = (np.arange(n) // n_subsets + 1).astype(np.int16)
    lib = ctx.lib
    dist = C.c_void_p()
    ctx.check(lib.lamb_hip_dist_create(ctx.h, n, N.ptr(blk, C.c_int8), n_subsets,
                                       N.ptr(subset, C.c_int32), N.ptr(local, C.c_int16),
                                       None, None, C.byref(dist)), "dist_create")
    mat = C.c_void_p()
    try:
        ctx.check(lib.lamb_hip_matrix_create(dist, C.byref(mat)), "matrix_create")
        ctx.check(lib.lamb_hip_matrix_commit(mat), "commit")
        out = torch.zeros(1, dtype=torch.float64, device="cuda")
        rc = lib.lamb_hip_trace_noalloc(mat, C.c_void_p(out.data_ptr()))
        assert rc != 0
        lvl, typ = C.c_int32(), C.c_int32()
        lib.lamb_hip_last_error(ctx.h, C.byref(lvl), C.byref(typ), None, 0)
        assert typ.value == WRONG_ARGS
    finally:
        if mat:
            lib.lamb_hip_matrix_destroy(mat)
        lib.lamb_hip_dist_destroy(dist)
