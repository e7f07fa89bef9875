"""CPU FP64 oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py
import this package, and only as the checker; the product path never does.
It wraps oracle/lamb_oracle.c (built by oracle/Makefile into oracle/_build/),
whose header states the semantics it restates and the reference lines each
part follows.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "_build" / "liblamb_oracle.so"


def build() -> pathlib.Path:
    subprocess.run(["make", "-s", "-C", os.fspath(HERE)], check=True)
    return LIB


class _Mat(C.Structure):
    _fields_ = [("n_atoms", C.c_int32), ("blk", C.POINTER(C.c_int32)),
                ("row_p", C.POINTER(C.c_int64)), ("col", C.POINTER(C.c_int32)),
                ("off", C.POINTER(C.c_int64)), ("data", C.POINTER(C.c_double))]


class _Res(C.Structure):
    _fields_ = [("nblk", C.c_int64), ("ndata", C.c_int64),
                ("row_p", C.POINTER(C.c_int64)), ("col", C.POINTER(C.c_int32)),
                ("off", C.POINTER(C.c_int64)), ("data", C.POINTER(C.c_double)),
                ("nsq", C.POINTER(C.c_double)),
                ("n_triples_kept", C.c_int64), ("n_triples_filtered", C.c_int64),
                ("n_struct_blocks", C.c_int64), ("flops", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(os.fspath(LIB))
        L.orc_multiply.argtypes = [C.POINTER(_Mat), C.POINTER(_Mat), C.POINTER(_Mat),
                                   C.c_double, C.c_double, C.c_double, C.POINTER(_Res)]
        L.orc_multiply.restype = C.c_int
        L.orc_result_free.argtypes = [C.POINTER(_Res)]
        L.orc_block_nsq.argtypes = [C.POINTER(_Mat), C.POINTER(C.c_double)]
        L.orc_trace.argtypes = [C.POINTER(_Mat)]
        L.orc_trace.restype = C.c_double
        L.orc_gershgorin.argtypes = [C.POINTER(_Mat), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_adjoint.argtypes = [C.POINTER(_Mat), C.POINTER(_Mat), C.POINTER(_Mat),
                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_nsq1.argtypes = [C.POINTER(C.c_double), C.c_int64]
        L.orc_nsq1.restype = C.c_double
        L.orc_num_threads.restype = C.c_int
        L.orc_set_num_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _mat(csr):
    """Build the C view of a BlockCSR; returns (struct, keep-alive arrays)."""
    blk = np.ascontiguousarray(csr.blk, dtype=np.int32)
    row_p = np.ascontiguousarray(csr.row_p, dtype=np.int64)
    col = np.ascontiguousarray(csr.col, dtype=np.int32)
    off = np.ascontiguousarray(csr.blk_off, dtype=np.int64)
    data = np.ascontiguousarray(csr.data, dtype=np.float64)
    if len(col) == 0:
        col = np.zeros(1, np.int32)
        off = np.zeros(1, np.int64)
    if len(data) == 0:
        data = np.zeros(1)
    m = _Mat(len(blk), _p(blk, C.c_int32), _p(row_p, C.c_int64), _p(col, C.c_int32),
             _p(off, C.c_int64), _p(data, C.c_double))
    return m, (blk, row_p, col, off, data)


def set_threads(n: int):
    lib().orc_set_num_threads(int(n))


def num_threads() -> int:
    return int(lib().orc_num_threads())


def block_nsq(csr) -> np.ndarray:
    m, keep = _mat(csr)
    out = np.zeros(max(csr.n_blocks, 1))
    lib().orc_block_nsq(C.byref(m), _p(out, C.c_double))
    return out[:csr.n_blocks]


def nsq1(x: np.ndarray) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    return float(lib().orc_nsq1(_p(x, C.c_double), len(x)))


def multiply(a, b, *, eps: float = 0.0, alpha: float = 1.0, beta: float = 0.0, c_old=None):
    """C = alpha*A*B + beta*C_old with the prefilter and post-filter.

    Returns (BlockCSR, stats dict)."""
    from lamb_amd.systems import BlockCSR
    ma, ka = _mat(a)
    mb, kb = _mat(b)
    if c_old is not None:
        mc, kc = _mat(c_old)
        pc = C.byref(mc)
    else:
        pc = None
    r = _Res()
    rc = lib().orc_multiply(C.byref(ma), C.byref(mb), pc, float(alpha), float(beta),
                            float(eps), C.byref(r))
    if rc != 0:
        raise MemoryError(f"oracle multiply failed ({rc})")
    n = a.n_atoms
    try:
        row_p = np.ctypeslib.as_array(r.row_p, shape=(n + 1,)).copy()
        nb = int(r.nblk)
        col = np.ctypeslib.as_array(r.col, shape=(max(nb, 1),))[:nb].copy()
        off = np.ctypeslib.as_array(r.off, shape=(max(nb, 1),))[:nb].copy()
        nsq = np.ctypeslib.as_array(r.nsq, shape=(max(nb, 1),))[:nb].copy()
        nd = int(r.ndata)
        data = np.ctypeslib.as_array(r.data, shape=(max(nd, 1),))[:nd].copy()
        stats = dict(n_triples_kept=int(r.n_triples_kept),
                     n_triples_filtered=int(r.n_triples_filtered),
                     n_struct_blocks=int(r.n_struct_blocks), flops=float(r.flops),
                     nsq=nsq)
    finally:
        lib().orc_result_free(C.byref(r))
    return BlockCSR(np.asarray(a.blk, np.int32).copy(), row_p, col, off, data), stats


def trace(csr) -> float:
    m, keep = _mat(csr)
    return float(lib().orc_trace(C.byref(m)))


def gershgorin(csr):
    m, keep = _mat(csr)
    lo, hi = C.c_double(), C.c_double()
    lib().orc_gershgorin(C.byref(m), C.byref(lo), C.byref(hi))
    return lo.value, hi.value


def adjoint(a, b, g):
    """(dA, dB) data arrays laid out like a.data / b.data."""
    ma, ka = _mat(a)
    mb, kb = _mat(b)
    mg, kg = _mat(g)
    da = np.zeros(max(len(a.data), 1))
    db = np.zeros(max(len(b.data), 1))
    lib().orc_adjoint(C.byref(ma), C.byref(mb), C.byref(mg), _p(da, C.c_double),
                      _p(db, C.c_double))
    return da[:len(a.data)], db[:len(b.data)]


def scale_shift(csr, a: float, b: float):
    """a * X + b * I (every diagonal block must be stored when b != 0)."""
    from lamb_amd.systems import BlockCSR
    data = a * csr.data
    if b != 0.0:
        rows = csr.rows()
        for k in np.flatnonzero(rows == csr.col):
            m = int(csr.blk[rows[k]])
            idx = csr.blk_off[k] + np.arange(m) * (m + 1)
            data[idx] += b
        missing = set(np.flatnonzero(csr.blk > 0)) - set(rows[rows == csr.col])
        if missing:
            raise ValueError(f"{len(missing)} atoms without a stored diagonal block")
    return BlockCSR(csr.blk.copy(), csr.row_p.copy(), csr.col.copy(), csr.blk_off.copy(), data)


def purification_start(h):
    """P0 = (H - lo I) / (hi - lo), lo / hi the Gershgorin bounds of H."""
    lo, hi = gershgorin(h)
    return scale_shift(h, 1.0 / (hi - lo), -lo / (hi - lo))


def diff_norm(x, y) -> float:
    """||X - Y||_F over the union pattern (blocks compared by (row, col))."""
    def blocks(m):
        rows = m.rows()
        out = {}
        for k in range(m.n_blocks):
            z = int(m.blk[rows[k]]) * int(m.blk[m.col[k]])
            out[(int(rows[k]), int(m.col[k]))] = m.data[m.blk_off[k]:m.blk_off[k] + z]
        return out
    bx, by = blocks(x), blocks(y)
    s = 0.0
    for key in set(bx) | set(by):
        d = bx.get(key, 0.0) - by.get(key, 0.0)
        s += float(np.sum(np.square(d)))
    return float(np.sqrt(s))


def mcweeny(p, steps: int, *, eps: float = 0.0):
    """P <- 3 P^2 - 2 P^3, steps times, with the same two filtered multiplies
    per step as the device chain (P2 = P*P; T = -2 P2*P + 3 P2). Returns
    (P, traces [steps], idempotency errors ||P^2 - P||_F of each step's
    input [steps])."""
    traces, idem = [], []
    for _ in range(steps):
        p2, _ = multiply(p, p, eps=eps)
        idem.append(diff_norm(p2, p))
        p, _ = multiply(p2, p, eps=eps, alpha=-2.0, beta=3.0, c_old=p2)
        traces.append(trace(p))
    return p, np.array(traces), np.array(idem)
