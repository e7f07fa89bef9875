"""torch.autograd binding of the sector multiply (a thin test binding).

A matrix is a pair (values, Pattern): ``values`` a float64 CUDA tensor in the
library's canonical order (stored blocks in CSR order of the internal atom
order, each column-major, back to back), ``Pattern`` a device matrix whose
value pool is released between uses so that only the index stays resident.
``matmul`` is C = A*B as a torch.autograd.Function whose backward calls the
hand-written adjoint kernels (lamb_hip_multiply_adjoint: dA = G B^T on A's
pattern, dB = A^T G on B's pattern); ``project`` moves values between
patterns (its adjoint is the projection back). ``mcweeny`` builds the
differentiable purification chain P <- 3 P^2 - 2 P^3 from H, with torch
checkpointing of every ``checkpoint_every`` steps so that only segment
boundaries are kept for the backward pass.

The library runs on its own HIP stream: every call waits for torch's current
stream first and returns after its own work is done.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch
import torch.utils.checkpoint

from . import _native as N
from . import device as D


class Pattern:
    """A stored-block pattern on the device (values released when idle)."""

    def __init__(self, dm: D.DeviceMatrix):
        self.dm = dm
        self.n_blocks, self.n_values = dm.info()
        self.release()

    def release(self):
        self.dm.ctx.check(self.dm.ctx.lib.lamb_hip_matrix_release_values(self.dm.h),
                          "release_values")

    def load(self, values: torch.Tensor):
        v = values.detach().contiguous()
        if v.dtype != torch.float64 or not v.is_cuda or v.numel() != self.n_values:
            raise ValueError("values must be a float64 CUDA tensor of the pattern's size")
        torch.cuda.current_stream().synchronize()
        rc = self.dm.ctx.lib.lamb_hip_matrix_set_values(self.dm.h, C.c_void_p(v.data_ptr()),
                                                        self.n_values)
        self.dm.ctx.check(rc, "set_values")

    def block_atoms(self):
        """Global (row, col) atom of each stored block, canonical order."""
        r = np.zeros(self.n_blocks, np.int32)
        c = np.zeros(self.n_blocks, np.int32)
        rc = self.dm.ctx.lib.lamb_hip_matrix_block_atoms(self.dm.h, N.ptr(r, C.c_int32),
                                                         N.ptr(c, C.c_int32))
        self.dm.ctx.check(rc, "block_atoms")
        return r, c


def values_of(dm: D.DeviceMatrix) -> torch.Tensor:
    """The values of a device matrix (with its pool present) as a tensor."""
    _, n = dm.info()
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    torch.cuda.current_stream().synchronize()
    rc = dm.ctx.lib.lamb_hip_matrix_get_values(dm.h, C.c_void_p(out.data_ptr()), n)
    dm.ctx.check(rc, "get_values")
    return out


def leaf(dm: D.DeviceMatrix):
    """(values, Pattern) of an existing device matrix; dm becomes the pattern."""
    v = values_of(dm)
    return v, Pattern(dm)


class _MatMul(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a, b, pa, pb, flags, out):
        pa.load(a)
        pb.load(b)
        c_dm = D.DeviceMatrix(pa.dm.ddist)
        D.multiply(pa.dm, pb.dm, c_dm, flags=flags)
        c = values_of(c_dm)
        pc = Pattern(c_dm)
        pa.release()
        pb.release()
        out.append(pc)
        ctx.save_for_backward(a, b)
        ctx.pats = (pa, pb, pc)
        ctx.flags = flags
        return c

    @staticmethod
    def backward(ctx, g):
        a, b = ctx.saved_tensors
        pa, pb, pc = ctx.pats
        pa.load(a)
        pb.load(b)
        pc.load(g)
        dist = pa.dm.ddist
        da_dm, db_dm = D.DeviceMatrix(dist), D.DeviceMatrix(dist)
        st = N.Stats()
        lib = pa.dm.ctx.lib
        need_a, need_b = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        rc = lib.lamb_hip_multiply_adjoint(pa.dm.ctx.h, pa.dm.h, pb.dm.h, pc.dm.h,
                                           da_dm.h if need_a else None,
                                           db_dm.h if need_b else None, int(ctx.flags),
                                           C.byref(st))
        pa.dm.ctx.check(rc, "multiply_adjoint")
        da = values_of(da_dm) if need_a else None
        db = values_of(db_dm) if need_b else None
        da_dm.close()
        db_dm.close()
        pa.release()
        pb.release()
        pc.release()
        return da, db, None, None, None, None


def matmul(a, pa: Pattern, b, pb: Pattern, flags: int = D.FP64):
    """C = A*B (unfiltered): returns (values, Pattern) of C."""
    out = []
    c = _MatMul.apply(a, b, pa, pb, flags, out)
    return c, out[0]


class _Project(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, px, target):
        px.load(x)
        y_dm = D.DeviceMatrix(px.dm.ddist)
        lib = px.dm.ctx.lib
        px.dm.ctx.check(lib.lamb_hip_matrix_project(px.dm.h, target.dm.h, y_dm.h), "project")
        y = values_of(y_dm)
        y_dm.close()
        px.release()
        ctx.pats = (px, target)
        return y

    @staticmethod
    def backward(ctx, g):
        px, target = ctx.pats
        target.load(g)
        x_dm = D.DeviceMatrix(px.dm.ddist)
        lib = px.dm.ctx.lib
        px.dm.ctx.check(lib.lamb_hip_matrix_project(target.dm.h, px.dm.h, x_dm.h), "project")
        gx = values_of(x_dm)
        x_dm.close()
        target.release()
        return gx, None, None


def project(x, px: Pattern, target: Pattern):
    """x (on px's pattern) moved onto target's pattern (zeros where absent)."""
    return _Project.apply(x, px, target)


def diagonal_mask(p: Pattern, blk: np.ndarray) -> torch.Tensor:
    """1.0 at the diagonal elements of diagonal blocks, canonical order."""
    rows, cols = p.block_atoms()
    sizes = blk[rows].astype(np.int64) * blk[cols]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    m = np.zeros(p.n_values)
    for k in np.flatnonzero(rows == cols):
        n = int(blk[rows[k]])
        m[offs[k] + np.arange(n) * (n + 1)] = 1.0
    return torch.from_numpy(m).to("cuda")


def mcweeny_step(p, pp: Pattern, flags: int = D.FP64):
    """One step P <- 3 P^2 - 2 P^3 = 3 P2 - 2 P2 P on values; returns (values, Pattern)."""
    p2, pp2 = matmul(p, pp, p, pp, flags)
    c1, pc1 = matmul(p2, pp2, p, pp, flags)
    t = 3.0 * project(p2, pp2, pc1) - 2.0 * c1
    return t, pc1


def mcweeny(p0, pp0: Pattern, steps: int, flags: int = D.FP64, checkpoint_every: int = 2):
    """The differentiable chain from P0; checkpointed in segments of
    ``checkpoint_every`` steps (0: no checkpointing). Returns (values, Pattern)."""
    p, pp = p0, pp0
    k = 0
    while k < steps:
        n = min(checkpoint_every or steps, steps - k)
        if checkpoint_every and torch.is_grad_enabled():
            holder = {}

            def segment(x, pp=pp, n=n, holder=holder):
                q, pq = x, pp
                for _ in range(n):
                    q, pq = mcweeny_step(q, pq, flags)
                holder["pattern"] = pq
                return q

            p = torch.utils.checkpoint.checkpoint(segment, p, use_reentrant=False)
            pp = holder["pattern"]
        else:
            for _ in range(n):
                p, pp = mcweeny_step(p, pp, flags)
        k += n
    return p, pp


def purification_loss(h, ph: Pattern, blk: np.ndarray, lo: float, hi: float, steps: int,
                      flags: int = D.FP64, checkpoint_every: int = 2):
    """L = sum over H's stored blocks of <P_steps, H> (= tr(P_steps H) for a
    symmetric H), P0 = (H - lo I) / (hi - lo) with lo, hi held constant."""
    scale = 1.0 / (hi - lo)
    p0 = scale * h - lo * scale * diagonal_mask(ph, blk)
    p, pp = mcweeny(p0, ph, steps, flags, checkpoint_every)
    return torch.dot(project(p, pp, ph), h)
