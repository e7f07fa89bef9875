"""Device handles over the C-ABI: context, distribution, matrix."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _native as N
from .matrix import Distribution, LambMatrix, Sector

# error levels and types of the reference (src/lamb_error.F90:116-159)
LEVELS = {0: "ok", 1: "note", 2: "warning", 3: "failure", 4: "memory", 5: "fatal"}
FP64 = 0
BF16_SPLIT = 1


def bf16_split(slices: int = 5) -> int:
    """multiply flags for the BF16-split form with the given slice count (3..5)."""
    return BF16_SPLIT | (int(slices) << 8)


class LambError(RuntimeError):
    def __init__(self, level: int, type_: int, msg: str):
        super().__init__(f"[{LEVELS.get(level, level)} / type {type_}] {msg}")
        self.level = level
        self.type = type_
        self.msg = msg


class Context:
    """One HIP device + stream (lamb_hip_init)."""

    def __init__(self, device: int = 0):
        self.lib = N.load()
        h = C.c_void_p()
        rc = self.lib.lamb_hip_init(device, C.byref(h))
        if rc != 0:
            raise LambError(rc, -2000, f"lamb_hip_init(device={device}) failed")
        self.h = h
        self.device = device

    def check(self, rc: int, what: str = ""):
        if rc == 0:
            return
        lvl, typ = C.c_int32(), C.c_int32()
        buf = C.create_string_buffer(512)
        self.lib.lamb_hip_last_error(self.h, C.byref(lvl), C.byref(typ), buf, 512)
        raise LambError(rc, typ.value, f"{what}: {buf.value.decode(errors='replace')}")

    def close(self):
        if self.h:
            self.lib.lamb_hip_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceDistribution:
    def __init__(self, ctx: Context, dist: Distribution, spatial: bool = True):
        self.ctx = ctx
        self.dist = dist
        blk = np.ascontiguousarray(dist.blk.astype(np.int8))
        sm = np.ascontiguousarray(dist.subset_map.astype(np.int32))
        om = np.ascontiguousarray(dist.order_map.astype(np.int16))
        xyz = np.ascontiguousarray(dist.coords.astype(np.float32).reshape(-1)) if spatial else None
        box = np.full(3, dist.box, np.float32)
        h = C.c_void_p()
        rc = ctx.lib.lamb_hip_dist_create(ctx.h, dist.n_atoms, N.ptr(blk, C.c_int8),
                                          dist.n_subsets, N.ptr(sm, C.c_int32),
                                          N.ptr(om, C.c_int16), N.ptr(xyz, C.c_float),
                                          N.ptr(box, C.c_float), C.byref(h))
        ctx.check(rc, "dist_create")
        self.h = h
        self.n_groups = 1

    def set_groups(self, group_of_atom: np.ndarray, n_groups: int | None = None):
        """Group (independent system) id in 0..n_groups-1 of each global atom."""
        g = np.ascontiguousarray(group_of_atom, dtype=np.int32)
        n = int(g.max()) + 1 if n_groups is None else int(n_groups)
        self.ctx.check(self.ctx.lib.lamb_hip_dist_set_groups(self.h, n, N.ptr(g, C.c_int32)),
                       "dist_set_groups")
        self.n_groups = n

    def close(self):
        if self.h:
            self.ctx.lib.lamb_hip_dist_destroy(self.h)
            self.h = None


class DeviceMatrix:
    """A device-resident block matrix (lamb_hip_mat)."""

    def __init__(self, ddist: DeviceDistribution):
        self.ddist = ddist
        self.ctx = ddist.ctx
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.lamb_hip_matrix_create(ddist.h, C.byref(h)), "matrix_create")
        self.h = h

    @classmethod
    def from_lamb(cls, ddist: DeviceDistribution, m: LambMatrix) -> "DeviceMatrix":
        dm = cls(ddist)
        dm.upload(m)
        return dm

    @classmethod
    def from_dbcsr(cls, ddist: DeviceDistribution, path) -> "DeviceMatrix":
        """A DBCSR dump loaded into the device pool (lamb_hip_dbcsr_load)."""
        dm = cls(ddist)
        try:
            dm.ctx.check(dm.ctx.lib.lamb_hip_dbcsr_load(dm.h, os.fsencode(os.fspath(path))),
                         "dbcsr_load")
        except Exception:
            dm.close()
            raise
        return dm

    def upload(self, m: LambMatrix):
        lib = self.ctx.lib
        for sec in m.sectors:
            r, s = sec.subsets
            row_p = np.ascontiguousarray(sec.row_p, dtype=np.int64)
            col_i = np.ascontiguousarray(sec.col_i, dtype=np.int16)
            ext = np.ascontiguousarray(sec.extents, dtype=np.int64).reshape(-1)
            data = np.ascontiguousarray(sec.data, dtype=np.float64)
            rc = lib.lamb_hip_sector_upload(self.h, r, s, sec.n_blks, N.ptr(row_p, C.c_int64),
                                            N.ptr(col_i, C.c_int16), N.ptr(ext, C.c_int64),
                                            N.ptr(data, C.c_double), len(data))
            self.ctx.check(rc, f"sector_upload({r},{s})")
        self.ctx.check(lib.lamb_hip_matrix_commit(self.h), "matrix_commit")

    def download(self) -> LambMatrix:
        lib = self.ctx.lib
        dist = self.ddist.dist
        sectors = []
        for (r, s) in dist.sectors:
            nb, nd = C.c_int64(), C.c_int64()
            self.ctx.check(lib.lamb_hip_sector_query(self.h, r, s, C.byref(nb), C.byref(nd)),
                           "sector_query")
            n_rows = len(dist.subset_atoms[r - 1])
            row_p = np.zeros(n_rows + 1, np.int64)
            col_i = np.zeros(nb.value, np.int16)
            ext = np.zeros((nb.value, 2), np.int64)
            data = np.zeros(nd.value, np.float64)
            rc = lib.lamb_hip_sector_download(self.h, r, s, N.ptr(row_p, C.c_int64),
                                              N.ptr(col_i, C.c_int16), N.ptr(ext, C.c_int64),
                                              N.ptr(data, C.c_double))
            self.ctx.check(rc, f"sector_download({r},{s})")
            sectors.append(Sector((r, s), row_p, col_i, ext, data))
        return LambMatrix(dist, sectors)

    def goods_pack(self, subset_row: int, subset_col: int):
        """One sector in the wire format of lamb_goods_X_create
        (src/lamb_goods_X_types.F90:119-268): (index int64 words, data f64)."""
        lib = self.ctx.lib
        ni, nd = C.c_int64(), C.c_int64()
        self.ctx.check(lib.lamb_hip_goods_size(self.h, subset_row, subset_col, C.byref(ni),
                                               C.byref(nd)), "goods_size")
        index = np.zeros(ni.value, np.int64)
        data = np.zeros(nd.value, np.float64)
        self.ctx.check(lib.lamb_hip_goods_pack(self.h, subset_row, subset_col,
                                               N.ptr(index, C.c_int64), len(index),
                                               N.ptr(data, C.c_double), len(data)),
                       f"goods_pack({subset_row},{subset_col})")
        return index, data

    def goods_unpack(self, subset_row: int, subset_col: int, index, data, commit=True):
        """Stage one sector from the wire format (lamb_goods_X_unpack,
        src/lamb_goods_X_methods.F90:166-218); commit makes it live."""
        lib = self.ctx.lib
        index = np.ascontiguousarray(index, dtype=np.int64)
        data = np.ascontiguousarray(data, dtype=np.float64)
        self.ctx.check(lib.lamb_hip_goods_unpack(self.h, subset_row, subset_col,
                                                 N.ptr(index, C.c_int64), len(index),
                                                 N.ptr(data, C.c_double), len(data)),
                       f"goods_unpack({subset_row},{subset_col})")
        if commit:
            self.ctx.check(lib.lamb_hip_matrix_commit(self.h), "matrix_commit")

    def copy_from(self, other: "DeviceMatrix"):
        self.ctx.check(self.ctx.lib.lamb_hip_matrix_copy(other.h, self.h), "matrix_copy")

    def info(self):
        nb, ne = C.c_int64(), C.c_int64()
        self.ctx.check(self.ctx.lib.lamb_hip_matrix_info(self.h, C.byref(nb), C.byref(ne)), "info")
        return nb.value, ne.value

    def trace(self) -> float:
        t = C.c_double()
        self.ctx.check(self.ctx.lib.lamb_hip_trace(self.h, C.byref(t)), "trace")
        return t.value

    def _per_group(self, fn, *args) -> np.ndarray:
        out = np.zeros(self.ddist.n_groups)
        self.ctx.check(fn(self.h, *args, N.ptr(out, C.c_double)), fn.__name__)
        return out

    def group_trace(self) -> np.ndarray:
        return self._per_group(self.ctx.lib.lamb_hip_group_trace)

    def group_norm(self) -> np.ndarray:
        return self._per_group(self.ctx.lib.lamb_hip_group_norm)

    def group_diff_norm(self, other: "DeviceMatrix") -> np.ndarray:
        return self._per_group(self.ctx.lib.lamb_hip_group_diff_norm, other.h)

    def group_gershgorin(self):
        lo = np.zeros(self.ddist.n_groups)
        hi = np.zeros(self.ddist.n_groups)
        rc = self.ctx.lib.lamb_hip_group_gershgorin(self.h, N.ptr(lo, C.c_double),
                                                    N.ptr(hi, C.c_double))
        self.ctx.check(rc, "group_gershgorin")
        return lo, hi

    def group_scale_shift(self, a, b):
        a = np.ascontiguousarray(np.broadcast_to(a, self.ddist.n_groups), dtype=np.float64)
        b = np.ascontiguousarray(np.broadcast_to(b, self.ddist.n_groups), dtype=np.float64)
        rc = self.ctx.lib.lamb_hip_group_scale_shift(self.h, N.ptr(a, C.c_double),
                                                     N.ptr(b, C.c_double))
        self.ctx.check(rc, "group_scale_shift")

    def close(self):
        if self.h:
            self.ctx.lib.lamb_hip_matrix_destroy(self.h)
            self.h = None


def partition_atoms(ctx: Context, coords, box, dims):
    """(subset_map int32, order_map int16) of wrapped FP32 coordinates over
    dims cubes, computed on the device (lamb_hip_partition_atoms)."""
    xyz = np.ascontiguousarray(coords, dtype=np.float32).reshape(-1, 3)
    h = np.ascontiguousarray(np.broadcast_to(np.asarray(box, np.float32), (3,)))
    d = np.ascontiguousarray(dims, dtype=np.int32)
    n = len(xyz)
    subset = np.zeros(n, np.int32)
    order = np.zeros(n, np.int16)
    rc = ctx.lib.lamb_hip_partition_atoms(ctx.h, n, N.ptr(xyz, C.c_float), N.ptr(h, C.c_float),
                                          N.ptr(d, C.c_int32), N.ptr(subset, C.c_int32),
                                          N.ptr(order, C.c_int16))
    ctx.check(rc, "partition_atoms")
    return subset, order


def tuning_order(ctx: Context) -> int:
    """Row bands per block of the tile order of the context's last multiply."""
    rb = C.c_int32()
    ctx.check(ctx.lib.lamb_hip_tuning_order(ctx.h, C.byref(rb)), "tuning_order")
    return rb.value


def tuning_launch(ctx: Context):
    """(launch form, workgroups launched, C tiles) of the last numeric kernel.
    Forms: 0 one workgroup per tile, 1 persistent grid over the interleaved
    tile order, 2 persistent grid with one range of the order per XCD."""
    p, w, t = C.c_int32(), C.c_int32(), C.c_int32()
    ctx.check(ctx.lib.lamb_hip_tuning_launch(ctx.h, C.byref(p), C.byref(w), C.byref(t)),
              "tuning_launch")
    return p.value, w.value, t.value


def tuning_mode(ctx: Context, tune_shapes: bool) -> None:
    """Let the autotuner choose the FP64 tile shape per workload class."""
    ctx.check(ctx.lib.lamb_hip_tuning_mode(ctx.h, 1 if tune_shapes else 0), "tuning_mode")


def tuning_save(ctx: Context, path) -> None:
    """Write the context's tuning table (shapes, orders, launch forms)."""
    ctx.check(ctx.lib.lamb_hip_tuning_save(ctx.h, os.fsencode(os.fspath(path))), "tuning_save")


def tuning_load(ctx: Context, path) -> None:
    """Read a tuning table written by tuning_save into the context."""
    ctx.check(ctx.lib.lamb_hip_tuning_load(ctx.h, os.fsencode(os.fspath(path))), "tuning_load")


def tuning_info(ctx: Context):
    """(FP64 kernel variant of the last multiply, workload classes tuned)."""
    v, n = C.c_int32(), C.c_int32()
    ctx.check(ctx.lib.lamb_hip_tuning_info(ctx.h, C.byref(v), C.byref(n)), "tuning_info")
    return v.value, n.value


def purification_start(h: DeviceMatrix) -> DeviceMatrix:
    """P0 = (H - lo I) / (hi - lo) per group, lo / hi the Gershgorin bounds."""
    lo, hi = h.group_gershgorin()
    p = DeviceMatrix(h.ddist)
    p.copy_from(h)
    span = hi - lo
    p.group_scale_shift(1.0 / span, -lo / span)
    return p


def mcweeny(p: DeviceMatrix, steps: int, *, eps_filter: float = 0.0, flags: int = FP64):
    """steps McWeeny iterations on p in place; returns (traces, idempotency
    errors), both [steps, n_groups], and the accumulated statistics."""
    ctx = p.ctx
    g = p.ddist.n_groups
    traces = np.zeros((steps, g))
    idem = np.zeros((steps, g))
    st = N.Stats()
    rc = ctx.lib.lamb_hip_mcweeny(ctx.h, p.h, int(steps), float(eps_filter), int(flags),
                                  N.ptr(traces, C.c_double), N.ptr(idem, C.c_double),
                                  C.byref(st))
    ctx.check(rc, "mcweeny")
    return traces, idem, st.as_dict()


class ChainGraph:
    """A McWeeny chain captured as a HIP graph (lamb_hip_mcweeny_capture)."""

    def __init__(self, p: DeviceMatrix, steps: int, *, eps_filter: float = 0.0,
                 flags: int = FP64):
        self.ctx = p.ctx
        self.steps = int(steps)
        self.n_groups = p.ddist.n_groups
        h = C.c_void_p()
        rc = self.ctx.lib.lamb_hip_mcweeny_capture(self.ctx.h, p.h, self.steps, float(eps_filter),
                                                   int(flags), C.byref(h))
        self.ctx.check(rc, "mcweeny_capture")
        self.h = h

    def replay(self, p_in: DeviceMatrix, p_out: DeviceMatrix):
        """Run the graph on p_in (the captured pattern) into p_out; returns
        (traces, idempotency errors) [steps, n_groups] and the device ms."""
        traces = np.zeros((self.steps, self.n_groups))
        idem = np.zeros((self.steps, self.n_groups))
        ms = C.c_double()
        rc = self.ctx.lib.lamb_hip_graph_replay(self.h, p_in.h, p_out.h,
                                                N.ptr(traces, C.c_double),
                                                N.ptr(idem, C.c_double), C.byref(ms))
        self.ctx.check(rc, "graph_replay")
        return traces, idem, ms.value

    def info(self):
        steps, sizes, flops = C.c_int32(), C.c_int64(), C.c_double()
        self.ctx.check(self.ctx.lib.lamb_hip_graph_info(self.h, C.byref(steps), C.byref(sizes),
                                                        C.byref(flops)), "graph_info")
        return dict(steps=steps.value, recorded_sizes=sizes.value, flops=flops.value)

    def close(self):
        if self.h:
            self.ctx.lib.lamb_hip_graph_destroy(self.h)
            self.h = None


def multiply(a: DeviceMatrix, b: DeviceMatrix, c: DeviceMatrix, *, alpha: float = 1.0,
             beta: float = 0.0, eps_filter: float = 0.0, flags: int = FP64) -> dict:
    """C = alpha*A*B + beta*C on the device; returns the call's statistics."""
    ctx = a.ctx
    st = N.Stats()
    rc = ctx.lib.lamb_hip_multiply(ctx.h, a.h, b.h, c.h, float(alpha), float(beta),
                                   float(eps_filter), int(flags), C.byref(st))
    ctx.check(rc, "multiply")
    return st.as_dict()


def multiply_test(ctx: Context, path, *, eps_filter: float = 0.0) -> dict:
    """lamb_multiply_test (src/lamb_multiplication.F90:85-114): the DBCSR dump
    at path as A, copied to B and C, then C = A*B; returns the statistics."""
    st = N.Stats()
    rc = ctx.lib.lamb_hip_multiply_test(ctx.h, os.fsencode(os.fspath(path)), float(eps_filter),
                                        C.byref(st))
    ctx.check(rc, "multiply_test")
    return st.as_dict()


def sector_multiply(a: DeviceMatrix, b: DeviceMatrix, c: DeviceMatrix, subset_m: int,
                    subset_k: int, subset_n: int, *, eps_filter: float = 0.0,
                    flags: int = FP64) -> dict:
    """C(m, n) += A(m, k) * B(k, n) with the prefilter and no post-filter
    (lamb_hip_sector_multiply; the reference's sector_multiply)."""
    st = N.Stats()
    rc = a.ctx.lib.lamb_hip_sector_multiply(a.ctx.h, a.h, b.h, c.h, subset_m, subset_k, subset_n,
                                            eps_filter, flags, C.byref(st))
    a.ctx.check(rc, f"sector_multiply({subset_m},{subset_k},{subset_n})")
    return st.as_dict()


def matrix_filter(c: DeviceMatrix, eps_filter: float) -> int:
    """Drop C's blocks with nsq < eps_filter^2; returns how many were dropped."""
    n = C.c_int64()
    c.ctx.check(c.ctx.lib.lamb_hip_matrix_filter(c.h, eps_filter, C.byref(n)), "matrix_filter")
    return n.value


class Plexus:
    """Sector ownership of process my_proc of n_procs (lamb_hip_plexus_*):
    region pools, realm map, catalog, orbit, sectors and multiply schedule.
    Regions / subsets are 1-based, processes 0-based."""

    def __init__(self, n_procs: int, my_proc: int, n_regions: int = 0):
        self.lib = N.load()
        h = C.c_void_p()
        rc = self.lib.lamb_hip_plexus_create(n_procs, my_proc, n_regions, C.byref(h))
        if rc != 0:
            raise LambError(rc, 100, f"plexus_create(n_procs={n_procs}, my_proc={my_proc}, "
                                     f"n_regions={n_regions})")
        self.h = h
        self.n_procs = n_procs
        self.my_proc = my_proc
        nr, ncat, norb, nsec = (C.c_int32() for _ in range(4))
        nsch = C.c_int64()
        self.lib.lamb_hip_plexus_info(h, C.byref(nr), C.byref(ncat), C.byref(norb),
                                      C.byref(nsec), C.byref(nsch))
        self.n_regions = nr.value
        n = self.n_regions
        rmap = np.zeros(n * n, np.int32)
        psz = np.zeros(n, np.int32)
        pools = np.zeros(n * n_procs, np.int32)
        burden = np.zeros(n_procs, np.int32)
        self.lib.lamb_hip_plexus_maps(h, N.ptr(rmap, C.c_int32), N.ptr(psz, C.c_int32),
                                      N.ptr(pools, C.c_int32), N.ptr(burden, C.c_int32))
        self.realm_map = rmap.reshape(n, n)  # [r - 1, s - 1]
        pools = pools.reshape(n, n_procs)
        self.pools = {r + 1: pools[r, :psz[r]].tolist() for r in range(n)}
        self.burden = burden.tolist()
        cat = np.zeros(ncat.value, np.int32)
        orb = np.zeros(norb.value, np.int32)
        sec = np.zeros(2 * nsec.value, np.int32)
        self.lib.lamb_hip_plexus_lists(h, N.ptr(cat, C.c_int32), N.ptr(orb, C.c_int32),
                                       N.ptr(sec, C.c_int32))
        self.catalog = cat.tolist()
        self.orbit = orb.tolist()
        self.sectors = [tuple(x) for x in sec.reshape(-1, 2).tolist()]
        sch = np.zeros(6 * nsch.value, np.int32)
        self.lib.lamb_hip_plexus_sched(h, N.ptr(sch, C.c_int32))
        self.sched = [tuple(x) for x in sch.reshape(-1, 6).tolist()]

    def close(self):
        if self.h:
            self.lib.lamb_hip_plexus_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
