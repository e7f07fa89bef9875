"""Host-side mirror of the reference's matrix, sector and distribution types.

The reference stores a matrix as a list of *sectors*: one sub-matrix per
(row subset, column subset) pair, subsets being spatial cubes of atoms.
This module restates, in numpy, the pieces that fix that layout:

* ``n_cubes_for(P)``  -- src/lamb_plexus_operations.F90:115
  (``root = (sqrt(real(1+8P)) - 1)/2*2``, single precision, truncated)
* ``partition_dims(n)`` -- src/lamb_partition_operations.F90:62-108
  (prime factors dealt round-robin over the three dimensions)
* ``which_cube``     -- src/lamb_partition_methods.F90:85-107
  (``mod(int(x*multiplier), n_d)`` in FP32, then 1 + cx + cy*nx + cz*nx*ny)
* ``set_partition``  -- src/lamb_set_partition_methods.F90:69-144
  (subset of each atom, local index = arrival order inside the subset)
* ``my_sectors``     -- src/lamb_plexus_operations.F90:454-488 at one process
  (catalog x catalog, row subset outer)
* sector index       -- src/lamb_index_types.F90:86-105,
  src/lamb_index_methods.F90:198-274: ``row_p`` int64 0-based exclusive
  offsets built by a counting sort of the rows, ``col_i`` int16 1-based
  local column atom, extents int64 (lo, hi) 1-based inclusive into the
  sector's FP64 store.
* ``get_blk_n``      -- src/lamb_index_methods.F90:73-107 (binary search in an
  ascending ``col_i`` row).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .systems import BlockCSR


def n_cubes_for(n_procs: int) -> int:
    """Number of spatial cubes for ``n_procs`` processes."""
    root = (np.sqrt(np.float32(1 + 8 * n_procs)) - np.float32(1.0)) / np.float32(2.0) * np.float32(2.0)
    return int(np.float32(0.0) + np.float32(root))


def partition_dims(n_cubes: int) -> tuple[int, int, int]:
    """Distribute the prime factors of ``n_cubes`` over three dimensions."""
    n = int(n_cubes)
    dims = [1, 1, 1]
    p = 0
    product = 1
    i = 2
    tgt = n
    while product < n:
        if tgt % i == 0:
            dims[p] *= i
            product *= i
            p = (p + 1) % 3
            if product >= n:
                break
            tgt //= i
        else:
            i += 1
    return tuple(dims)


def which_cube(coords: np.ndarray, dims, box) -> np.ndarray:
    """1-based cube index of each (wrapped, FP32) coordinate triple."""
    c = np.asarray(coords, dtype=np.float32)
    h = np.broadcast_to(np.asarray(box, dtype=np.float32), (3,))
    mult = np.asarray(dims, dtype=np.float32) / h  # real(ncubes, space_k) / h(d,d)
    cube = np.zeros(c.shape[:-1] + (3,), dtype=np.int64)
    for d in range(3):
        x = (c[..., d] * mult[d]).astype(np.float32)
        cube[..., d] = np.fmod(np.trunc(x).astype(np.int64), dims[d])
    return 1 + cube[..., 0] + cube[..., 1] * dims[0] + cube[..., 2] * dims[0] * dims[1]


def index_compress(rows_1based: np.ndarray, n_rows: int) -> np.ndarray:
    """Counting sort of 1-based row indices into an exclusive-offset row_p."""
    counts = np.bincount(np.asarray(rows_1based, dtype=np.int64), minlength=n_rows + 1)[1:n_rows + 1]
    row_p = np.zeros(n_rows + 1, dtype=np.int64)
    np.cumsum(counts, out=row_p[1:])
    return row_p


def get_blk_n(row_p: np.ndarray, col_i: np.ndarray, row: int, col: int):
    """1-based block number of (row, col) in a CSR index, and whether found."""
    first = int(row_p[row - 1]) + 1
    last = int(row_p[row])
    blk = (first + last) // 2
    while first <= blk <= last:
        c = int(col_i[blk - 1])
        if c == col:
            return blk, True
        if c < col:
            first = blk + 1
        else:
            last = blk - 1
        blk = (first + last) // 2
    return blk, False


@dataclasses.dataclass
class Distribution:
    """Atom set plus its spatial partition into subsets (one process)."""

    blk: np.ndarray          # int32 block size per global atom
    coords: np.ndarray       # float32 [n, 3] wrapped positions
    box: float
    n_procs: int = 1

    def __post_init__(self):
        self.blk = np.asarray(self.blk, dtype=np.int32)
        self.coords = np.asarray(self.coords, dtype=np.float32)
        self.n_cubes = n_cubes_for(self.n_procs)
        self.dims = partition_dims(self.n_cubes)
        self.n_subsets = int(np.prod(self.dims))
        subset = which_cube(self.coords, self.dims, self.box).astype(np.int32)
        # local index = arrival order within the subset (order_map)
        order = np.zeros(self.n_atoms, dtype=np.int32)
        count = np.zeros(self.n_subsets + 1, dtype=np.int64)
        for g in range(self.n_atoms):
            count[subset[g]] += 1
            order[g] = count[subset[g]]
        if count.max(initial=0) > np.iinfo(np.int16).max:
            raise ValueError("too many atoms in one subset for the int16 local index")
        self.subset_map = subset
        self.order_map = order.astype(np.int16)
        self.subset_atoms = [np.flatnonzero(subset == s + 1).astype(np.int64)
                             for s in range(self.n_subsets)]
        # one process owns every sector: catalog x catalog, row subset outer
        self.sectors = [(r, s) for r in range(1, self.n_subsets + 1)
                        for s in range(1, self.n_subsets + 1)]

    @property
    def n_atoms(self) -> int:
        return int(self.blk.shape[0])

    def sector_index(self, subsets) -> int:
        """1-based local sector index of a subset pair, 0 if absent
        (lamb_plexus_get_sector_i, src/lamb_plexus_methods.F90:111-132)."""
        for i, s in enumerate(self.sectors):
            if tuple(s) == tuple(subsets):
                return i + 1
        return 0


@dataclasses.dataclass
class Sector:
    """One sub-matrix (row subset, col subset) in the reference layout."""

    subsets: tuple
    row_p: np.ndarray    # int64 [n_rows + 1]
    col_i: np.ndarray    # int16 [n_blks], 1-based
    extents: np.ndarray  # int64 [n_blks, 2], 1-based inclusive
    data: np.ndarray     # float64

    @property
    def n_blks(self) -> int:
        return int(self.col_i.shape[0])


class LambMatrix:
    """A matrix as the reference holds it: a list of sectors."""

    def __init__(self, dist: Distribution, sectors: list[Sector] | None = None):
        self.dist = dist
        self.sectors = sectors if sectors is not None else [
            Sector(s, np.zeros(len(dist.subset_atoms[s[0] - 1]) + 1, np.int64),
                   np.zeros(0, np.int16), np.zeros((0, 2), np.int64), np.zeros(0))
            for s in dist.sectors]
        self.device = None  # device-side copy, managed by lamb_amd.multiplication

    # -- conversions -------------------------------------------------------
    @classmethod
    def from_global(cls, dist: Distribution, csr: BlockCSR) -> "LambMatrix":
        rows = csr.rows().astype(np.int64)
        cols = csr.col.astype(np.int64)
        sr = dist.subset_map[rows]
        sc = dist.subset_map[cols]
        lr = dist.order_map[rows].astype(np.int64)
        lc = dist.order_map[cols].astype(np.int64)
        order = np.lexsort((lc, lr, sc, sr))
        sizes = (dist.blk[rows] * dist.blk[cols]).astype(np.int64)
        sectors = []
        for (r, s) in dist.sectors:
            sel = order[(sr[order] == r) & (sc[order] == s)]
            n_rows = len(dist.subset_atoms[r - 1])
            row_p = index_compress(lr[sel], n_rows)
            col_i = lc[sel].astype(np.int16)
            sz = sizes[sel]
            ends = np.cumsum(sz)
            ext = np.stack([ends - sz + 1, ends], axis=1).astype(np.int64)
            total = int(ends[-1]) if len(ends) else 0
            if total:
                starts = csr.blk_off[sel]
                idx = np.repeat(starts - (ends - sz), sz) + np.arange(total, dtype=np.int64)
                data = csr.data[idx]
            else:
                data = np.zeros(0)
            sectors.append(Sector((r, s), row_p, col_i, ext, data))
        return cls(dist, sectors)

    def to_global(self) -> BlockCSR:
        dist = self.dist
        rr, cc, src_off, sizes, parts = [], [], [], [], []
        base = 0
        for sec in self.sectors:
            r, s = sec.subsets
            n_rows = len(sec.row_p) - 1
            lrow = np.repeat(np.arange(n_rows), np.diff(sec.row_p))
            g_r = dist.subset_atoms[r - 1][lrow]
            g_c = dist.subset_atoms[s - 1][sec.col_i.astype(np.int64) - 1]
            rr.append(g_r)
            cc.append(g_c)
            src_off.append(sec.extents[:, 0] - 1 + base)
            sizes.append(sec.extents[:, 1] - sec.extents[:, 0] + 1)
            parts.append(sec.data)
            base += len(sec.data)
        rows = np.concatenate(rr) if rr else np.zeros(0, np.int64)
        cols = np.concatenate(cc) if cc else np.zeros(0, np.int64)
        soff = np.concatenate(src_off) if src_off else np.zeros(0, np.int64)
        sz = np.concatenate(sizes) if sizes else np.zeros(0, np.int64)
        alldata = np.concatenate(parts) if parts else np.zeros(0)
        order = np.lexsort((cols, rows))
        rows, cols, soff, sz = rows[order], cols[order], soff[order], sz[order]
        ends = np.cumsum(sz)
        total = int(ends[-1]) if len(ends) else 0
        idx = np.repeat(soff - (ends - sz), sz) + np.arange(total, dtype=np.int64)
        data = alldata[idx] if total else np.zeros(0)
        row_p = np.zeros(dist.n_atoms + 1, np.int64)
        np.add.at(row_p, rows + 1, 1)
        row_p = np.cumsum(row_p)
        return BlockCSR(dist.blk.copy(), row_p, cols.astype(np.int32),
                        (ends - sz).astype(np.int64), data)

    @property
    def n_blks(self) -> int:
        return sum(s.n_blks for s in self.sectors)
