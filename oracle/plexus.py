"""Restatement of the reference's plexus (sector ownership maps), for tests.

TEST INFRASTRUCTURE ONLY: the product path never imports this module.

Follows src/lamb_plexus_operations.F90:
- n_regions from the process count: int((sqrt(1 + 8P) - 1) / 2 * 2) in
  single precision (:115), or a forced count;
- fill_region_pools (:198-320): a recursive walk over the realms (region
  pairs r <= s) -- trr_tr on the diagonal triangle, trr_sq on the square
  between two halves, quadrants in the order (lo,lo), (lo,hi), (hi,hi),
  (hi,lo), midpoints (first + last - 1) / 2 -- numbering realms 1, 2, ... in
  walk order; realm t goes to process mod(int(real(t - 1) * p_per_realms), P)
  with p_per_realms = real(P) / (real(nr) * (real(nr) + 1.0) / 2.0), all
  single precision; that process joins the pools of both regions (sorted,
  no repeats, add_to_pool :322-347);
- catalog(p) = the regions whose pool holds p; orbit(p) = the union of the
  pools of its catalog (:246-253);
- fill_realm_map (:378-449): for r1 <= r2 in row order, the common processes
  of the two pools (a merge of the sorted lists); the least burdened of them,
  scanning from the last down and replacing only on a strictly smaller burden
  (so ties go to the highest process id); map(r1, r2) = map(r2, r1) = it;
  its burden grows by 2;
- fill_my_sectors (:454-488): (r, s) over catalog x catalog in order with
  map(r, s) = p;
- lamb_plexus_create_sched (:492-571): for each sector (m, n) of p and each
  k = 1..n_regions, one era holding GET A(m, k) from map(m, k) and GET
  B(k, n) from map(k, n).

Processes are 0-based, regions and subsets 1-based, as in the reference.
Pinned by the hand-derived plexus for P = 1 and P = 3 written out in
tests/test_plexus.py and by the invariants checked there.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32


def n_regions_for(n_procs: int) -> int:
    root = (np.sqrt(F32(1 + 8 * n_procs), dtype=F32) - F32(1.0)) / F32(2.0) * F32(2.0)
    return int(root)


def _add_sorted(lst: list, val: int):
    if val not in lst:
        lst.append(val)
        lst.sort()


def region_pools(n_regions: int, n_procs: int):
    """(pools[r] sorted process lists for r = 1..n, realm -> process in walk
    order as a list of (r, s, p))."""
    nr = F32(n_regions)
    p_per_realms = F32(n_procs) / (nr * (nr + F32(1.0)) / F32(2.0))
    pools = {r: [] for r in range(1, n_regions + 1)}
    realms = []
    counter = [1]

    def which_p(t):
        return int(F32(t - 1) * p_per_realms) % n_procs

    def reg1(r):
        p = which_p(counter[0])
        _add_sorted(pools[r], p)
        realms.append((r, r, p))
        counter[0] += 1

    def reg2(r, s):
        p = which_p(counter[0])
        _add_sorted(pools[r], p)
        if r != s:
            _add_sorted(pools[s], p)
        realms.append((r, s, p))
        counter[0] += 1

    def tr(a, b):
        if b < a:
            return
        if a == b:
            reg1(a)
            return
        m = (a + b - 1) // 2
        tr(a, m)
        sq(a, m, m + 1, b)
        tr(m + 1, b)

    def sq(r0, r1, s0, s1):
        if r1 < r0 or s1 < s0:
            return
        if r0 == r1 and s0 == s1:
            reg2(r0, s0)
            return
        rm = (r0 + r1 - 1) // 2
        sm = (s0 + s1 - 1) // 2
        sq(r0, rm, s0, sm)
        sq(r0, rm, sm + 1, s1)
        sq(rm + 1, r1, sm + 1, s1)
        sq(rm + 1, r1, s0, sm)

    tr(1, n_regions)
    return pools, realms


def realm_map(pools: dict, n_regions: int, n_procs: int):
    """(map[n+1][n+1] with 1-based regions, burdens per process)."""
    m = np.full((n_regions + 1, n_regions + 1), -1, np.int64)
    burden = [0] * n_procs
    for r1 in range(1, n_regions + 1):
        for r2 in range(r1, n_regions + 1):
            a, b = pools[r1], pools[r2]
            i = j = 0
            conj = []
            while i < len(a) and j < len(b):
                if a[i] == b[j]:
                    conj.append(a[i])
                    i += 1
                    j += 1
                elif a[i] > b[j]:
                    j += 1
                else:
                    i += 1
            assert conj, "no common process of two regions"
            best = conj[-1]
            for q in reversed(conj[:-1]):
                if burden[q] < burden[best]:
                    best = q
            m[r1, r2] = m[r2, r1] = best
            burden[best] += 2
    return m, burden


class Plexus:
    """The maps of one process p of a P-process plexus."""

    def __init__(self, n_procs: int, my_proc: int, n_regions: int | None = None):
        self.n_procs = n_procs
        self.my_proc = my_proc
        self.n_regions = n_regions_for(n_procs) if n_regions is None else n_regions
        self.pools, self.realms = region_pools(self.n_regions, n_procs)
        self.catalog = [r for r in range(1, self.n_regions + 1) if my_proc in self.pools[r]]
        orbit = set()
        for r in self.catalog:
            orbit.update(self.pools[r])
        self.orbit = sorted(orbit)
        self.map, self.burden = realm_map(self.pools, self.n_regions, n_procs)
        self.sectors = [(r, s) for r in self.catalog for s in self.catalog
                        if self.map[r, s] == my_proc]

    def schedule(self):
        """[(era, src process, dst process, which (1 = A, 2 = B), row subset,
        col subset)] in the order lamb_plexus_create_sched adds them."""
        out = []
        era = 0
        for (m, n) in self.sectors:
            for k in range(1, self.n_regions + 1):
                era += 1
                out.append((era, int(self.map[m, k]), self.my_proc, 1, m, k))
                out.append((era, int(self.map[k, n]), self.my_proc, 2, k, n))
        return out
