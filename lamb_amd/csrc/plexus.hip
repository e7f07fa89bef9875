// Sector ownership maps (the reference's plexus) for a forced process count:
// which regions each process lies in, which process owns each sector, and
// the schedule of operand sectors a process fetches for its multiply.
//
// Follows src/lamb_plexus_operations.F90:87-571 (lamb_plexus_calculate with
// force_n_procs, fill_region_pools, fill_realm_map, fill_my_sectors,
// lamb_plexus_create_sched). These are small sequential greedy walks over
// the n(n+1)/2 region pairs, run once per distribution, so they are host
// code; the sector products they schedule run on the device
// (lamb_hip_sector_multiply). Processes are 0-based, regions 1-based.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <iterator>
#include <vector>

#include "../../include/lamb_hip.h"

struct lamb_hip_plexus {
  int32_t n_procs = 0, my_proc = 0, n_regions = 0;
  std::vector<std::vector<int32_t>> pools;  // [region - 1]: sorted processes
  std::vector<int32_t> realm_map;           // n_regions^2, row-major, (r-1, s-1)
  std::vector<int32_t> burdens;             // per process
  std::vector<int32_t> catalog, orbit;      // sorted
  std::vector<int32_t> sectors;             // (r, s) pairs
  std::vector<int32_t> sched;               // 6 per entry
};

namespace {

void add_sorted(std::vector<int32_t>& v, int32_t x) {
  auto it = std::lower_bound(v.begin(), v.end(), x);
  if (it == v.end() || *it != x) v.insert(it, x);
}

// Region pools: the realms (r <= s) in the order of the reference's recursive
// walk (trr_tr over the triangle, trr_sq over squares, quadrants (lo,lo),
// (lo,hi), (hi,hi), (hi,lo), midpoint (first + last - 1) / 2), realm t on
// process mod(int(real(t - 1) * p_per_realms), P) in single precision.
struct PoolWalk {
  int32_t n_procs;
  float p_per_realms;
  int64_t realm = 1;
  std::vector<std::vector<int32_t>>* pools;

  int32_t which_p() const {
    const float x = (float)(realm - 1) * p_per_realms;
    return (int32_t)((int64_t)x % n_procs);
  }
  void reg(int32_t r, int32_t s) {
    const int32_t p = which_p();
    add_sorted((*pools)[r - 1], p);
    if (r != s) add_sorted((*pools)[s - 1], p);
    ++realm;
  }
  void sq(int32_t r0, int32_t r1, int32_t s0, int32_t s1) {
    if (r1 < r0 || s1 < s0) return;
    if (r0 == r1 && s0 == s1) {
      reg(r0, s0);
      return;
    }
    const int32_t rm = (r0 + r1 - 1) / 2, sm = (s0 + s1 - 1) / 2;
    sq(r0, rm, s0, sm);
    sq(r0, rm, sm + 1, s1);
    sq(rm + 1, r1, sm + 1, s1);
    sq(rm + 1, r1, s0, sm);
  }
  void tr(int32_t a, int32_t b) {
    if (b < a) return;
    if (a == b) {
      reg(a, a);
      return;
    }
    const int32_t m = (a + b - 1) / 2;
    tr(a, m);
    sq(a, m, m + 1, b);
    tr(m + 1, b);
  }
};

}  // namespace

extern "C" {

int32_t lamb_hip_plexus_create(int32_t n_procs, int32_t my_proc, int32_t n_regions,
                               lamb_hip_plexus** out) {
  if (!out) return LAMB_HIP_FAILURE;
  *out = nullptr;
  if (n_procs < 1 || my_proc < 0 || my_proc >= n_procs || n_regions < 0) return LAMB_HIP_FAILURE;
  if (n_regions == 0) {
    // n_cubes = int((sqrt(1 + 8P) - 1) / 2 * 2) in single precision (:115)
    const float root = (std::sqrt((float)(1 + 8 * (int64_t)n_procs)) - 1.0f) / 2.0f * 2.0f;
    n_regions = (int32_t)root;
  }
  if (n_regions < 1 || n_regions > 32767) return LAMB_HIP_FAILURE;  // cube_k = int16
  auto* px = new lamb_hip_plexus();
  px->n_procs = n_procs;
  px->my_proc = my_proc;
  px->n_regions = n_regions;
  px->pools.assign(n_regions, {});
  const float nr = (float)n_regions;
  PoolWalk w{n_procs, (float)n_procs / (nr * (nr + 1.0f) / 2.0f), 1, &px->pools};
  w.tr(1, n_regions);
  for (int32_t r = 1; r <= n_regions; ++r)
    if (std::binary_search(px->pools[r - 1].begin(), px->pools[r - 1].end(), my_proc))
      px->catalog.push_back(r);
  for (int32_t r : px->catalog)
    for (int32_t q : px->pools[r - 1]) add_sorted(px->orbit, q);
  // realm map: the least burdened common process of the two pools, scanned
  // from the last down, replaced only on a strictly smaller burden
  px->realm_map.assign((size_t)n_regions * n_regions, -1);
  px->burdens.assign(n_procs, 0);
  std::vector<int32_t> conj;
  for (int32_t r1 = 1; r1 <= n_regions; ++r1)
    for (int32_t r2 = r1; r2 <= n_regions; ++r2) {
      const auto& a = px->pools[r1 - 1];
      const auto& b = px->pools[r2 - 1];
      conj.clear();
      std::set_intersection(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(conj));
      if (conj.empty()) {
        delete px;
        return LAMB_HIP_FATAL;  // the walk gives every realm a process of both pools
      }
      int32_t best = conj.back();
      for (size_t i = conj.size() - 1; i-- > 0;)
        if (px->burdens[conj[i]] < px->burdens[best]) best = conj[i];
      px->realm_map[(size_t)(r1 - 1) * n_regions + (r2 - 1)] = best;
      px->realm_map[(size_t)(r2 - 1) * n_regions + (r1 - 1)] = best;
      px->burdens[best] += 2;
    }
  for (int32_t r : px->catalog)
    for (int32_t s : px->catalog)
      if (px->realm_map[(size_t)(r - 1) * n_regions + (s - 1)] == my_proc) {
        px->sectors.push_back(r);
        px->sectors.push_back(s);
      }
  // schedule: per sector (m, n) and k = 1..n_regions one era fetching A(m, k)
  // from its owner and B(k, n) from its owner
  int32_t era = 0;
  for (size_t i = 0; i < px->sectors.size(); i += 2) {
    const int32_t m = px->sectors[i], n = px->sectors[i + 1];
    for (int32_t k = 1; k <= n_regions; ++k) {
      ++era;
      const int32_t sa = px->realm_map[(size_t)(m - 1) * n_regions + (k - 1)];
      const int32_t sb = px->realm_map[(size_t)(k - 1) * n_regions + (n - 1)];
      px->sched.insert(px->sched.end(), {era, sa, my_proc, 1, m, k});
      px->sched.insert(px->sched.end(), {era, sb, my_proc, 2, k, n});
    }
  }
  *out = px;
  return LAMB_HIP_OK;
}

int32_t lamb_hip_plexus_destroy(lamb_hip_plexus* px) {
  delete px;
  return LAMB_HIP_OK;
}

int32_t lamb_hip_plexus_info(const lamb_hip_plexus* px, int32_t* n_regions,
                             int32_t* catalog_size, int32_t* orbit_size, int32_t* n_sectors,
                             int64_t* n_sched) {
  if (!px) return LAMB_HIP_FAILURE;
  if (n_regions) *n_regions = px->n_regions;
  if (catalog_size) *catalog_size = (int32_t)px->catalog.size();
  if (orbit_size) *orbit_size = (int32_t)px->orbit.size();
  if (n_sectors) *n_sectors = (int32_t)(px->sectors.size() / 2);
  if (n_sched) *n_sched = (int64_t)(px->sched.size() / 6);
  return LAMB_HIP_OK;
}

int32_t lamb_hip_plexus_maps(const lamb_hip_plexus* px, int32_t* realm_map, int32_t* pool_sizes,
                             int32_t* pools, int32_t* burdens) {
  if (!px) return LAMB_HIP_FAILURE;
  const int32_t n = px->n_regions;
  if (realm_map) std::copy(px->realm_map.begin(), px->realm_map.end(), realm_map);
  for (int32_t r = 0; r < n; ++r) {
    const auto& p = px->pools[r];
    if (pool_sizes) pool_sizes[r] = (int32_t)p.size();
    if (pools) {
      int32_t* row = pools + (size_t)r * px->n_procs;
      std::fill(row, row + px->n_procs, -1);
      std::copy(p.begin(), p.end(), row);
    }
  }
  if (burdens) std::copy(px->burdens.begin(), px->burdens.end(), burdens);
  return LAMB_HIP_OK;
}

int32_t lamb_hip_plexus_lists(const lamb_hip_plexus* px, int32_t* catalog, int32_t* orbit,
                              int32_t* sectors) {
  if (!px) return LAMB_HIP_FAILURE;
  if (catalog) std::copy(px->catalog.begin(), px->catalog.end(), catalog);
  if (orbit) std::copy(px->orbit.begin(), px->orbit.end(), orbit);
  if (sectors) std::copy(px->sectors.begin(), px->sectors.end(), sectors);
  return LAMB_HIP_OK;
}

int32_t lamb_hip_plexus_sched(const lamb_hip_plexus* px, int32_t* entries) {
  if (!px) return LAMB_HIP_FAILURE;
  if (entries) std::copy(px->sched.begin(), px->sched.end(), entries);
  return LAMB_HIP_OK;
}

}  // extern "C"
