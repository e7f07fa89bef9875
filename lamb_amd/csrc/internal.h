// Host-side objects behind the opaque C-ABI handles.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "common.h"

namespace lamb {

// Device allocation that only grows.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t want);
  hipError_t upload(hipStream_t s, const void* src, size_t n);
  void release();
};

struct StagedSector {
  int32_t sr = 0, sc = 0;
  std::vector<int64_t> row_p;
  std::vector<int16_t> col_i;
  std::vector<int64_t> ext;
  std::vector<double> data;
};

struct StagedList {
  std::vector<int32_t> subset_idx;
  std::vector<int16_t> local_idx;
  std::vector<int64_t> ext;
  std::vector<double> data;
};

int32_t set_error(lamb_hip_ctx* ctx, int32_t level, int32_t type, const char* fmt, ...);
int32_t hip_error(lamb_hip_ctx* ctx, hipError_t e, const char* what);

// lamb_hip_multiply restricted to the blocks of mask's pattern (mask may be
// nullptr: the plain multiply). post_eps >= 0 sets the post-filter threshold
// apart from the prefilter's eps (0 keeps every block); < 0 uses eps.
int32_t multiply_masked(lamb_hip_ctx* ctx, const lamb_hip_mat* A, const lamb_hip_mat* B,
                        lamb_hip_mat* C, double alpha, double beta, double eps, int32_t flags,
                        lamb_hip_stats* stats, const lamb_hip_mat* mask,
                        double post_eps = -1.0);

// Per-atom partials (chain.hip): ||X - Y||_F^2 (Y may be nullptr) and the
// diagonal sums, written to out[n_atoms] in internal atom order.
hipError_t launch_diff_partials(hipStream_t s, const lamb_hip_mat* X, const lamb_hip_mat* Y,
                                double* out);
hipError_t launch_trace_partials(hipStream_t s, const lamb_hip_mat* X, double* out);

// A matrix whose value pool was released (lamb_hip_matrix_release_values)
// while it stores blocks: entry points that read values refuse it.
bool values_missing(const lamb_hip_mat* m);
int32_t refuse_missing_values(lamb_hip_ctx* ctx, const char* what);

// Releases the device scratch of the symbolic phase held by a context.
void release_sym_state(lamb_hip_ctx* ctx);

// Device address of the no-allocation scratch array (noalloc.hip), taken once
// per device with that device current.
hipError_t noalloc_setup(int device);

// Recursive coordinate bisection order of n atoms (coords: xyz per atom).
void spatial_order(int32_t n, const float* coords, std::vector<int32_t>& order);

}  // namespace lamb

struct lamb_hip_ctx {
  int32_t device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[8] = {};
  int32_t level = 0;
  int32_t type = 0;
  std::string msg;
  // scratch reused across multiplies (symbolic.h)
  lamb::DevBuf d_kept, d_apack, d_bpack;
  lamb::DevBuf d_aexp, d_bexp;  // BF16 split: exponent per A row / B column
  lamb::DevBuf d_aslice, d_bslice;  // BF16 split: slice packs
  lamb::DevBuf d_slot_counter;      // persistent numeric grids: slot counters, XCD ranges
  void* sym = nullptr;  // lamb::SymState*
  // autotuner: tile order (row bands per block) and launch form (one
  // workgroup per tile, the resident workgroups looping over the order
  // interleaved, or over one range of it per XCD: lamb::kForm*) per
  // workload class (multiply.hip)
  struct TuneChoice {
    int32_t rowblock = 8;
    int32_t form = 0;
  };
  std::map<std::array<int64_t, 4>, TuneChoice> tune;
  // tile shape (FP64 variant: 32 x 32, 64 x 32 or 64 x 64 per wave) per
  // workload class, when shape tuning is on or a table was loaded
  int32_t tune_shapes = 0;
  std::map<std::array<int64_t, 5>, int32_t> shape_tune;
  int32_t last_variant = -1;   // FP64 variant of the last multiply
  int32_t last_rowblock = -1;  // tile-order row bands per block of the last multiply
  int32_t last_grid = -1;      // workgroups of the last numeric launch
  int32_t last_form = -1;      // its launch form
  int32_t last_tiles = -1;     // C tiles of the last multiply
  // Sizes the host reads back from the device (graph.hip). kSizesLive: read
  // them (stream sync); kSizesRecord: read and append to size_log;
  // kSizesReplay (during a graph capture): take them from size_log in order
  // and enqueue a check that raises d_guard when the device value differs.
  int32_t size_mode = 0;
  std::vector<int64_t> size_log;
  size_t size_cursor = 0;
  unsigned long long* guard = nullptr;  // the device's guard flag (install_guard)
};

namespace lamb {
constexpr int32_t kSizesLive = 0, kSizesRecord = 1, kSizesReplay = 2;
// n int64 sizes at device address src into dst, according to ctx->size_mode
hipError_t read_sizes(lamb_hip_ctx* ctx, const void* src, int64_t* dst, int n);
// Waits for the context stream, except while a graph is being captured.
hipError_t stream_sync(lamb_hip_ctx* ctx);
}  // namespace lamb

struct lamb_hip_dist {
  lamb_hip_ctx* ctx = nullptr;
  int32_t n = 0;
  int32_t n_subsets = 0;
  int32_t max_blk = 0;
  // block sizes present: bit s of size_mask[s / 64] for size s (0..64)
  uint64_t size_mask[2] = {0, 0};
  std::vector<int32_t> blk;        // by global atom
  std::vector<int32_t> subset_of;  // 1-based, by global atom
  std::vector<int32_t> local_of;   // 1-based, by global atom
  std::vector<std::vector<int32_t>> subset_atoms;  // [subset][local-1] -> global atom
  std::vector<int32_t> perm;       // global -> internal
  std::vector<int32_t> iperm;      // internal -> global
  std::vector<int32_t> blk_int;    // by internal atom
  std::vector<lamb::Band> bands;
  std::vector<int32_t> band_of;      // by internal atom
  std::vector<int32_t> pos_in_band;  // by internal atom
  lamb::DevBuf d_blk;    // int32 by internal atom
  lamb::DevBuf d_bands;  // Band[]
  lamb::DevBuf d_band_of, d_pos_in_band;  // int32 by internal atom
  lamb::DevBuf d_perm;   // int32 by global atom: its internal atom
  lamb::DevBuf d_elem0;  // int64 [n+1]: first matrix row of each internal atom
  int64_t n_elem = 0;    // matrix dimension (sum of block sizes)
  // groups of atoms (independent systems of a batch), by internal atom
  int32_t n_groups = 1;
  std::vector<int32_t> group_int;
  lamb::DevBuf d_group;
};

struct lamb_hip_mat {
  lamb_hip_dist* dist = nullptr;
  // internal block-CSR (rows/cols are internal atom indices, cols ascending);
  // the device arrays are authoritative, the host vectors a cache
  bool host_index_valid = true;
  std::vector<int64_t> row_p;
  std::vector<int32_t> col;
  std::vector<int64_t> off;
  std::vector<int32_t> rows;
  int64_t nblk = 0;
  int64_t ndata = 0;   // size of the data pool (may hold unused gaps)
  int64_t nelems = 0;  // values in stored blocks
  lamb::DevBuf d_data, d_row_p, d_col, d_off, d_rows, d_nsq;
  // spare data / index buffers a multiply writes C into before swapping
  // (kept only while a graph is recorded or captured)
  lamb::DevBuf s_data, s_row_p, s_col, s_off, s_rows;
  std::vector<double> nsq;  // host copy of the block norms
  bool nsq_valid = false;
  std::vector<double> host_data;
  bool host_data_valid = false;
  std::vector<lamb::StagedSector> staged_sectors;
  std::vector<lamb::StagedList> staged_lists;

  hipError_t upload_index(hipStream_t s);
  hipError_t download_index(hipStream_t s);
  void release_device();
  void release_spares();
  void invalidate_derived();
};
