// Internal declarations shared by the C-ABI layer and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/lamb_hip.h"

namespace lamb {

// Tile geometry of the numeric kernel: C tiles are (row band) x (col band),
// a band being up to kBandRows rows made of whole atoms, at most
// kBandAtoms atoms.
constexpr int kBandRows = 64;
constexpr int kBandAtoms = 16;
// K columns staged in LDS per chunk of the numeric kernel.
constexpr int kChunkK = 16;
// LDS row pitch (doubles) of the staged panels; +1 breaks bank aliasing.
constexpr int kPitch = kBandRows + 1;

// One (tile, k) strip pair of the numeric kernel. a_col / b_col are the
// first panel columns of the packed A strip [kdim][64] and packed B strip
// [kdim][64] (element offset = col * 64); the entry covers K columns
// [kstart, next entry's kstart) of its tile (the tile's ktot for the last).
// rmask / cmask select the atoms of the row / column band whose blocks take
// part (bit = position of the atom inside its band).
// a_col: A panel column of the entry's first K column; bit 31 (kEntryPartial)
// is set when the prefilter or an output mask cut the entry's row / column
// atoms below its strips' masks. Only such entries need per-lane masking:
// the panels already hold zeros outside the strips' atoms.
constexpr uint32_t kEntryPartial = 0x80000000u;

struct Entry {
  uint32_t a_col;
  uint32_t b_col;
  int32_t kstart;
  uint16_t rmask;
  uint16_t cmask;
};
static_assert(sizeof(Entry) == 16, "Entry layout");

struct TileDesc {
  int32_t I, J;       // row band, col band
  int32_t e0, ne;     // entries [e0, e0+ne)
  int32_t ktot;       // sum of k_k over the entries
  int32_t pad;
  int64_t out_base;   // element offset of this tile's region in the C pool
  uint16_t pattern[kBandAtoms];  // structural C pattern: bit j of row i
};
static_assert(sizeof(TileDesc) == 64, "TileDesc layout");

// Strips: the blocks of one band restricted to one atom k.
// A strip (I, k): blocks A_ik for atoms i of row band I.
// B strip (k, J): blocks B_kj for atoms j of col band J.
// off[pos] is the element offset of the block at band position pos.
struct Strip {
  int32_t k;       // internal atom index of the shared K atom
  int32_t band;    // I (A strips) or J (B strips)
  uint16_t mask;   // positions present
  uint16_t pad0;
  int32_t pad1;
  double min_nsq;  // smallest nsq over the present blocks
  int64_t off[kBandAtoms];
  double nsq[kBandAtoms];
};

// The fields of a strip that the symbolic kernels read for every strip pair
// (16 bytes in place of the 280 of a Strip): its band, atom mask, the rows
// (A) or columns (B) of the band the mask covers, and its smallest block norm.
struct StripHead {
  int32_t band;
  uint16_t mask;
  uint16_t rows;
  double min_nsq;
};
static_assert(sizeof(StripHead) == 16, "StripHead layout");

struct Band {
  int32_t first;  // first internal atom
  int32_t natoms;
  int16_t rs[kBandAtoms + 1];  // row start per position; rs[natoms] = rows
  int16_t pad[7];
};
static_assert(sizeof(Band) == 56, "Band layout");

// Graph replays (graph.hip) must stop all later work once a replayed size
// differs from the recorded one: every kernel that can run inside a captured
// chain starts with LAMB_GUARD, which returns when the device's guard flag is
// raised. Each translation unit holds its own pointer to the flag, set by
// its lamb_guard_set_<unit> (all of them called by install_guard).
#define LAMB_GUARD_UNIT(unit)                                                  \
  static __device__ const unsigned long long* lamb_guard_flag = nullptr;       \
  hipError_t lamb_guard_set_##unit(const unsigned long long* p) {              \
    return hipMemcpyToSymbol(HIP_SYMBOL(lamb_guard_flag), &p, sizeof(p));      \
  }
#define LAMB_GUARD                                   \
  do {                                               \
    if (lamb_guard_flag && *lamb_guard_flag) return; \
  } while (0)
hipError_t lamb_guard_set_symbolic(const unsigned long long* p);
hipError_t lamb_guard_set_scan(const unsigned long long* p);
hipError_t lamb_guard_set_kernels(const unsigned long long* p);
hipError_t lamb_guard_set_bf16split(const unsigned long long* p);
hipError_t lamb_guard_set_chain(const unsigned long long* p);
hipError_t lamb_guard_set_adjoint(const unsigned long long* p);
// the guard flag of a device (allocated and installed on first use)
hipError_t install_guard(int device, unsigned long long** flag);

// Kernel launchers (kernels.hip).
// Pack strips into dense [kdim][64] panels: A strips put band rows along the
// 64 (row_side = true), B strips put band columns along the 64.
hipError_t launch_pack_strips(hipStream_t s, int64_t n_strips, const Strip* strips,
                              const int64_t* pk_off, const Band* bands, const int32_t* d_blk,
                              const double* data, bool row_side, double* packed);
hipError_t launch_block_nsq(hipStream_t s, int64_t nblk, const int32_t* d_rows,
                            const int32_t* d_cols, const int32_t* d_blk, const int64_t* d_off,
                            const double* d_data, double* d_nsq);
// Launch forms of the numeric kernels; every form computes every tile the
// same way (same bits).
// - kFormGrid: one workgroup per tile, workgroup b on position b of the tile
//   order.
// - kFormPersistent: the device's resident workgroups, each looping over
//   positions x, x + 8, x + 16, ... for its residue x = blockIdx.x mod 8 (the
//   XCD the hardware deals it to): each XCD walks the positions it walks in
//   kFormGrid.
// - kFormXcdRanges: resident workgroups too, residue x walking the range
//   [xr[x], xr[x + 1]) of the order alone (xr from build_xcd_ranges), so an
//   XCD's tiles share row and column bands in its own L2.
// The persistent forms take their later positions from eight device
// counters (slot_counter[0..7], reset by the launch) and read the ranges at
// slot_counter + kXcdRangeWords; the buffer holds kSlotControlWords int32.
constexpr int kFormGrid = 0;
constexpr int kFormPersistent = 1;
constexpr int kFormXcdRanges = 2;
constexpr int kXcdRangeWords = 16;
constexpr int kSlotControlWords = 32;
// FP64 numeric kernel variants: per-wave 32 x 32 quadrants, per-wave 64 x 32
// halves with two waves splitting each half's K steps, or whole-tile waves
// splitting the K steps four ways. whole: the product has no partial entry
// (the half and K-split kernels then run without per-lane masks). form: one
// of the launch forms above; grid_used gets the grid launched.
constexpr int kF64Quadrant = 0;
constexpr int kF64Ksplit = 1;
constexpr int kF64Half = 2;
constexpr int kF64Lanes = 3;  // vector-ALU FMAs, K columns over lane groups (whole products)
constexpr int kF64Fine = 4;   // half kernel on 4x4x4 MFMAs, 16-row x 4-column groups (whole)
// Panels addressable by 32-bit byte offsets (a_zero, b_zero: the zero column
// after each side's panels)
bool lanes_panels_fit(uint32_t a_zero, uint32_t b_zero);
hipError_t launch_numeric_f64(hipStream_t s, int32_t n_tiles, const TileDesc* tiles,
                              const int32_t* perm, const Entry* entries, const Band* bands,
                              const double* a_pack, const double* b_pack, uint32_t a_zero,
                              uint32_t b_zero, double alpha, double beta, double eps2,
                              const int64_t* old_row_p, const int32_t* old_col,
                              const int64_t* old_off, const double* old_data, double* c_data,
                              uint16_t* kept, int variant, bool whole, int form,
                              int32_t* slot_counter, int32_t* grid_used);
// Enqueues the reset of a persistent grid's eight slot counters (one per
// residue of blockIdx.x mod 8) for a grid of `grid` workgroups.
void launch_slot_counter_reset(hipStream_t s, int32_t* counter, int32_t grid);
// Workgroups of a 256-thread kernel resident on the current device at once
// (the grid of a persistent launch; 0 when unknown).
int32_t resident_workgroups(const void* kernel);
// BF16-split numeric phase (bf16split.hip). ex: int32 per matrix row (A,
// row_side) or column (B); slice packs: slices x (n_chunks + 1) x 64 x 16 B.
hipError_t launch_slice_exponents(hipStream_t s, int64_t nblk, const int32_t* rows,
                                  const int32_t* cols, const int32_t* blk, const int64_t* elem0,
                                  const int64_t* off, const double* data, bool row_side,
                                  int64_t n_elem, int32_t* ex);
hipError_t launch_pack_slices(hipStream_t s, int slices, int64_t n_strips, const Strip* strips,
                              const int64_t* pk_off, const Band* bands, const int32_t* blk,
                              const int64_t* elem0, const int32_t* ex, const double* data,
                              bool row_side, int64_t n_chunks, void* out);
hipError_t launch_numeric_bf16(hipStream_t s, int slices, int32_t n_tiles, const TileDesc* tiles,
                               const int32_t* perm, const Entry* entries, const Band* bands,
                               const void* a_sl, const void* b_sl, uint32_t a_zero,
                               uint32_t b_zero, double alpha, double beta, double eps2,
                               const int64_t* old_row_p, const int32_t* old_col,
                               const int64_t* old_off, const double* old_data, double* c_data,
                               uint16_t* kept, int form, int32_t* slot_counter,
                               int32_t* grid_used);
hipError_t launch_trace(hipStream_t s, int32_t n_atoms, const int64_t* row_p,
                        const int32_t* col, const int64_t* off, const int32_t* blk,
                        const double* data, double* out);

}  // namespace lamb
