// Device symbolic phase: strips, tiles / entries, C index.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "common.h"
#include "internal.h"

namespace lamb {

// device arrays of a distribution
struct DistView {
  int32_t n = 0;       // atoms
  int32_t nbands = 0;  // bands
  const int32_t* blk = nullptr;          // by internal atom
  const Band* bands = nullptr;
  const int32_t* band_of = nullptr;      // by internal atom
  const int32_t* pos_in_band = nullptr;  // by internal atom
};

// device arrays of a matrix (internal block-CSR)
struct MatView {
  int64_t nblk = 0;
  const int64_t* row_p = nullptr;
  const int32_t* col = nullptr;
  const int64_t* off = nullptr;
  const int32_t* rows = nullptr;
  const double* nsq = nullptr;
};

struct StripSet {
  int64_t n = 0;         // strips
  int64_t pk_total = 0;  // packed panel elements
  DevBuf strips;         // Strip[n]
  DevBuf head;           // StripHead[n]
  DevBuf ptr;            // int64: per band (A) or per row atom (B), +1
  DevBuf pk;             // int64[n+1]: packed panel offsets
  void release() {
    strips.release();
    head.release();
    ptr.release();
    pk.release();
  }
};

struct SymScratch {
  DevBuf merged, flag, excl, scan, gslot;
  DevBuf cnt, ksum, pat, stats, band4, band4s, cells, cellcnt, cellbase, mpat, cidx, tw;
  DevBuf flag_b, excl_b, gslot_b, sizes;  // B side of the strip build; counts read back
  void release() {
    for (DevBuf* b : {&merged, &flag, &excl, &scan, &gslot, &cnt, &ksum, &pat, &stats, &band4,
                      &band4s, &cells, &cellcnt, &cellbase, &mpat, &cidx, &tw, &flag_b, &excl_b,
                      &gslot_b, &sizes})
      b->release();
  }
};

struct SymResult {
  int64_t n_tiles = 0, n_entries = 0, c_size = 0, n_struct = 0;
  int64_t kept = 0, dropped = 0;
  int64_t n_partial = 0;  // entries flagged kEntryPartial
  double flops = 0.0;
  DevBuf tiles, entries, tile_base, perm;
  void release() {
    perm.release();
    tiles.release();
    entries.release();
    tile_base.release();
  }
};

struct CIndex {
  int64_t nblk = 0, nelems = 0;
  DevBuf row_p, col, off, rows;
};

// The strips of A (a_side) and of B, built side by side so that their counts
// and their packed panel sizes come back in one read each.
hipError_t build_strips(lamb_hip_ctx* ctx, SymScratch& W, const MatView& MA, const MatView& MB,
                        const DistView& D, StripSet& A, StripSet& B);
// mask (optional): only blocks in the mask matrix's pattern are produced
hipError_t build_tiles(lamb_hip_ctx* ctx, SymScratch& W, const DistView& D, const StripSet& A,
                       const StripSet& B, const MatView* Cold, double eps2, SymResult& r,
                       const MatView* mask = nullptr);
// perm[k] = tile run k-th by the numeric kernel (blocks of rb row bands,
// column band by column band)
hipError_t build_tile_order(lamb_hip_ctx* ctx, SymScratch& W, const DistView& D, SymResult& r,
                            int rb);
// xr[0..8]: the XCD ranges of the tile order perm. Residue x of a persistent
// grid walks positions [xr[x], xr[x + 1]), one contiguous eighth of the
// order by weight (each tile weighs its K columns, ktot).
hipError_t build_xcd_ranges(lamb_hip_ctx* ctx, SymScratch& W, const SymResult& r, int32_t* xr);
hipError_t build_c_index(lamb_hip_ctx* ctx, SymScratch& W, const DistView& D,
                         const SymResult& r, const uint16_t* kept, CIndex& c);

}  // namespace lamb
