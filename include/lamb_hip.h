/*
 * lamb_hip.h -- C-ABI of the MI355X (gfx950) sector block-sparse multiply.
 *
 * Drop-in boundary for the reference's Fortran 2003 multiplication path.
 * Every entry point is extern "C", takes plain pointers and fixed-width
 * integers, and returns an int32 status: 0 = ok, otherwise the error level
 * of the reference's error module (src/lamb_error.F90:116-137):
 *   1 note, 2 warning, 3 failure, 4 memory, 5 fatal.
 * The error *type* and a message are available from lamb_hip_last_error;
 * types follow src/lamb_error.F90:141-159 (caller 1, wrong_args 100,
 * precondition 200, internal -1, overflow -2, unimplemented -1000,
 * external -2000).
 *
 * Integer widths follow the reference's kinds (src/lamb_kinds.F90:155-179):
 * allatm_k int64 (global atom counts), locatm_k int16 (subset-local atom
 * index), subset_k int32, blksz_k int8 (block dimension), blkn_k int64
 * (block counts / CSR offsets), dataptr_k int64 (data extents),
 * space_k float32 (coordinates).
 *
 * Ownership: the caller owns every host array it passes; the library copies
 * what it needs into device memory before returning and keeps no host
 * pointer (the reference's data_X%mine = .true. semantics,
 * src/lamb_data_X_types.F90:77-83). Device memory belongs to the handle
 * that allocated it and is released by the matching _destroy.
 *
 * Threading: one host thread per context. Calls are ordered on the
 * context's HIP stream and return after the device work they start has
 * finished (blocking), like the reference's synchronous lamb_multiply.
 */
#ifndef LAMB_HIP_H
#define LAMB_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lamb_hip_ctx lamb_hip_ctx;
typedef struct lamb_hip_dist lamb_hip_dist;
typedef struct lamb_hip_mat lamb_hip_mat;

/* error levels, src/lamb_error.F90:116-137 */
#define LAMB_HIP_OK 0
#define LAMB_HIP_NOTE 1
#define LAMB_HIP_WARNING 2
#define LAMB_HIP_FAILURE 3
#define LAMB_HIP_MEMORY 4
#define LAMB_HIP_FATAL 5

/* error types, src/lamb_error.F90:141-159 */
#define LAMB_HIP_CALLER_ERROR 1
#define LAMB_HIP_WRONG_ARGS 100
#define LAMB_HIP_PRECONDITION_FAILED 200
#define LAMB_HIP_INTERNAL_ERROR (-1)
#define LAMB_HIP_OVERFLOW_ERROR (-2)
#define LAMB_HIP_UNIMPLEMENTED (-1000)
#define LAMB_HIP_EXTERNAL_ERROR (-2000)

/* multiply flags: mode in bits 0-7, BF16 slice count in bits 8-11 */
#define LAMB_HIP_FP64 0          /* plain FP64 block products */
#define LAMB_HIP_BF16_SPLIT 1    /* exponent-aligned BF16 slices, FP32 MFMA, FP64 across chains */
#define LAMB_HIP_SLICES(n) ((n) << 8) /* with BF16_SPLIT: 3..5 slices (0 = default 5) */

/* Statistics of one multiply (filled when a non-NULL pointer is given). */
typedef struct lamb_hip_stats {
  int64_t n_triples_kept;      /* atom-level (i,k,j) products kept by the prefilter */
  int64_t n_triples_filtered;  /* stored (i,k),(k,j) pairs dropped by the prefilter */
  int64_t n_c_blocks_struct;   /* C blocks in the structural pattern */
  int64_t n_c_blocks_kept;     /* C blocks stored after the post-filter */
  int64_t n_c_elems_kept;      /* FP64 values in the stored C blocks */
  int64_t n_tiles;             /* C tiles processed by the numeric kernel */
  int64_t n_tile_entries;      /* (tile, k) strip pairs processed */
  double flops;                /* sum over kept triples of 2*m_i*k_k*n_j */
  double ms_symbolic;          /* device time of the symbolic phase */
  double ms_numeric;           /* device time of the numeric kernel */
  double ms_index;             /* device time of the C index build */
  double ms_total;             /* wall time of the whole call */
} lamb_hip_stats;

/* --- context ------------------------------------------------------------ */
int32_t lamb_hip_init(int32_t device, lamb_hip_ctx** ctx);
int32_t lamb_hip_finalize(lamb_hip_ctx* ctx);
/* Last error of this context: level, type and a NUL-terminated message. */
int32_t lamb_hip_last_error(const lamb_hip_ctx* ctx, int32_t* level, int32_t* type,
                            char* msg, int32_t msg_len);

/* --- distribution: atoms, their subsets and positions --------------------
 * Replaces the atom set + set partition the reference builds in
 * lamb_set_partition_update (src/lamb_set_partition_methods.F90:69-144):
 * subset_map(g) is the 1-based subset of global atom g, order_map(g) its
 * 1-based local index inside that subset. coords (3*n_atoms, xyz per atom)
 * and box (3 edge lengths) are optional (NULL): when given, the device pool
 * orders atoms along a recursive coordinate bisection so that MFMA tiles
 * cover compact groups of atoms. */
int32_t lamb_hip_dist_create(lamb_hip_ctx* ctx, int64_t n_atoms, const int8_t* blk_size,
                             int32_t n_subsets, const int32_t* subset_map,
                             const int16_t* order_map, const float* coords,
                             const float* box, lamb_hip_dist** dist);
int32_t lamb_hip_dist_destroy(lamb_hip_dist* dist);

/* --- matrices ------------------------------------------------------------
 * A square block matrix over one distribution (rows and columns share the
 * atom set, as lamb_multiply requires of A, B, C,
 * src/lamb_multiplication.F90:136-146). */
int32_t lamb_hip_matrix_create(lamb_hip_dist* dist, lamb_hip_mat** mat);
int32_t lamb_hip_matrix_destroy(lamb_hip_mat* mat);
/* Device-side copy (lamb_matrix_copy, src/lamb_matrix_X_types.F90:144-187). */
int32_t lamb_hip_matrix_copy(const lamb_hip_mat* src, lamb_hip_mat* dst);

/* Stage one sector (subset_row, subset_col) in the reference's layout
 * (lamb_sector_X_type, src/lamb_sector_X_types.F90:79-87; index
 * src/lamb_index_types.F90:86-105): row_p int64[n_rows+1] 0-based
 * exclusive offsets, col_i int16[n_blks] 1-based subset-local column atom,
 * extents int64[2*n_blks] (lo, hi) 1-based inclusive into data, data the
 * sector's FP64 store of data_len values, each block column-major.
 * n_rows is the size of subset_row. Replaces the block insertion of
 * fill_local_sectors (src/lamb_matrix_operations.F90:294-375). */
int32_t lamb_hip_sector_upload(lamb_hip_mat* mat, int32_t subset_row, int32_t subset_col,
                               int64_t n_blks, const int64_t* row_p, const int16_t* col_i,
                               const int64_t* extents, const double* data, int64_t data_len);
/* Block-list form of internal_distribute (src/lamb_matrix_operations.F90:88-97)
 * for one data store: n blocks, subset_idx int32[2*n] (row, col subset),
 * local_idx int16[2*n] (1-based local row/col atom), extents int64[2*n]
 * 1-based inclusive into data. Blocks may arrive in any order. */
int32_t lamb_hip_distribute(lamb_hip_mat* mat, int64_t n, const int32_t* subset_idx,
                            const int16_t* local_idx, const int64_t* extents,
                            const double* data, int64_t data_len);
/* Build the device pool from everything staged since the last commit
 * (replaces lamb_index_convert_to_csr, src/lamb_index_methods.F90:224-246). */
int32_t lamb_hip_matrix_commit(lamb_hip_mat* mat);

/* Sizes of sector (subset_row, subset_col) as it would be downloaded. */
int32_t lamb_hip_sector_query(lamb_hip_mat* mat, int32_t subset_row, int32_t subset_col,
                              int64_t* n_blks, int64_t* data_len);
/* Fill caller-allocated arrays (sizes from lamb_hip_sector_query) with the
 * sector in the reference's layout; col_i ascending per row, blocks stored
 * back to back in row order. */
int32_t lamb_hip_sector_download(lamb_hip_mat* mat, int32_t subset_row, int32_t subset_col,
                                 int64_t* row_p, int16_t* col_i, int64_t* extents,
                                 double* data);
/* Sector wire format (lamb_goods_X_create, src/lamb_goods_X_types.F90:119-268;
 * lamb_goods_X_unpack, src/lamb_goods_X_methods.F90:166-218): an int64 index
 * array (12 header words of 1-based offsets and sizes, then row_p's first
 * n_rows entries, col_i as int16 packed little-endian into words, extents as
 * (lo, hi) pairs, one zero flag byte per block; each array max(1, ceil(bytes/8))
 * words) and the sector's FP64 data. goods_size gives both lengths,
 * goods_pack fills caller buffers of exactly those lengths, goods_unpack
 * validates the header against subset_row and stages the sector like
 * lamb_hip_sector_upload (lamb_hip_matrix_commit makes it live). */
int32_t lamb_hip_goods_size(lamb_hip_mat* mat, int32_t subset_row, int32_t subset_col,
                            int64_t* n_index_words, int64_t* n_data);
int32_t lamb_hip_goods_pack(lamb_hip_mat* mat, int32_t subset_row, int32_t subset_col,
                            int64_t* index, int64_t n_index_words, double* data,
                            int64_t n_data);
int32_t lamb_hip_goods_unpack(lamb_hip_mat* mat, int32_t subset_row, int32_t subset_col,
                              const int64_t* index, int64_t n_index_words,
                              const double* data, int64_t n_data);
/* Totals of the whole matrix. */
int32_t lamb_hip_matrix_info(const lamb_hip_mat* mat, int64_t* n_blks, int64_t* n_elems);

/* --- operations ----------------------------------------------------------
 * C = alpha * A*B + beta * C (lamb_multiply, src/lamb_multiplication.F90:116-204,
 * whose per-sector work sector_multiply :206-225 this implements).
 * Triple (i,k,j) is kept iff nsq(A_ik)*nsq(B_kj) >= eps_filter^2; C_ij is
 * stored iff nsq(C_ij) >= eps_filter^2, where nsq is the sum of squares of a
 * block's elements in column-major order, one fused multiply-add per step.
 * beta == 0 ignores C's old contents. C may not alias A or B. Products of
 * a distribution with an atom of more than 64 rows (the reference's int8
 * block size allows up to 126) are refused with LAMB_HIP_UNIMPLEMENTED:
 * the MFMA tiles hold at most 64 rows of one atom. Storage, exchange and the
 * group functions accept such atoms. */
int32_t lamb_hip_multiply(lamb_hip_ctx* ctx, const lamb_hip_mat* A, const lamb_hip_mat* B,
                          lamb_hip_mat* C, double alpha, double beta, double eps_filter,
                          int32_t flags, lamb_hip_stats* stats);

/* One sector product, C(m, n) += A(m, k) * B(k, n), for row subset m, inner
 * subset k and column subset n (sector_multiply, src/lamb_multiplication.F90:
 * 206-225, called from lamb_multiply's sector-pair loop :180-201, whose body
 * is a stub in the reference). Triples are prefiltered as in
 * lamb_hip_multiply (nsq(A_ik)*nsq(B_kj) >= eps_filter^2); C is an
 * accumulator between sector products, so no post-filter runs here: apply
 * lamb_hip_matrix_filter once every sector pair is in. C keeps every other
 * sector unchanged. */
int32_t lamb_hip_sector_multiply(lamb_hip_ctx* ctx, const lamb_hip_mat* A,
                                 const lamb_hip_mat* B, lamb_hip_mat* C, int32_t subset_m,
                                 int32_t subset_k, int32_t subset_n, double eps_filter,
                                 int32_t flags, lamb_hip_stats* stats);
/* The post-filter of lamb_hip_multiply on its own: drops the blocks of C
 * with nsq(C_ij) < eps_filter^2 (nsq as above); *n_dropped (may be NULL)
 * gets their count. */
int32_t lamb_hip_matrix_filter(lamb_hip_mat* C, double eps_filter, int64_t* n_dropped);

/* tr(A): sum of the diagonal elements of the diagonal blocks (FP64). */
int32_t lamb_hip_trace(const lamb_hip_mat* A, double* trace);

/* tr(A) without allocating: no counterpart in the reference. For callers
 * that cannot allocate (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). The per-atom partial sums go to a fixed
 * scratch array of the library, LAMB_HIP_NOALLOC_MAX_ATOMS doubles (8 MiB)
 * on every GPU the library runs on, held until the process ends; A's
 * distribution may hold at most that many atoms (more: LAMB_HIP_WRONG_ARGS).
 * dev_trace is the device address of one double. The call enqueues its
 * kernels on the context's stream and returns without waiting: dev_trace
 * holds, once the stream gets there, the value lamb_hip_trace returns, bit
 * for bit. Calls that overlap on one GPU (several contexts or host threads
 * at once) share the scratch array and race. */
#define LAMB_HIP_NOALLOC_MAX_ATOMS 1048576
int32_t lamb_hip_trace_noalloc(const lamb_hip_mat* A, double* dev_trace);

/* Kernel choice and autotuner state (no counterpart in the reference).
 * last_variant: the FP64 numeric kernel of the context's last multiply
 * (0 quadrant: per-wave 32 x 32 quadrants; 1 ksplit: whole-tile waves
 * splitting the K steps four ways; 2 half: per-wave 64 x 32 halves, two waves
 * splitting each half's K steps; 3 lanes: vector-ALU FMAs, per-wave 16 x 32
 * with the K columns of a step spread over the lane groups; 4 fine: the half
 * kernel on v_mfma_f64_4x4x4_4b, 16-row x 4-column live groups; lanes and
 * fine serve whole products whose packed panels stay below 4 GB, others run
 * half). The variant is a
 * function of the product (K-split for whole products of at least 1e7 flop
 * per tile, else half), so a product's bits do not depend on timings;
 * LAMB_HIP_F64_KERNEL=quadrant|half|ksplit|lanes|fine forces one (the variants add
 * K products in different orders: results then differ at rounding level). n_classes: workload classes
 * tuned so far. On the first product of a class the autotuner times the
 * tile orders (row bands per block of the order in which workgroups take
 * tiles: 4, 8, 16) as a persistent grid of the device's resident workgroups,
 * each looping over tiles with the XCDs sharing one window of the order;
 * then the fastest order as one workgroup per tile and as a persistent grid
 * whose XCDs each walk their own work-balanced range of the order. It keeps
 * the fastest of the five (the XCD ranges form, which moves fewer bytes,
 * wins ties within 0.5 %); order and launch form only map tiles to
 * workgroups, so every candidate writes the same bits. LAMB_HIP_TILE_ROWBLOCK=n and LAMB_HIP_PERSISTENT=0|1|2 force
 * them. */
int32_t lamb_hip_tuning_info(const lamb_hip_ctx* ctx, int32_t* last_variant,
                             int32_t* n_classes);
/* Row bands per block of the tile order the context's last multiply used (the
 * autotuner's choice for its workload class, or LAMB_HIP_TILE_ROWBLOCK). */
int32_t lamb_hip_tuning_order(const lamb_hip_ctx* ctx, int32_t* last_rowblock);
/* Tile-shape tuning (no counterpart in the reference). The per-wave tile of
 * the FP64 kernel (quadrant 32 x 32, half 64 x 32, K-split 64 x 64, lanes
 * 16 x 32, fine 64 x 32 on 4 x 4 x 4 MFMAs; see lamb_hip_tuning_info) is by default chosen by the rule, because the shapes
 * add K products in different orders. tune_shapes = 1: the first product of
 * each workload class (form, whether every entry is whole, the set of block
 * sizes of the distribution, i.e. the (m, n, k) shapes of its blocks, and
 * half-octave buckets of flops and of C blocks per tile) times the shapes
 * (lanes and fine only for whole products) and keeps the fastest for that class; results then depend, at
 * rounding level, on which shape won. tune_shapes = 0 (default): rule, or a
 * loaded table's entry. */
int32_t lamb_hip_tuning_mode(lamb_hip_ctx* ctx, int32_t tune_shapes);
/* The context's tuning table (tile shapes, tile orders, launch forms per
 * workload class) written to / read from a text file (no counterpart in the
 * reference). Loaded entries take effect for their classes as if tuned in
 * this context: a process that loads the table of an earlier run makes the
 * same choices, without tuning launches, whatever the mode. A file that is
 * not a table, or holds a malformed entry, is refused (LAMB_HIP_WRONG_ARGS)
 * and nothing is loaded; an unreadable or unwritable path gives
 * LAMB_HIP_EXTERNAL_ERROR. */
int32_t lamb_hip_tuning_save(const lamb_hip_ctx* ctx, const char* path);
int32_t lamb_hip_tuning_load(lamb_hip_ctx* ctx, const char* path);
/* Launch form of the context's last numeric kernel (no counterpart in the
 * reference): form = 0 for one workgroup per tile, 1 for a persistent grid
 * whose XCDs share one window of the tile order, 2 for a persistent grid
 * whose XCDs each walk their own range of it (work-balanced by K columns);
 * workgroups = the grid launched, tiles = the product's C tiles (-1 before
 * any multiply). The forms give the same bits. */
int32_t lamb_hip_tuning_launch(const lamb_hip_ctx* ctx, int32_t* form,
                               int32_t* workgroups, int32_t* tiles);

/* --- sector ownership (plexus) -------------------------------------------
 * The maps of lamb_plexus_calculate with force_n_procs
 * (src/lamb_plexus_operations.F90:87-185) for process my_proc (0-based) of
 * n_procs: region pools (fill_region_pools :198-320), the realm map that
 * gives each sector (region pair) its owner (fill_realm_map :378-449), the
 * process's catalog, orbit and sectors (fill_my_sectors :454-488) and its
 * multiply schedule (lamb_plexus_create_sched :492-571). Regions and subsets
 * are 1-based and one to one, as in the reference. n_regions = 0 takes the
 * reference's count for n_procs (int((sqrt(1 + 8 P) - 1) / 2 * 2)), else it
 * is forced (force_n_cubes). Host-side maps: no GPU call. Getters copy into
 * caller arrays, any of which may be NULL:
 *   realm_map [n_regions * n_regions], row-major (r - 1, s - 1);
 *   pool_sizes [n_regions]; pools [n_regions * n_procs], row r - 1 holding
 *   its sorted processes then -1; burdens [n_procs];
 *   catalog [catalog_size], orbit [orbit_size] (sorted);
 *   sectors [2 * n_sectors] as (row subset, col subset) pairs;
 *   sched [6 * n_sched]: era, source process, this process, operand
 *   (1 = A, 2 = B), row subset, col subset. */
typedef struct lamb_hip_plexus lamb_hip_plexus;
int32_t lamb_hip_plexus_create(int32_t n_procs, int32_t my_proc, int32_t n_regions,
                               lamb_hip_plexus** plexus);
int32_t lamb_hip_plexus_destroy(lamb_hip_plexus* plexus);
int32_t lamb_hip_plexus_info(const lamb_hip_plexus* plexus, int32_t* n_regions,
                             int32_t* catalog_size, int32_t* orbit_size, int32_t* n_sectors,
                             int64_t* n_sched);
int32_t lamb_hip_plexus_maps(const lamb_hip_plexus* plexus, int32_t* realm_map,
                             int32_t* pool_sizes, int32_t* pools, int32_t* burdens);
int32_t lamb_hip_plexus_lists(const lamb_hip_plexus* plexus, int32_t* catalog, int32_t* orbit,
                              int32_t* sectors);
int32_t lamb_hip_plexus_sched(const lamb_hip_plexus* plexus, int32_t* sched);

/* --- batches and purification chains --------------------------------------
 * None of these has a counterpart in the reference, whose multiply computes
 * nothing; they are the matrix functions a McWeeny purification chain needs
 * (P <- 3P^2 - 2P^3, started from H scaled into [0,1]). A batch of
 * independent systems is one distribution whose atoms carry a group (system)
 * id: matrices are block diagonal over groups, every multiply runs all
 * groups in one launch per kernel, and these functions report per group.
 * Results are FP64 and deterministic (per-atom partials summed in order). */
/* group_of_atom[g] in 0..n_groups-1 for each global atom g (default: all 0). */
int32_t lamb_hip_dist_set_groups(lamb_hip_dist* dist, int32_t n_groups,
                                 const int32_t* group_of_atom);
/* out[n_groups]: per-group trace, Frobenius norm, ||X - Y||_F (union pattern). */
int32_t lamb_hip_group_trace(const lamb_hip_mat* A, double* out);
int32_t lamb_hip_group_norm(const lamb_hip_mat* A, double* out);
int32_t lamb_hip_group_diff_norm(const lamb_hip_mat* X, const lamb_hip_mat* Y, double* out);
/* Per-group Gershgorin bounds of the spectrum: lo = min_r (H_rr - R_r),
 * hi = max_r (H_rr + R_r), R_r = sum over c != r of |H_rc|. Warning level
 * (precondition) when an atom has no stored diagonal block. */
int32_t lamb_hip_group_gershgorin(const lamb_hip_mat* A, double* lo, double* hi);
/* A <- a[g] * A + b[g] * I per group; every atom needs its diagonal block
 * when some b[g] != 0 (else failure, precondition). */
int32_t lamb_hip_group_scale_shift(lamb_hip_mat* A, const double* a, const double* b);
/* steps McWeeny iterations on P in place, every multiply filtered with
 * eps_filter and computed in the form flags selects. Step k writes
 * idem[k*n_groups + g] = ||P^2 - P||_F of its input and
 * traces[k*n_groups + g] = tr of its output (either may be NULL). stats
 * accumulates the 2*steps multiplies. */
int32_t lamb_hip_mcweeny(lamb_hip_ctx* ctx, lamb_hip_mat* P, int32_t steps, double eps_filter,
                         int32_t flags, double* traces, double* idem, lamb_hip_stats* stats);

/* McWeeny chains as HIP graphs. capture records `steps` iterations from P
 * (sizes the host reads are logged), grows the chain's buffers to their
 * maximum and captures the same chain on the context stream; P is not
 * modified. replay runs the graph on P_in (which must have the captured
 * pattern; its values may differ) and copies the result to P_out, with the
 * per-step traces / idempotency errors of lamb_hip_mcweeny and the graph's
 * device time in ms. A replay whose sizes differ from the recorded ones
 * (other filter decisions) fails with a precondition error. */
typedef struct lamb_hip_graph lamb_hip_graph;
int32_t lamb_hip_mcweeny_capture(lamb_hip_ctx* ctx, const lamb_hip_mat* P, int32_t steps,
                                 double eps_filter, int32_t flags, lamb_hip_graph** graph);
int32_t lamb_hip_graph_replay(lamb_hip_graph* graph, const lamb_hip_mat* P_in,
                              lamb_hip_mat* P_out, double* traces, double* idem, double* ms);
/* steps, number of recorded sizes, flops of one replay */
int32_t lamb_hip_graph_info(const lamb_hip_graph* graph, int32_t* steps, int64_t* sizes,
                            double* flops);
int32_t lamb_hip_graph_destroy(lamb_hip_graph* graph);

/* --- differentiation ------------------------------------------------------
 * No counterpart in the reference. Reverse-mode adjoint of C = A*B: with G
 * the gradient of a loss with respect to C (stored on C's pattern),
 *   dA = (G B^T) sampled on A's pattern,  dB = (A^T G) sampled on B's pattern,
 * each an SDDMM computed as one masked multiply by an explicit transpose, in
 * the form flags selects (FP64 or BF16 split), unfiltered. dA / dB (either
 * may be NULL) get exactly A's / B's pattern. */
int32_t lamb_hip_multiply_adjoint(lamb_hip_ctx* ctx, const lamb_hip_mat* A,
                                  const lamb_hip_mat* B, const lamb_hip_mat* G,
                                  lamb_hip_mat* dA, lamb_hip_mat* dB, int32_t flags,
                                  lamb_hip_stats* stats);
/* XT = X^T (distinct matrices of one distribution). */
int32_t lamb_hip_matrix_transpose(const lamb_hip_mat* X, lamb_hip_mat* XT);
/* dst = like's pattern holding src's blocks (zero blocks where src has none). */
int32_t lamb_hip_matrix_project(const lamb_hip_mat* src, const lamb_hip_mat* like,
                                lamb_hip_mat* dst);
/* Values in canonical order (stored blocks in CSR order of the internal atom
 * order, each column-major, back to back; n = n_elems of matrix_info) to or
 * from DEVICE memory of the library's device. */
int32_t lamb_hip_matrix_get_values(const lamb_hip_mat* X, double* dev_values, int64_t n);
int32_t lamb_hip_matrix_set_values(lamb_hip_mat* X, const double* dev_values, int64_t n);
/* Frees X's value pool, keeping its pattern (set_values brings it back). */
int32_t lamb_hip_matrix_release_values(lamb_hip_mat* X);
/* Global (0-based) row and column atom of each stored block, canonical order. */
int32_t lamb_hip_matrix_block_atoms(lamb_hip_mat* X, int32_t* rows, int32_t* cols);

/* --- atom partition on the device ---------------------------------------
 * The atom loop of lamb_set_partition_update (src/lamb_set_partition_methods.
 * F90:98-120) for atoms that move: subset_map(g) = 1 + cx + cy*nx +
 * cz*nx*ny with c_d = mod(int(x_d * mult_d), n_d), mult_d = n_d / box_d in
 * FP32 (lamb_partition_which_cube1, src/lamb_partition_methods.F90:85-107),
 * and order_map(g) the arrival order of g within its subset. coords
 * (FP32 xyz per atom, wrapped into the box) and the outputs are host arrays;
 * dims are the cube counts per axis (at most 4096 cubes). Atoms outside the
 * cubes are an error (wrong args), as is a subset beyond the int16 local
 * index (overflow). */
int32_t lamb_hip_partition_atoms(lamb_hip_ctx* ctx, int64_t n_atoms, const float* coords,
                                 const float* box, const int32_t* dims, int32_t* subset_map,
                                 int16_t* order_map);

/* --- DBCSR files ---------------------------------------------------------
 * The dump format that lamb_load_dbcsr reads (src/lamb_io_operations.F90:
 * 296-330; chunk offsets :557-597), in native byte order:
 *   sub-header1: version char[10], name char[80], matrix type char,
 *                int32 n_chunks, data_type, block rows, block columns,
 *                int32 row block sizes[block rows], column block sizes;
 *   one sub-header2 per chunk: int32 nblks, data area size (FP64 values);
 *   one sub-header3 per chunk: int32 row_p[block rows + 1] (0-based
 *                exclusive offsets), col_i[nblks] (1-based block column),
 *                blk_p[nblks] (1-based start in the chunk's data area);
 *   one FP64 data area per chunk.
 * The two readers need no device and no context: on failure they return a
 * status and, when msg is not NULL, a NUL-terminated message. */
typedef struct lamb_hip_dbcsr_info {
  char version[11];     /* NUL-terminated */
  char name[81];        /* NUL-terminated, trailing blanks kept */
  char matrix_type;     /* 'N' (no symmetry), 'S', 'A', 'H' or 'K' */
  int32_t n_chunks;     /* numnodes of the writer */
  int32_t data_type;
  int32_t n_block_rows, n_block_cols;
  int64_t n_blocks;     /* over all chunks */
  int64_t n_values;     /* sum of m*n over all blocks */
} lamb_hip_dbcsr_info;
/* Sub-headers 1 and 2 of every chunk (lamb_load_dbcsr :353-384 and
 * load_dbcsr_chunks :568-585, with the reference's checks: block sizes
 * below 127, every block inside its chunk's data area). row_blk_size /
 * col_blk_size (int32[n_block_rows] / [n_block_cols]) may be NULL. */
int32_t lamb_hip_dbcsr_read_info(const char* path, lamb_hip_dbcsr_info* info,
                                 int32_t* row_blk_size, int32_t* col_blk_size, char* msg,
                                 int32_t msg_len);
/* Every block of every chunk, in file order (chunk, then row, then the
 * chunk's col_i order): rows / cols int32[n_blocks] 1-based block row and
 * column; values FP64[n_values] the blocks back to back, each column-major
 * (the coordinate index load_dbcsr_chunks builds, :642-656). */
int32_t lamb_hip_dbcsr_read_blocks(const char* path, int64_t n_blocks, int32_t* rows,
                                   int32_t* cols, int64_t n_values, double* values, char* msg,
                                   int32_t msg_len);
/* load_dbcsr_chunks into a device matrix (:520-740): each chunk's blocks are
 * staged as one data store, like lamb_hip_distribute, with (subset, local)
 * taken from mat's distribution, and the matrix is committed. The file must
 * be of type 'N' (the reference's load rejects the symmetric types as
 * unimplemented, :409-412) and square over the distribution's atoms, with
 * its block sizes. */
int32_t lamb_hip_dbcsr_load(lamb_hip_mat* mat, const char* path);
/* lamb_multiply_test (src/lamb_multiplication.F90:85-114): the DBCSR dump at
 * path loaded as A (lamb_load_dbcsr), copied to B and C (lamb_matrix_copy),
 * then C = A*B with eps_filter, FP64; stats may be NULL. The distribution
 * is the file's atoms in file order, in subsets of at most 32,767 atoms.
 * All three matrices are released before the call returns. */
int32_t lamb_hip_multiply_test(lamb_hip_ctx* ctx, const char* path, double eps_filter,
                               lamb_hip_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* LAMB_HIP_H */
