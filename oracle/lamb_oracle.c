/*
 * CPU FP64 oracle for the sector block-sparse multiply C = alpha*A*B + beta*C.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py load this library, and only as the checker.
 * The product path never calls it.
 *
 * The reference's per-sector multiply is an empty stub
 * (src/lamb_multiplication.F90:206-225): it defines which sector pairs
 * multiply (the loop at :180-201) but no arithmetic. This file restates
 * the semantics the product implements, in plain C:
 *
 *  - Block (i,j) is an m_i x n_j FP64 array, column-major, the DBCSR
 *    convention of the reference's loader (src/lamb_io_operations.F90:680-682).
 *  - Matrices are block-CSR: row_p int64[n+1] with 0-based exclusive offsets
 *    (the reference's index_compress/count_to_offset,
 *    src/lamb_index_methods.F90:248-274), col int32 ascending per row (the
 *    binary search precondition of get_blk_n, :73-107), off int64 element
 *    offset of each block in data.
 *  - nsq(X) = sum of squares of X's elements in column-major order, each step
 *    one fused multiply-add: s = fma(x, x, s), starting from s = +0.
 *  - Prefilter: triple (i,k,j) with A_ik and B_kj stored is kept iff
 *    nsq(A_ik) * nsq(B_kj) >= eps * eps (each product rounded once).
 *  - Numeric: acc_ij = sum over kept k ascending of A_ik * B_kj; within one
 *    block product the inner index kk runs ascending and every step is
 *    acc = fma(a, b, acc).
 *  - C_ij = alpha * acc_ij + beta * Cold_ij (Cold only read when beta != 0),
 *    over the union of the kept-triple pattern and (beta != 0) Cold's pattern.
 *  - Post-filter: C_ij is stored iff nsq(C_ij) >= eps * eps.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  int32_t n_atoms;
  const int32_t* blk;    /* block size per atom (rows == cols) */
  const int64_t* row_p;  /* n_atoms + 1 */
  const int32_t* col;    /* nblk */
  const int64_t* off;    /* nblk */
  const double* data;
} orc_mat;

typedef struct {
  int64_t nblk;
  int64_t ndata;
  int64_t* row_p;  /* n_atoms + 1 */
  int32_t* col;
  int64_t* off;
  double* data;
  double* nsq;
  /* statistics */
  int64_t n_triples_kept;
  int64_t n_triples_filtered;
  int64_t n_struct_blocks;
  double flops; /* sum over kept triples of 2 m_i k_k n_j */
} orc_result;

double orc_nsq1(const double* x, int64_t len) {
  double s = 0.0;
  for (int64_t e = 0; e < len; ++e) s = fma(x[e], x[e], s);
  return s;
}

/* nsq of every stored block, in storage order. */
void orc_block_nsq(const orc_mat* A, double* out) {
  const int32_t n = A->n_atoms;
#pragma omp parallel for schedule(dynamic, 16)
  for (int32_t i = 0; i < n; ++i) {
    for (int64_t b = A->row_p[i]; b < A->row_p[i + 1]; ++b) {
      const int64_t len = (int64_t)A->blk[i] * A->blk[A->col[b]];
      out[b] = orc_nsq1(A->data + A->off[b], len);
    }
  }
}

/* acc (m x n) += X (m x kk) * Y (kk x n), all column-major. */
static void gemm_acc(double* acc, const double* X, const double* Y, int m, int kk, int n) {
  for (int c = 0; c < n; ++c) {
    double* a = acc + (int64_t)c * m;
    for (int q = 0; q < kk; ++q) {
      const double y = Y[q + (int64_t)c * kk];
      const double* x = X + (int64_t)q * m;
      for (int r = 0; r < m; ++r) a[r] = fma(x[r], y, a[r]);
    }
  }
}

typedef struct {
  int32_t ncnt;
  int32_t* cols;
  double* vals;   /* concatenated kept blocks */
  int64_t nval;
  double* nsq;
} row_out;

static int cmp_i32(const void* a, const void* b) {
  const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* Returns 0 on success, nonzero on allocation failure. */
int orc_multiply(const orc_mat* A, const orc_mat* B, const orc_mat* Cold,
                 double alpha, double beta, double eps, orc_result* R) {
  const int32_t n = A->n_atoms;
  const double eps2 = eps * eps;
  const int use_old = (beta != 0.0) && Cold != NULL;
  memset(R, 0, sizeof(*R));
  int64_t nA = A->row_p[n], nB = B->row_p[n];
  double* nsqA = (double*)malloc(sizeof(double) * (size_t)(nA ? nA : 1));
  double* nsqB = (double*)malloc(sizeof(double) * (size_t)(nB ? nB : 1));
  row_out* rows = (row_out*)calloc((size_t)n, sizeof(row_out));
  if (!nsqA || !nsqB || !rows) return 1;
  orc_block_nsq(A, nsqA);
  orc_block_nsq(B, nsqB);
  int64_t kept = 0, filtered = 0, nstruct = 0;
  double flops = 0.0;
  int fail = 0;
  int32_t maxblk = 0;
  for (int32_t i = 0; i < n; ++i) if (A->blk[i] > maxblk) maxblk = A->blk[i];

#pragma omp parallel reduction(+ : kept, filtered, nstruct, flops) reduction(| : fail)
  {
    int32_t* slot = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* touched = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    double* ws = (double*)malloc(sizeof(double) * (size_t)n * maxblk * maxblk);
    if (!slot || !touched || !ws) fail |= 1;
    if (slot) for (int32_t j = 0; j < n; ++j) slot[j] = -1;
#pragma omp for schedule(dynamic, 4)
    for (int32_t i = 0; i < n; ++i) {
      if (fail) continue;
      const int m = A->blk[i];
      int32_t nt = 0;
      for (int64_t ab = A->row_p[i]; ab < A->row_p[i + 1]; ++ab) {
        const int32_t k = A->col[ab];
        const int kk = A->blk[k];
        const double* a = A->data + A->off[ab];
        for (int64_t bb = B->row_p[k]; bb < B->row_p[k + 1]; ++bb) {
          const int32_t j = B->col[bb];
          if (nsqA[ab] * nsqB[bb] >= eps2) {
            const int nn = B->blk[j];
            if (slot[j] < 0) {
              slot[j] = j;
              touched[nt++] = j;
              memset(ws + (int64_t)j * maxblk * maxblk, 0, sizeof(double) * (size_t)m * nn);
            }
            gemm_acc(ws + (int64_t)j * maxblk * maxblk, a, B->data + B->off[bb], m, kk, nn);
            kept += 1;
            flops += 2.0 * m * kk * nn;
          } else {
            filtered += 1;
          }
        }
      }
      if (use_old) {
        for (int64_t cb = Cold->row_p[i]; cb < Cold->row_p[i + 1]; ++cb) {
          const int32_t j = Cold->col[cb];
          if (slot[j] < 0) {
            slot[j] = j;
            touched[nt++] = j;
            memset(ws + (int64_t)j * maxblk * maxblk, 0, sizeof(double) * (size_t)m * B->blk[j]);
          }
        }
      }
      qsort(touched, (size_t)nt, sizeof(int32_t), cmp_i32);
      nstruct += nt;
      /* combine with Cold, post-filter */
      row_out* ro = &rows[i];
      int64_t tot = 0;
      for (int32_t t = 0; t < nt; ++t) tot += (int64_t)m * B->blk[touched[t]];
      ro->cols = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nt ? nt : 1));
      ro->nsq = (double*)malloc(sizeof(double) * (size_t)(nt ? nt : 1));
      ro->vals = (double*)malloc(sizeof(double) * (size_t)(tot ? tot : 1));
      if (!ro->cols || !ro->nsq || !ro->vals) { fail |= 1; continue; }
      int64_t cold_b = use_old ? Cold->row_p[i] : 0;
      const int64_t cold_e = use_old ? Cold->row_p[i + 1] : 0;
      for (int32_t t = 0; t < nt; ++t) {
        const int32_t j = touched[t];
        const int nn = B->blk[j];
        double* w = ws + (int64_t)j * maxblk * maxblk;
        const double* old = NULL;
        if (use_old) {
          while (cold_b < cold_e && Cold->col[cold_b] < j) ++cold_b;
          if (cold_b < cold_e && Cold->col[cold_b] == j) old = Cold->data + Cold->off[cold_b];
        }
        const int64_t len = (int64_t)m * nn;
        for (int64_t e = 0; e < len; ++e) {
          double v = alpha * w[e];
          if (old) v = v + beta * old[e];
          w[e] = v;
        }
        const double s = orc_nsq1(w, len);
        if (s >= eps2) {
          memcpy(ro->vals + ro->nval, w, sizeof(double) * (size_t)len);
          ro->cols[ro->ncnt] = j;
          ro->nsq[ro->ncnt] = s;
          ro->ncnt += 1;
          ro->nval += len;
        }
        slot[j] = -1;
      }
    }
    free(slot);
    free(touched);
    free(ws);
  }
  free(nsqA);
  free(nsqB);
  if (fail) {
    for (int32_t i = 0; i < n; ++i) { free(rows[i].cols); free(rows[i].vals); free(rows[i].nsq); }
    free(rows);
    return 2;
  }
  int64_t nblk = 0, ndata = 0;
  for (int32_t i = 0; i < n; ++i) { nblk += rows[i].ncnt; ndata += rows[i].nval; }
  R->row_p = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
  R->col = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nblk ? nblk : 1));
  R->off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nblk ? nblk : 1));
  R->nsq = (double*)malloc(sizeof(double) * (size_t)(nblk ? nblk : 1));
  R->data = (double*)malloc(sizeof(double) * (size_t)(ndata ? ndata : 1));
  if (!R->row_p || !R->col || !R->off || !R->nsq || !R->data) return 3;
  int64_t b = 0, d = 0;
  R->row_p[0] = 0;
  for (int32_t i = 0; i < n; ++i) {
    const int m = A->blk[i];
    memcpy(R->data + d, rows[i].vals, sizeof(double) * (size_t)rows[i].nval);
    for (int32_t t = 0; t < rows[i].ncnt; ++t) {
      R->col[b] = rows[i].cols[t];
      R->nsq[b] = rows[i].nsq[t];
      R->off[b] = d;
      d += (int64_t)m * B->blk[rows[i].cols[t]];
      ++b;
    }
    R->row_p[i + 1] = b;
    free(rows[i].cols);
    free(rows[i].vals);
    free(rows[i].nsq);
  }
  free(rows);
  R->nblk = nblk;
  R->ndata = ndata;
  R->n_triples_kept = kept;
  R->n_triples_filtered = filtered;
  R->n_struct_blocks = nstruct;
  R->flops = flops;
  return 0;
}

void orc_result_free(orc_result* R) {
  free(R->row_p);
  free(R->col);
  free(R->off);
  free(R->data);
  free(R->nsq);
  memset(R, 0, sizeof(*R));
}

/* Sum of the diagonal elements of the diagonal blocks, atoms ascending. */
double orc_trace(const orc_mat* A) {
  double t = 0.0;
  for (int32_t i = 0; i < A->n_atoms; ++i) {
    for (int64_t b = A->row_p[i]; b < A->row_p[i + 1]; ++b) {
      if (A->col[b] != i) continue;
      const int m = A->blk[i];
      const double* x = A->data + A->off[b];
      for (int r = 0; r < m; ++r) t += x[r + (int64_t)r * m];
    }
  }
  return t;
}

/* Gershgorin bounds over scalar rows: lo = min(a_rr - R_r), hi = max(a_rr + R_r),
 * R_r = sum over c != r of |a_rc|. */
void orc_gershgorin(const orc_mat* A, double* lo, double* hi) {
  int64_t nrows = 0;
  for (int32_t i = 0; i < A->n_atoms; ++i) nrows += A->blk[i];
  double* rad = (double*)calloc((size_t)nrows, sizeof(double));
  double* cen = (double*)calloc((size_t)nrows, sizeof(double));
  int64_t r0 = 0;
  for (int32_t i = 0; i < A->n_atoms; ++i) {
    const int m = A->blk[i];
    for (int64_t b = A->row_p[i]; b < A->row_p[i + 1]; ++b) {
      const int32_t j = A->col[b];
      const int nn = A->blk[j];
      const double* x = A->data + A->off[b];
      for (int c = 0; c < nn; ++c)
        for (int r = 0; r < m; ++r) {
          const double v = x[r + (int64_t)c * m];
          if (j == i && r == c) cen[r0 + r] = v;
          else rad[r0 + r] += fabs(v);
        }
    }
    r0 += m;
  }
  double l = INFINITY, h = -INFINITY;
  for (int64_t r = 0; r < nrows; ++r) {
    if (cen[r] - rad[r] < l) l = cen[r] - rad[r];
    if (cen[r] + rad[r] > h) h = cen[r] + rad[r];
  }
  *lo = l;
  *hi = h;
  free(rad);
  free(cen);
}

/*
 * Adjoint of C = A*B (no filtering) for upstream G on C's pattern:
 *   dA_ik = sum_j G_ij * B_kj^T   on the stored pattern of A
 *   dB_kj = sum_i A_ik^T * G_ij   on the stored pattern of B
 * Outputs are laid out exactly like A->data / B->data (same offsets).
 * Sums run over j (resp. i) ascending, fma per step.
 */
void orc_adjoint(const orc_mat* A, const orc_mat* B, const orc_mat* G,
                 double* dA, double* dB) {
  const int32_t n = A->n_atoms;
#pragma omp parallel for schedule(dynamic, 4)
  for (int32_t i = 0; i < n; ++i) {
    const int m = A->blk[i];
    for (int64_t ab = A->row_p[i]; ab < A->row_p[i + 1]; ++ab) {
      const int32_t k = A->col[ab];
      const int kk = A->blk[k];
      double* out = dA + A->off[ab];
      for (int64_t e = 0; e < (int64_t)m * kk; ++e) out[e] = 0.0;
      /* walk G row i and B row k together (both ascending in j) */
      int64_t g = G->row_p[i], ge = G->row_p[i + 1];
      int64_t bb = B->row_p[k], be = B->row_p[k + 1];
      while (g < ge && bb < be) {
        const int32_t jg = G->col[g], jb = B->col[bb];
        if (jg < jb) { ++g; continue; }
        if (jb < jg) { ++bb; continue; }
        const int nn = B->blk[jg];
        const double* gx = G->data + G->off[g];   /* m x nn */
        const double* bx = B->data + B->off[bb];  /* kk x nn */
        for (int q = 0; q < kk; ++q)
          for (int c = 0; c < nn; ++c) {
            const double y = bx[q + (int64_t)c * kk];
            for (int r = 0; r < m; ++r)
              out[r + (int64_t)q * m] = fma(gx[r + (int64_t)c * m], y, out[r + (int64_t)q * m]);
          }
        ++g;
        ++bb;
      }
    }
  }
  /* dB_kj = sum_i A_ik^T G_ij over i ascending: column k of A comes from a
   * CSC copy of A's index; each k is owned by one thread. */
  int64_t nA = A->row_p[n];
  int64_t* cp = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  int32_t* ci = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nA ? nA : 1));
  int64_t* cb = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nA ? nA : 1));
  int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
  for (int64_t b = 0; b < nA; ++b) cp[A->col[b] + 1] += 1;
  for (int32_t k = 0; k < n; ++k) cp[k + 1] += cp[k];
  for (int32_t k = 0; k < n; ++k) fill[k] = cp[k];
  for (int32_t i = 0; i < n; ++i)
    for (int64_t b = A->row_p[i]; b < A->row_p[i + 1]; ++b) {
      const int32_t k = A->col[b];
      ci[fill[k]] = i;
      cb[fill[k]] = b;
      fill[k] += 1;
    }
#pragma omp parallel for schedule(dynamic, 4)
  for (int32_t k = 0; k < n; ++k) {
    const int kk = B->blk[k];
    for (int64_t bb = B->row_p[k]; bb < B->row_p[k + 1]; ++bb) {
      const int nn = B->blk[B->col[bb]];
      double* out = dB + B->off[bb];
      for (int64_t e = 0; e < (int64_t)kk * nn; ++e) out[e] = 0.0;
    }
    for (int64_t t = cp[k]; t < cp[k + 1]; ++t) {
      const int32_t i = ci[t];
      const int m = A->blk[i];
      const double* ax = A->data + A->off[cb[t]];  /* m x kk */
      int64_t g = G->row_p[i], ge = G->row_p[i + 1];
      int64_t bb = B->row_p[k], be = B->row_p[k + 1];
      while (g < ge && bb < be) {
        const int32_t jg = G->col[g], jb = B->col[bb];
        if (jg < jb) { ++g; continue; }
        if (jb < jg) { ++bb; continue; }
        const int nn = B->blk[jg];
        const double* gx = G->data + G->off[g];  /* m x nn */
        double* out = dB + B->off[bb];           /* kk x nn */
        for (int c = 0; c < nn; ++c)
          for (int q = 0; q < kk; ++q) {
            double s = out[q + (int64_t)c * kk];
            for (int r = 0; r < m; ++r) s = fma(ax[r + (int64_t)q * m], gx[r + (int64_t)c * m], s);
            out[q + (int64_t)c * kk] = s;
          }
        ++g;
        ++bb;
      }
    }
  }
  free(cp);
  free(ci);
  free(cb);
  free(fill);
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}
