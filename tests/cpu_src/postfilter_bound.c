// The post-filter shortcut of the tile epilogue: when |sum - eps2| > margin,
// the contract's sequential FMA chain is on the same side of eps2 as the
// epilogue's cheap sum. That sum is formed as on the GPU: the 16 lanes of a
// wave row, lane l one FMA per element e = l, l + 16, ..., then the DPP row
// prefix sums by row_shr 1, 2, 4, 8, total in lane 15. Random blocks up to
// 64 x 64, eps2 placed 1e-17..1e-9 (relative) from the sum, or equal to the
// chain; exit 1 on any disagreement.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
static double rnd(void) { return (double)rand() / RAND_MAX; }
int main(void) {
  srand(12345);
  long trials = 0, certain = 0, wrong = 0, near = 0;
  for (int it = 0; it < 100000; ++it) {
    const int m = 1 + rand() % 64, n = 1 + rand() % 64, len = m * n;
    static double x[4096];
    const double scale = ldexp(1.0, (rand() % 200) - 100);
    for (int e = 0; e < len; ++e) x[e] = (2 * rnd() - 1) * scale * (rnd() < 0.1 ? 1e-8 : 1.0);
    double seq = 0.0;
    for (int e = 0; e < len; ++e) seq = fma(x[e], x[e], seq);
    double v[16];
    for (int l = 0; l < 16; ++l) {
      v[l] = 0.0;
      for (int e = l; e < len; e += 16) v[l] = fma(x[e], x[e], v[l]);
    }
    for (int k = 1; k <= 8; k <<= 1) {  // row_shr k: lane l reads l - k, else 0
      double nv[16];
      for (int l = 0; l < 16; ++l) nv[l] = v[l] + (l >= k ? v[l - k] : 0.0);
      for (int l = 0; l < 16; ++l) v[l] = nv[l];
    }
    const double sum = v[15];
    const double rel = ldexp(1.0, -(rand() % 30) - 27) * (rand() % 2 ? 1 : -1);
    const double eps2 = it % 4 == 0 ? seq : sum * (1.0 + rel);
    const double margin = sum * ((double)(len + (len + 15) / 16 + 8) * 0x1p-52);
    const int cert = sum > 0x1p-930 && sum <= 0x1p+1000 && fabs(sum - eps2) > margin;
    ++trials;
    if (cert) {
      ++certain;
      if ((sum >= eps2) != (seq >= eps2)) ++wrong;
    } else {
      ++near;
    }
  }
  printf("trials %ld certain %ld fallback %ld wrong %ld\n", trials, certain, near, wrong);
  return wrong != 0;
}
