"""Writes tests/golden/integer_product.npz: a tiny block-sparse product with
integer-valued blocks, computed with Python integers (exact), plus the
kept-triple count under the norm prefilter (eps = 4.5) and the post-filter.

Run from the repository root: python tests/golden/make_integer_product.py
"""
import random

import numpy as np

rng = random.Random(20261203)
n = 10
blk = [rng.choice([1, 2, 3, 4]) for _ in range(n)]


def rand_pattern(p):
    return sorted({(i, j) for i in range(n) for j in range(n) if i == j or rng.random() < p})


def rand_block(m, k):
    return [[rng.randint(-3, 3) for _ in range(k)] for _ in range(m)]


def csr(pattern, blocks):
    row_p = [0] * (n + 1)
    col, off, data = [], [], []
    o = 0
    for (i, j) in pattern:
        row_p[i + 1] += 1
        col.append(j)
        off.append(o)
        b = blocks[(i, j)]
        for c in range(blk[j]):          # column-major
            for r in range(blk[i]):
                data.append(b[r][c])
        o += blk[i] * blk[j]
    for i in range(n):
        row_p[i + 1] += row_p[i]
    return row_p, col, off, data


pa = rand_pattern(0.35)
pb = rand_pattern(0.35)
A = {(i, j): rand_block(blk[i], blk[j]) for (i, j) in pa}
B = {(i, j): rand_block(blk[i], blk[j]) for (i, j) in pb}
eps2 = 4.5 * 4.5


def nsq(b):
    return sum(v * v for row in b for v in row)


acc, kept = {}, 0
for (i, k) in pa:
    for (k2, j) in pb:
        if k2 != k or nsq(A[(i, k)]) * nsq(B[(k, j)]) < eps2:
            continue
        kept += 1
        c = acc.setdefault((i, j), [[0] * blk[j] for _ in range(blk[i])])
        for r in range(blk[i]):
            for q in range(blk[j]):
                c[r][q] += sum(A[(i, k)][r][t] * B[(k, j)][t][q] for t in range(blk[k]))
C = {key: v for key, v in acc.items() if nsq(v) >= eps2}
pc = sorted(C)
a_row_p, a_col, a_off, a_data = csr(pa, A)
b_row_p, b_col, b_off, b_data = csr(pb, B)
c_row_p, c_col, c_off, c_data = csr(pc, C)
np.savez("tests/golden/integer_product.npz",
         blk=np.array(blk, np.int32), eps=np.float64(4.5), kept=np.int64(kept),
         a_row_p=np.array(a_row_p, np.int64), a_col=np.array(a_col, np.int32),
         a_off=np.array(a_off, np.int64), a_data=np.array(a_data, np.float64),
         b_row_p=np.array(b_row_p, np.int64), b_col=np.array(b_col, np.int32),
         b_off=np.array(b_off, np.int64), b_data=np.array(b_data, np.float64),
         c_row_p=np.array(c_row_p, np.int64), c_col=np.array(c_col, np.int32),
         c_off=np.array(c_off, np.int64), c_data=np.array(c_data, np.float64))
print("kept", kept, "C blocks", len(pc), "of", len(acc))
