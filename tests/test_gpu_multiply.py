"""GPU parity: the HIP multiply through the C-ABI against the CPU oracle.

Bar: identical stored block pattern (row_p, col) and kept-triple counts;
values within 1e-12 relative Frobenius error (FP64, different but fixed
summation order); integer-valued and identity cases bit-exact.
"""
import numpy as np
import pytest

import oracle
from __graft_entry__ import compare
from lamb_amd import device as D
from lamb_amd import systems as S
from lamb_amd.matrix import Distribution, LambMatrix

pytestmark = pytest.mark.gpu
TOL = 1e-12


def gpu_multiply(ctx, atoms, a, b, *, eps=0.0, alpha=1.0, beta=0.0, c_old=None, spatial=True,
                 n_procs=1, flags=None):
    dist = Distribution(atoms.blk, atoms.pos, atoms.box, n_procs=n_procs)
    dd = D.DeviceDistribution(ctx, dist, spatial=spatial)
    da = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, a))
    db = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, b))
    dc = D.DeviceMatrix(dd)
    if c_old is not None:
        dc.upload(LambMatrix.from_global(dist, c_old))
    st = D.multiply(da, db, dc, alpha=alpha, beta=beta, eps_filter=eps,
                    flags=D.FP64 if flags is None else flags)
    got = dc.download().to_global()
    for h in (dc, db, da, dd):
        h.close()
    return got, st


def check(ctx, atoms, a, b, *, eps=0.0, alpha=1.0, beta=0.0, c_old=None, spatial=True,
          exact=False, n_procs=1):
    print(f"case: atoms={atoms.n_atoms} blocks={a.n_blocks} eps={eps} alpha={alpha} beta={beta}"
          f" n_procs={n_procs}", flush=True)
    got, st = gpu_multiply(ctx, atoms, a, b, eps=eps, alpha=alpha, beta=beta, c_old=c_old,
                           spatial=spatial, n_procs=n_procs)
    ref, rst = oracle.multiply(a, b, eps=eps, alpha=alpha, beta=beta, c_old=c_old)
    same, err = compare(got, ref)
    assert same, "stored block pattern differs from the oracle"
    if exact:
        assert np.array_equal(got.data, ref.data)
    else:
        assert err <= TOL, err
    assert st["n_triples_kept"] == rst["n_triples_kept"]
    assert st["n_triples_filtered"] == rst["n_triples_filtered"]
    assert st["flops"] == rst["flops"]
    assert st["n_c_blocks_kept"] == ref.n_blocks
    return st, rst


def test_tiny_mixed(ctx):
    atoms, a = S.small_system(20, 1, sizes=(5, 13))
    b = S.decay_matrix(atoms, 6.0, 0.5, 2)
    check(ctx, atoms, a, b)


def test_water_eps(ctx):
    atoms = S.water_system(96, seed=11)
    a = S.decay_matrix(atoms, 8.0, 0.4, 12)
    b = S.decay_matrix(atoms, 8.0, 0.4, 13)
    check(ctx, atoms, a, b, eps=1e-7)


@pytest.mark.parametrize("n_procs", [2, 8])
def test_several_subsets(ctx, n_procs):
    """A distribution with several spatial subsets (one process owning every
    sector), as used for batches beyond the int16 local index of one subset."""
    atoms = S.water_system(300, seed=41)
    a = S.decay_matrix(atoms, 7.0, 0.4, 42)
    b = S.decay_matrix(atoms, 7.0, 0.4, 43)
    assert Distribution(atoms.blk, atoms.pos, atoms.box, n_procs=n_procs).n_subsets > 1
    check(ctx, atoms, a, b, eps=1e-6, n_procs=n_procs)


def test_three_kinds(ctx):
    atoms = S.water_system(300, seed=21, heavy_fraction=0.1)
    assert set(np.unique(atoms.blk)) == {5, 13, 23}
    a = S.decay_matrix(atoms, 8.0, 0.4, 22)
    b = S.decay_matrix(atoms, 8.0, 0.4, 23)
    check(ctx, atoms, a, b, eps=1e-7)


def test_config1_symmetric(ctx):
    atoms, a, b, eps = S.config1()
    st, rst = check(ctx, atoms, a, b, eps=eps)
    assert st["n_triples_filtered"] == 0


def test_filters_bind(ctx):
    atoms, a = S.small_system(60, 31, cutoff=9.0, decay=1.5, sizes=(1, 5, 13))
    b = S.decay_matrix(atoms, 9.0, 1.3, 32)
    st, rst = check(ctx, atoms, a, b, eps=2e-3)
    assert rst["n_triples_filtered"] > 0
    assert rst["n_struct_blocks"] > st["n_c_blocks_kept"]  # the post-filter dropped blocks


def test_alpha_beta(ctx):
    atoms = S.water_system(150, seed=41)
    a = S.decay_matrix(atoms, 6.0, 0.4, 42)
    b = S.decay_matrix(atoms, 6.0, 0.4, 43)
    old = S.decay_matrix(atoms, 9.0, 0.4, 44)
    check(ctx, atoms, a, b, eps=1e-9, alpha=-2.0, beta=3.0, c_old=old)


def test_arrival_order_without_coordinates(ctx):
    atoms = S.water_system(120, seed=51)
    a = S.decay_matrix(atoms, 8.0, 0.4, 52)
    b = S.decay_matrix(atoms, 8.0, 0.4, 53)
    check(ctx, atoms, a, b, eps=1e-7, spatial=False)


def test_identity_is_exact(ctx):
    atoms, a = S.small_system(30, 61, sizes=(5, 13, 23))
    n = a.n_atoms
    off = np.concatenate([[0], np.cumsum(a.blk.astype(np.int64) ** 2)[:-1]])
    data = np.concatenate([np.eye(int(m)).reshape(-1, order="F") for m in a.blk])
    eye = S.BlockCSR(a.blk.copy(), np.arange(n + 1, dtype=np.int64),
                     np.arange(n, dtype=np.int32), off.astype(np.int64), data)
    check(ctx, atoms, eye, a, exact=True)
    check(ctx, atoms, a, eye, exact=True)


def _lane_butterfly_nsq(x):
    """Sum of squares of a block of at most 64 values in the order of a
    lane-strided partial per lane followed by an xor-butterfly wave sum (the
    order an epilogue with one wave per block would take)."""
    v = np.zeros(64)
    v[:len(x)] = x * x  # fma(x, x, 0) rounds x * x once
    idx = np.arange(64)
    for o in (32, 16, 8, 4, 2, 1):
        v = v + v[idx ^ o]
    return float(v[0])


def _eps_between(lo, hi):
    """An eps with lo < fl(eps * eps) <= hi, or None."""
    e = float(np.sqrt(hi))
    for _ in range(64):
        e = float(np.nextafter(e, np.inf))
    for _ in range(128):
        if lo < e * e <= hi:
            return e
        e = float(np.nextafter(e, 0.0))
    return None


@pytest.mark.parametrize("direction", ["kept_by_sequential", "dropped_by_sequential"])
def test_post_filter_norm_order_near_threshold(ctx, direction):
    """C = I * B is exact, so only the summation order of nu^2(C_ij) decides a
    block whose nu^2 lies within an ulp of eps^2. eps is chosen so that
    fl(eps * eps) lies between the contract's sequential column-major FMA chain
    and the lane-strided butterfly order: the GPU pattern must be the
    oracle's (sequential) one, which the other order would not give."""
    atoms, b = S.small_system(40, 7, sizes=(5,), cutoff=7.0, decay=0.3)
    n = b.n_atoms
    eye = S.BlockCSR(b.blk.copy(), np.arange(n + 1, dtype=np.int64),
                     np.arange(n, dtype=np.int32), (25 * np.arange(n)).astype(np.int64),
                     np.concatenate([np.eye(5).reshape(-1, order="F")] * n))
    pick = None
    for t in range(b.n_blocks):
        x = b.data[b.blk_off[t]:b.blk_off[t] + 25]
        seq, alt = oracle.nsq1(x), _lane_butterfly_nsq(x)
        if direction == "kept_by_sequential" and alt < seq:
            eps = _eps_between(alt, seq)
        elif direction == "dropped_by_sequential" and seq < alt:
            eps = _eps_between(seq, alt)
        else:
            continue
        if eps is not None:
            pick = (t, seq, alt, eps)
            break
    assert pick is not None, "no block with distinguishable summation orders"
    t, seq, alt, eps = pick
    print(f"case: block {t} seq={seq!r} alt={alt!r} eps^2={eps * eps!r}", flush=True)
    st, rst = check(ctx, atoms, eye, b, eps=eps, exact=True)
    ref, _ = oracle.multiply(eye, b, eps=eps)
    i = int(np.searchsorted(b.row_p, t, side="right") - 1)
    j = int(b.col[t])
    cols = ref.col[ref.row_p[i]:ref.row_p[i + 1]]
    assert (j in cols) == (direction == "kept_by_sequential")


def test_integer_golden_is_exact(ctx):
    fx = np.load("tests/golden/integer_product.npz")
    blk = fx["blk"]
    rng = np.random.default_rng(0)
    atoms = S.AtomBox(rng.uniform(0, 10, (len(blk), 3)).astype(np.float32), blk, 10.0)
    a = S.BlockCSR(blk, fx["a_row_p"], fx["a_col"], fx["a_off"], fx["a_data"])
    b = S.BlockCSR(blk, fx["b_row_p"], fx["b_col"], fx["b_off"], fx["b_data"])
    got, st = gpu_multiply(ctx, atoms, a, b, eps=float(fx["eps"]))
    assert np.array_equal(got.row_p, fx["c_row_p"]) and np.array_equal(got.col, fx["c_col"])
    assert np.array_equal(got.data, fx["c_data"])
    assert st["n_triples_kept"] == int(fx["kept"])


def test_empty_operand(ctx):
    atoms, a = S.small_system(12, 71)
    empty = S.BlockCSR(a.blk.copy(), np.zeros(a.n_atoms + 1, np.int64), np.zeros(0, np.int32),
                       np.zeros(0, np.int64), np.zeros(0))
    got, st = gpu_multiply(ctx, atoms, empty, a)
    assert got.n_blocks == 0 and st["n_triples_kept"] == 0


def test_zero_size_blocks(ctx):
    atoms, a = S.small_system(25, 81, sizes=(0, 2, 3))
    b = S.decay_matrix(atoms, 6.0, 0.5, 82)
    check(ctx, atoms, a, b, eps=0.0)
    check(ctx, atoms, a, b, eps=1e-3)


def random_fill(blk, fill, seed):
    """Block matrix with each block (i, j) stored with probability fill and
    U[-1, 1] values: the shape of the reference's lamb_matrix_X_test matrices
    (src/lamb_matrix_X_tests.F90:102,128,202: up to 100 atoms, block sizes
    0..3, 70 % of blocks)."""
    rng = np.random.default_rng(seed)
    n = blk.size
    keep = rng.random((n, n)) < fill
    row_p = np.concatenate([[0], np.cumsum(keep.sum(axis=1))]).astype(np.int64)
    rows, col = np.nonzero(keep)
    sizes = blk[rows].astype(np.int64) * blk[col]
    off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    return S.BlockCSR(blk.copy(), row_p, col.astype(np.int32), off,
                      rng.uniform(-1.0, 1.0, int(sizes.sum())))


@pytest.mark.parametrize("n_procs", [1, 3])
def test_reference_test_shapes(ctx, n_procs):
    rng = np.random.default_rng(97 + n_procs)
    n = 100
    blk = rng.integers(0, 4, n).astype(np.int32)
    box = float((n / 0.1) ** (1.0 / 3.0))
    atoms = S.AtomBox(S._wrap32(rng.uniform(0.0, box, (n, 3)), box), blk, box)
    a = random_fill(blk, 0.7, 98)
    b = random_fill(blk, 0.7, 99)
    print(f"case: {n} atoms of sizes 0..3, 70 % fill, {n_procs} procs: A {a.n_blocks} "
          f"blocks", flush=True)
    for eps in (0.0, 0.3):
        check(ctx, atoms, a, b, eps=eps, n_procs=n_procs)


def test_upload_download_round_trip(ctx):
    atoms = S.water_system(200, seed=91, heavy_fraction=0.1)
    a = S.decay_matrix(atoms, 8.0, 0.4, 92)
    dist = Distribution(atoms.blk, atoms.pos, atoms.box)
    lm = LambMatrix.from_global(dist, a)
    dd = D.DeviceDistribution(ctx, dist)
    dm = D.DeviceMatrix.from_lamb(dd, lm)
    back = dm.download()
    for s0, s1 in zip(lm.sectors, back.sectors):
        assert s0.subsets == s1.subsets
        assert np.array_equal(s0.row_p, s1.row_p) and np.array_equal(s0.col_i, s1.col_i)
        assert np.array_equal(s0.extents, s1.extents) and np.array_equal(s0.data, s1.data)
    assert dm.trace() == pytest.approx(oracle.trace(a), rel=1e-14, abs=1e-14)
    dm.close()
    dd.close()


def test_atoms_above_64_rows_stored_but_not_multiplied(ctx):
    """Atoms of 70 and 126 rows (the reference's block size allows up to 126):
    upload, download and trace work; products are refused as unimplemented
    (the MFMA tiles hold at most 64 rows of one atom), with the library left
    usable."""
    base = S.water_system(60, seed=95)
    blk = base.blk.copy()
    blk[3], blk[17] = 70, 126
    atoms = S.AtomBox(base.pos, blk, base.box)
    a = S.decay_matrix(atoms, 8.0, 0.4, 96)
    print(f"case: atoms of {sorted(set(blk.tolist()))} rows, {a.n_blocks} blocks", flush=True)
    dist = Distribution(atoms.blk, atoms.pos, atoms.box)
    lm = LambMatrix.from_global(dist, a)
    dd = D.DeviceDistribution(ctx, dist)
    dm = D.DeviceMatrix.from_lamb(dd, lm)
    back = dm.download()
    for s0, s1 in zip(lm.sectors, back.sectors):
        assert np.array_equal(s0.row_p, s1.row_p) and np.array_equal(s0.col_i, s1.col_i)
        assert np.array_equal(s0.extents, s1.extents) and np.array_equal(s0.data, s1.data)
    assert dm.trace() == pytest.approx(oracle.trace(a), rel=1e-14, abs=1e-14)
    dc = D.DeviceMatrix(dd)
    with pytest.raises(D.LambError) as err:
        D.multiply(dm, dm, dc, eps_filter=1e-7)
    assert err.value.type == -1000  # unimplemented
    assert "block dimension 126 above 64" in str(err.value)
    for h in (dc, dm, dd):
        h.close()
    test_tiny_mixed(ctx)  # the context still multiplies


def test_default_kernel_is_deterministic_and_variants_agree(monkeypatch):
    """The FP64 variant is chosen by rule (half kernel on this sparse box) and
    the autotuner only times tile orders, which permute the tile-to-workgroup
    mapping: two contexts (two processes' worth of state), the tuning call and
    the tuned calls all return identical bits. Each forced variant meets the
    oracle bar on its own."""
    atoms = S.water_system(240, seed=101)
    a = S.decay_matrix(atoms, 8.0, 0.4, 102)
    b = S.decay_matrix(atoms, 8.0, 0.4, 103)
    ref, _ = oracle.multiply(a, b, eps=1e-7)
    got = []
    for _ in range(2):
        fresh = D.Context(0)  # empty tuning cache
        try:
            g, _ = gpu_multiply(fresh, atoms, a, b, eps=1e-7)
            v1, n1 = D.tuning_info(fresh)
            g2, _ = gpu_multiply(fresh, atoms, a, b, eps=1e-7)
            v2, n2 = D.tuning_info(fresh)
            print(f"variant {v1}, tuned classes {n1}", flush=True)
            assert (v1, n1) == (v2, n2) == (2, 1)
            assert np.array_equal(g.data, g2.data)
            got.append(g)
        finally:
            fresh.close()
    assert np.array_equal(got[0].row_p, got[1].row_p)
    assert np.array_equal(got[0].col, got[1].col)
    assert np.array_equal(got[0].data, got[1].data)
    for rb in ("4", "16"):
        monkeypatch.setenv("LAMB_HIP_TILE_ROWBLOCK", rb)
        fresh = D.Context(0)
        try:
            g, _ = gpu_multiply(fresh, atoms, a, b, eps=1e-7)
            assert np.array_equal(g.data, got[0].data), rb
        finally:
            fresh.close()
    monkeypatch.delenv("LAMB_HIP_TILE_ROWBLOCK")
    fresh = D.Context(0)
    try:
        for name in ("quadrant", "half", "ksplit", "lanes", "fine"):
            monkeypatch.setenv("LAMB_HIP_F64_KERNEL", name)
            g, _ = gpu_multiply(fresh, atoms, a, b, eps=1e-7)
            same, err = compare(g, ref)
            assert same and err <= TOL, (name, err)
    finally:
        fresh.close()


@pytest.mark.parametrize("variant", ["quadrant", "half", "ksplit", "lanes", "fine"])
def test_forced_variants_whole_and_partial_products(ctx, monkeypatch, variant):
    """Each FP64 variant on a product whose entries are all whole (unmasked
    loads) and on one where the prefilter cuts entries (per-lane masks)."""
    monkeypatch.setenv("LAMB_HIP_F64_KERNEL", variant)
    atoms = S.water_system(150, seed=111)
    a = S.decay_matrix(atoms, 8.0, 0.4, 112)
    b = S.decay_matrix(atoms, 8.0, 0.4, 113)
    check(ctx, atoms, a, b, eps=1e-7)
    atoms, a = S.small_system(60, 31, cutoff=9.0, decay=1.5, sizes=(1, 5, 13))
    b = S.decay_matrix(atoms, 9.0, 1.3, 32)
    st, rst = check(ctx, atoms, a, b, eps=2e-3)
    assert rst["n_triples_filtered"] > 0


@pytest.mark.parametrize("case", ["water_eps", "filters_bind", "alpha_beta", "three_kinds"])
def test_symbolic_global_memory_forms(ctx, monkeypatch, case):
    """The global-memory forms of the symbolic kernels (taken for systems with
    more bands than fit the per-band LDS buffers) against the oracle."""
    monkeypatch.setenv("LAMB_HIP_SYM_LDS", "0")
    {"water_eps": test_water_eps, "filters_bind": test_filters_bind,
     "alpha_beta": test_alpha_beta, "three_kinds": test_three_kinds}[case](ctx)


@pytest.mark.parametrize("size", [5, 23, 64])
def test_block_norm_order_near_threshold(ctx, size):
    """Block norms (the prefilter's and the stand-alone post-filter's) are one
    sequential column-major FMA chain per block, including blocks longer than
    one LDS chunk of the norm kernel (23 x 23 = 529 and 64 x 64 = 4,096
    values). eps is placed so that fl(eps * eps) lies between the chain's sum
    and the reversed chain's for one block: the filter must keep or drop it as
    the oracle's sequential norm says."""
    atoms, x = S.small_system(24 if size == 64 else 40, 17 + size, sizes=(size,),
                              cutoff=7.0, decay=0.3)
    seq = oracle.block_nsq(x)
    pick = None
    for t in range(x.n_blocks):
        v = x.data[x.blk_off[t]:x.blk_off[t] + size * size]
        alt = oracle.nsq1(v[::-1].copy())
        if alt != seq[t]:
            lo, hi = sorted((alt, float(seq[t])))
            eps = _eps_between(lo, hi)
            if eps is not None:
                pick = (t, eps)
                break
    assert pick is not None, "no block whose two summation orders differ"
    t, eps = pick
    keep = seq >= eps * eps
    print(f"case: {size}x{size} blocks, block {t} seq={seq[t]!r} eps^2={eps * eps!r} "
          f"kept {int(keep.sum())} of {x.n_blocks}", flush=True)
    dist = Distribution(atoms.blk, atoms.pos, atoms.box)
    dd = D.DeviceDistribution(ctx, dist)
    dm = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, x))
    try:
        dropped = D.matrix_filter(dm, eps)
        got = dm.download().to_global()
    finally:
        dm.close()
        dd.close()
    assert dropped == int((~keep).sum())
    rows = np.repeat(np.arange(x.n_atoms), np.diff(x.row_p))
    want = set(zip(rows[keep].tolist(), x.col[keep].tolist()))
    grows = np.repeat(np.arange(got.n_atoms), np.diff(got.row_p))
    assert set(zip(grows.tolist(), got.col.tolist())) == want


@pytest.mark.parametrize("sizes", [(64,), (33, 31), (47, 17), (63, 2, 1)])
@pytest.mark.parametrize("variant", ["quadrant", "half", "ksplit", "lanes", "fine", "bf16"])
def test_blocks_at_the_band_limit(ctx, monkeypatch, variant, sizes):
    """Atoms of up to 64 rows, the band height: bands of one 64-row atom, of
    two atoms that fill 64 rows exactly, and of many 1- and 2-row atoms (16
    atoms, the band's atom limit). A dense product (every pair within the
    cutoff) with a filter that binds near the cutoff, in each FP64 variant;
    the BF16 split without a filter, within 1e-9."""
    atoms, a = S.small_system(24, 90 + sizes[0], sizes=sizes, cutoff=7.0, decay=1.5)
    b = S.decay_matrix(atoms, 7.0, 1.5, 91 + sizes[0])
    if variant == "bf16":
        print(f"case: bf16 sizes={sizes} atoms={atoms.n_atoms} blocks={a.n_blocks}", flush=True)
        got, st = gpu_multiply(ctx, atoms, a, b, flags=D.BF16_SPLIT)
        ref, rst = oracle.multiply(a, b)
        same, err = compare(got, ref)
        assert same and err <= 1e-9, err
        assert st["n_triples_kept"] == rst["n_triples_kept"]
        return
    monkeypatch.setenv("LAMB_HIP_F64_KERNEL", variant)
    print(f"variant {variant}, sizes {sizes}", flush=True)
    check(ctx, atoms, a, b, eps=2e-3)
