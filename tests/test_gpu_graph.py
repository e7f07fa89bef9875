"""McWeeny chains captured as HIP graphs.

A replay recomputes, with the recorded sizes, exactly the chain the eager
path computes: the resulting matrix, traces and idempotency errors must be
bit-identical to the eager chain's (same kernels in the same order), on two
replays in a row, in FP64 and BF16-split form. A replay whose input values
change the filter decisions must fail cleanly (precondition error) and
leave the library usable: once with the values scaled down, so that the
replay's sizes are smaller than the recorded ones, and once scaled up under
a binding filter, so that they are larger.
"""
import numpy as np
import pytest

from lamb_amd import device as D
from lamb_amd import systems as S
from lamb_amd.matrix import Distribution, LambMatrix

pytestmark = pytest.mark.gpu


def batch_h(ctx, seed):
    b = S.config4(seed=seed, n_systems=3, lo=60, hi=180)
    dist = Distribution(b.atoms.blk, b.atoms.pos, b.atoms.box)
    dd = D.DeviceDistribution(ctx, dist)
    dd.set_groups(b.group, b.n_systems)
    dh = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, b.h))
    return b, dd, dh


@pytest.mark.parametrize("flags", [D.FP64, D.bf16_split(5)])
def test_replay_matches_eager_chain(ctx, flags):
    b, dd, dh = batch_h(ctx, 21)
    steps, eps = 4, 1e-7
    print(f"case: graph chain flags={flags} atoms={b.atoms.n_atoms} steps={steps}", flush=True)
    pe = D.purification_start(dh)
    te, ie, _ = D.mcweeny(pe, steps, eps_filter=eps, flags=flags)
    want = pe.download().to_global()
    p0 = D.purification_start(dh)
    g = D.ChainGraph(p0, steps, eps_filter=eps, flags=flags)
    print("  graph", g.info(), flush=True)
    out = D.DeviceMatrix(dd)
    for rep in range(2):
        tg, ig, ms = g.replay(p0, out)
        got = out.download().to_global()
        print(f"  replay {rep}: {ms:.3f} ms", flush=True)
        assert np.array_equal(got.row_p, want.row_p) and np.array_equal(got.col, want.col)
        assert np.array_equal(got.data, want.data)
        assert np.array_equal(tg, te) and np.array_equal(ig, ie)
    for x in (g, out, p0, pe, dh, dd):
        x.close()


def test_replay_with_other_filter_decisions_fails_cleanly(ctx):
    b, dd, dh = batch_h(ctx, 22)
    p0 = D.purification_start(dh)
    g = D.ChainGraph(p0, 3, eps_filter=1e-7)
    small = D.DeviceMatrix(dd)
    small.copy_from(p0)
    small.group_scale_shift(1e-3, 0.0)  # weaker blocks: the filters drop more
    out = D.DeviceMatrix(dd)
    with pytest.raises(D.LambError) as err:
        g.replay(small, out)
    assert err.value.type == 200  # precondition
    # the library keeps working: a replay of the captured input and an eager chain
    g.replay(p0, out)
    D.mcweeny(small, 2, eps_filter=1e-7)
    for x in (g, out, small, p0, dh, dd):
        x.close()


def test_replay_with_growing_sizes_fails_cleanly(ctx):
    """Values scaled up with a binding filter: the replay's products keep more
    triples and C blocks than the recorded chain, so its sizes exceed the
    buffers captured in the graph. The size check at the first differing read
    stops every later kernel before it writes through the larger size: the
    replay fails with a precondition error, and the graph afterwards still
    computes the eager chain bit for bit (nothing it holds was overwritten)."""
    b, dd, dh = batch_h(ctx, 23)
    steps, eps = 3, 1e-3
    p0 = D.purification_start(dh)
    big = D.DeviceMatrix(dd)
    big.copy_from(p0)
    big.group_scale_shift(10.0, 0.0)  # stronger blocks: the filters keep more
    c = D.DeviceMatrix(dd)
    s0 = D.multiply(p0, p0, c, eps_filter=eps)
    s1 = D.multiply(big, big, c, eps_filter=eps)
    print(f"case: growing replay, atoms={b.atoms.n_atoms}, eps {eps}: first product "
          f"triples {s0['n_triples_kept']} -> {s1['n_triples_kept']}, C blocks "
          f"{s0['n_c_blocks_kept']} -> {s1['n_c_blocks_kept']}", flush=True)
    assert s1["n_triples_kept"] > s0["n_triples_kept"]
    assert s1["n_c_blocks_kept"] > s0["n_c_blocks_kept"]
    pe = D.DeviceMatrix(dd)
    pe.copy_from(p0)
    te, ie, _ = D.mcweeny(pe, steps, eps_filter=eps)
    want = pe.download().to_global()
    g = D.ChainGraph(p0, steps, eps_filter=eps)
    out = D.DeviceMatrix(dd)
    with pytest.raises(D.LambError) as err:
        g.replay(big, out)
    assert err.value.type == 200  # precondition
    for rep in range(2):
        tg, ig, _ = g.replay(p0, out)
        got = out.download().to_global()
        print(f"  replay {rep} of the captured input after the failed one", flush=True)
        assert np.array_equal(got.row_p, want.row_p) and np.array_equal(got.col, want.col)
        assert np.array_equal(got.data, want.data)
        assert np.array_equal(tg, te) and np.array_equal(ig, ie)
    for x in (g, out, pe, c, big, p0, dh, dd):
        x.close()
