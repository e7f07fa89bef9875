"""Benchmark: one filtered block-sparse multiply C = A*B per step on one MI355X.

Workload: 2,048 H2O (6,144 atoms) in a 39.4 A periodic box, O 13x13 / H 5x5
blocks (dimension 47,104), block (i,j) stored iff r_ij < 8 A, entries
U[-1,1]*exp(-0.4 r), A and B from independent value seeds, eps_filter 1e-7.
A step is lamb_hip_multiply on device-resident A and B (symbolic phase,
numeric kernel, post-filter and C index). Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "block-SpGEMM GFLOP/s & HBM GB/s: FP64 plain vs BF16x3 batched at ≤1e-9 err"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of an FP64 numeric kernel on this workload, by kernel
# name, tile-order row block and launch form (0 one workgroup per tile, 1
# persistent grid over the interleaved order, 2 persistent grid with one range
# of the order per XCD):
# rocprofv3 FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 (scale factors from a
# device copy of known size in the same run), from
# profiles/numeric_f64_half_kernel_water6144.md. A kernel, order or launch
# form without a profile reports null.
PROFILED_TRAFFIC_BYTES = {
    ("numeric_f64_half_kernel", 4, 0): 109.94e9,
    ("numeric_f64_half_kernel", 4, 1): 117.64e9,
    ("numeric_f64_half_kernel", 4, 2): 87.37e9,
    ("numeric_f64_half_kernel", 8, 0): 108.31e9,
    ("numeric_f64_half_kernel", 8, 1): 114.90e9,
    ("numeric_f64_half_kernel", 8, 2): 97.30e9,
    ("numeric_f64_half_kernel", 16, 0): 121.87e9,
    ("numeric_f64_half_kernel", 16, 1): 128.66e9,
    ("numeric_f64_half_kernel", 16, 2): 106.19e9,
}
LAUNCH_FORMS = {0: "one workgroup per tile", 1: "persistent grid, XCDs share one window",
                2: "persistent grid, one range of the tile order per XCD"}
# FP64 numeric kernel of each variant (lamb_hip_tuning_info)
F64_KERNELS = {0: "numeric_f64_kernel", 1: "numeric_f64_ksplit_kernel",
               2: "numeric_f64_half_kernel", 3: "numeric_f64_lanes_kernel",
               4: "numeric_f64_half_kernel (fine)"}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--seed", type=int, default=2)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-rows", type=int, default=2,
                   help="CPU baseline samples every n-th block row of A")
    p.add_argument("--slices", type=int, default=5, help="BF16 slices of the split form")
    p.add_argument("--no-bf16", action="store_true", help="skip the BF16-split leg")
    p.add_argument("--workload", default="water6144",
                   choices=["water6144", "config3", "config4", "config5"],
                   help="water6144: the headline multiply; config3: BF16 split on the "
                        "512-atom box; config4: batched McWeeny chains (eager and graph); "
                        "config5: gradient of tr(P_k H) through the adjoint kernels")
    p.add_argument("--systems", type=int, default=64, help="config4: systems in the batch")
    p.add_argument("--lo", type=int, default=96, help="config4: smallest system (atoms)")
    p.add_argument("--hi", type=int, default=1024,
                   help="config4: largest system (atoms). Iterates fill in to dense after one "
                        "McWeeny step; with systems up to 3,072 atoms a dense iterate of the "
                        "64-system batch is 78 GB and a step does not fit one GPU")
    p.add_argument("--mem-groups", action="store_true",
                   help="config4: run the batch in groups of systems sized to device memory, "
                        "each group's chain captured as one graph (for the full 96..3,072 range)")
    p.add_argument("--replays", type=int, default=1,
                   help="config4 --mem-groups: graph replays per group (compared bit for bit)")
    p.add_argument("--mem-fraction", type=float, default=0.4,
                   help="config4 --mem-groups: share of free device memory a group may use "
                        "(by the estimate CHAIN_BYTES_PER_DENSE_BYTE; smaller groups also keep "
                        "each group's chain under a minute)")
    p.add_argument("--chain-steps", type=int, default=0,
                   help="McWeeny steps (default 20 for config4, 10 for config5)")
    p.add_argument("--tune-shapes", action="store_true",
                   help="let the autotuner choose the FP64 tile shape per workload class "
                        "(lamb_hip_tuning_mode); the shapes it kept are reported")
    p.add_argument("--atoms", type=int, default=3072, help="config5: atoms of the water box")
    p.add_argument("--fd-blocks", type=int, default=0, help="config5: blocks checked by FD")
    p.add_argument("--fd-first", type=int, default=0,
                   help="config5: first of the 32 drawn blocks checked by this run")
    return p.parse_args()


def cpu_baseline(a, b, eps, every):
    """Time the CPU FP64 oracle on a row sample of the same product, on all
    host threads and (on a 16x smaller sample) on one."""
    import oracle
    from lamb_amd.systems import BlockCSR

    def sample(step):
        rows = np.arange(a.n_atoms)
        keep = (rows % step) == 0
        cnt = np.where(keep, np.diff(a.row_p), 0)
        row_p = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
        sel = np.repeat(keep, np.diff(a.row_p))
        return BlockCSR(a.blk, row_p, a.col[sel], a.blk_off[sel], a.data), int(keep.sum())

    oracle.lib()
    cores = oracle.num_threads()
    sub, n_rows = sample(every)
    t0 = time.perf_counter()
    c, st = oracle.multiply(sub, b, eps=eps)
    dt = time.perf_counter() - t0
    sub1, n_rows1 = sample(16 * every)
    oracle.set_threads(1)
    try:
        t1 = time.perf_counter()
        _, st1 = oracle.multiply(sub1, b, eps=eps)
        dt1 = time.perf_counter() - t1
    finally:
        oracle.set_threads(cores)
    return {
        "value": st["flops"] / dt / 1e9,
        "unit": "GFLOP/s",
        "cores": cores,
        "kind": "port",
        "sample": f"every {every}th block row of A ({n_rows} of {a.n_atoms} rows), "
                  f"{st['flops']:.3e} flop in {dt:.2f} s, FP64 C oracle, OpenMP",
        "value_1_core": st1["flops"] / dt1 / 1e9,
        "sample_1_core": f"every {16 * every}th block row of A ({n_rows1} rows), "
                         f"{st1['flops']:.3e} flop in {dt1:.2f} s on one thread",
    }


def shape_choices(D, ctx):
    """{tile shape: classes} of the context's tuning table (shape entries)."""
    import tempfile
    names = {0: "quadrant 32x32", 1: "ksplit 64x64", 2: "half 64x32", 3: "lanes 16x32", 4: "fine 64x32"}
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "tuning.txt")
        D.tuning_save(ctx, path)
        lines = open(path).read().splitlines()[1:]
    out = {}
    for ln in lines:
        f = ln.split()
        if f[0] == "shape":
            out[names.get(int(f[-1]), f[-1])] = out.get(names.get(int(f[-1]), f[-1]), 0) + 1
    return out


def bf16_leg(D, da, db, dc, dd, eps, args, ms_step64, ms_num64):
    """The BF16-split form on the same product: throughput, speed-up over the
    FP64 path and relative Frobenius error against the FP64 result (dc)."""
    flags = D.bf16_split(args.slices)
    c16 = D.DeviceMatrix(dd)
    for _ in range(max(args.warmup, 1)):
        st = D.multiply(da, db, c16, eps_filter=eps, flags=flags)
    sts = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sts.append(D.multiply(da, db, c16, eps_filter=eps, flags=flags))
    ms_step = (time.perf_counter() - t0) * 1e3 / args.steps
    ms_num = float(np.mean([s["ms_numeric"] for s in sts]))
    flops = sts[-1]["flops"]
    err = float(c16.group_diff_norm(dc)[0] / dc.group_norm()[0])
    same_pattern = c16.info() == dc.info()
    c16.close()
    print(f"bf16 split ({args.slices} slices): {ms_step:.1f} ms/step, numeric {ms_num:.2f} ms, "
          f"error {err:.2e}", file=sys.stderr, flush=True)
    return {
        "slices": args.slices,
        "products_per_fp64_product": args.slices * (args.slices + 1) // 2,
        "gflops": flops / (ms_step * 1e-3) / 1e9,
        "ms_per_step": ms_step,
        "ms_numeric": ms_num,
        "speedup_step_vs_fp64": ms_step64 / ms_step,
        "speedup_numeric_vs_fp64": ms_num64 / ms_num,
        "rel_frobenius_error_vs_fp64": err,
        "same_block_count": bool(same_pattern),
        "mfma_bf16_achieved_tflops": flops * args.slices * (args.slices + 1) / 2 / (ms_num * 1e-3) / 1e12,
    }


def line(metric, value, unit, ms, args, config, dtype="f64", **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": dtype, "data": "synthetic (seeded water-like atom box)",
           "config": config}
    out.update(extra)
    out["library"] = library()
    return out


def library():
    """The HIP library this run used (path, hash, build record match)."""
    from lamb_amd import _native
    info = _native.library_info()
    if "sha256" in info:
        info["sha256"] = info["sha256"][:16]
    return info


def run_config3(args):
    """BF16 split (and FP64) on the 512-atom box, errors against the CPU oracle."""
    import oracle
    from __graft_entry__ import compare
    from lamb_amd import device as D
    from lamb_amd import systems as S
    from lamb_amd.matrix import Distribution, LambMatrix
    atoms, a, b, eps = S.config1(args.seed if args.seed != 2 else 1)
    ref, _ = oracle.multiply(a, b, eps=eps)
    dist = Distribution(atoms.blk, atoms.pos, atoms.box)
    ctx = D.Context(0)
    dd = D.DeviceDistribution(ctx, dist)
    da = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, a))
    db = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, b))
    res = {}
    for name, flags in (("fp64", D.FP64), ("bf16_split", D.bf16_split(args.slices))):
        dc = D.DeviceMatrix(dd)
        for _ in range(max(args.warmup, 1)):
            D.multiply(da, db, dc, eps_filter=eps, flags=flags)
        t0 = time.perf_counter()
        sts = [D.multiply(da, db, dc, eps_filter=eps, flags=flags) for _ in range(args.steps)]
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        same, err = compare(dc.download().to_global(), ref)
        res[name] = {"gflops": sts[-1]["flops"] / (ms * 1e-3) / 1e9, "ms_per_step": ms,
                     "ms_numeric": float(np.mean([x["ms_numeric"] for x in sts])),
                     "rel_frobenius_error_vs_oracle": err, "pattern_identical": bool(same)}
        dc.close()
    for h in (db, da, dd):
        h.close()
    ctx.close()
    r = res["bf16_split"]
    return line("block-SpGEMM GFLOP/s, BF16 split vs FP64 on the 512-atom box", r["gflops"],
                "GFLOP/s", r["ms_per_step"], args,
                {"workload": "512 atoms, U[0,17.4)^3, block 13, r<6 A, exp(-0.5r), A=B "
                             "symmetric + 2I, eps 1e-8", "slices": args.slices, **res},
                dtype="bf16x%d-split/f32-mfma/f64-accumulate" % args.slices)


def batch_distribution(atoms, Distribution):
    """The spatial partition with the fewest subsets whose local indices fit
    int16 (a batch of many systems can exceed 32,767 atoms in one subset)."""
    for n_procs in (1, 2, 4, 8, 16, 32, 64):
        try:
            return Distribution(atoms.blk, atoms.pos, atoms.box, n_procs=n_procs)
        except ValueError:
            continue
    raise ValueError("no spatial partition keeps every subset within the int16 local index")


# Device bytes per dense-iterate byte of a captured McWeeny chain: the
# iterate, its square and the update with a spare buffer each (6), the
# captured input and the two packed operand panels (about 2.5), the tile
# entries of a dense product (up to about 2 for 3,000-atom systems), and
# headroom for the symbolic scratch.
CHAIN_BYTES_PER_DENSE_BYTE = 12.0


def memory_groups(systems, budget):
    """Systems split into groups whose estimated chain memory fits budget
    bytes: largest first, each into the first group with room (first-fit
    decreasing). Returns lists of indices into systems."""
    need = []
    for atoms, _ in systems:
        dim = float(np.sum(atoms.blk))
        need.append(CHAIN_BYTES_PER_DENSE_BYTE * 8.0 * dim * dim)
    groups, room = [], []
    for i in sorted(range(len(systems)), key=lambda i: -need[i]):
        for g in range(len(groups)):
            if need[i] <= room[g]:
                groups[g].append(i)
                room[g] -= need[i]
                break
        else:
            if need[i] > budget:
                raise ValueError(f"system {i} needs {need[i] / 1e9:.1f} GB, above the "
                                 f"{budget / 1e9:.1f} GB budget")
            groups.append([i])
            room.append(budget - need[i])
    return [sorted(g) for g in groups], need


def run_config4_groups(args, bt, steps, eps):
    """The batch in memory-bounded groups: per group one distribution, the
    chain captured as a HIP graph (its record pass is the eager chain) and
    replayed --replays times; flops and the fastest replay of each group summed
    over groups."""
    import torch
    from lamb_amd import device as D
    from lamb_amd import systems as S
    from lamb_amd.matrix import Distribution, LambMatrix
    free, total = torch.cuda.mem_get_info(0)
    budget = args.mem_fraction * free
    groups, need = memory_groups(bt.systems, budget)
    print(f"{len(groups)} groups within {budget / 1e9:.0f} GB of {free / 1e9:.0f} GB free: "
          f"{[len(g) for g in groups]} systems", file=sys.stderr, flush=True)
    n = bt.n_systems
    n_classes = 0
    shapes = {}
    final_idem = np.full(n, np.nan)
    final_trace = np.full(n, np.nan)
    flops = ms_replay = ms_capture = 0.0
    replays_equal = True
    per_group = []
    for gi, idx in enumerate(groups):
        gb = S.batch([bt.systems[i] for i in idx])
        dist = batch_distribution(gb.atoms, Distribution)
        ctx = D.Context(0)  # a context per group: its scratch goes with it
        D.tuning_mode(ctx, args.tune_shapes)
        dd = D.DeviceDistribution(ctx, dist)
        dd.set_groups(gb.group, gb.n_systems)
        dh = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, gb.h))
        p0 = D.purification_start(dh)
        t0 = time.perf_counter()
        g = D.ChainGraph(p0, steps, eps_filter=eps)
        t_cap = (time.perf_counter() - t0) * 1e3
        print(f"group {gi + 1}/{len(groups)}: {len(idx)} systems, {gb.atoms.n_atoms} atoms; "
              f"recorded and captured in {t_cap:.0f} ms", file=sys.stderr, flush=True)
        out = D.DeviceMatrix(dd)
        reps = [g.replay(p0, out) for _ in range(max(args.replays, 1))]
        tr2, id2, _ = reps[-1]
        replays_equal &= all(np.array_equal(r[0], tr2) and np.array_equal(r[1], id2)
                             for r in reps)
        f = g.info()["flops"]
        ms = min(r[2] for r in reps)
        flops += f
        ms_replay += ms
        ms_capture += t_cap
        final_idem[idx] = id2[-1]
        final_trace[idx] = tr2[-1]
        per_group.append({"systems": len(idx), "atoms": int(gb.atoms.n_atoms),
                          "flops": f, "ms_replay": ms, "ms_capture": t_cap})
        print(f"group {gi + 1}/{len(groups)}: est. {sum(need[i] for i in idx) / 1e9:.0f} GB; "
              f"replay {ms:.0f} ms, {f / (ms * 1e-3) / 1e12:.1f} TFLOP/s", file=sys.stderr,
              flush=True)
        for h in (g, out, p0, dh, dd):
            h.close()
        n_classes += D.tuning_info(ctx)[1]
        for k, v in shape_choices(D, ctx).items():
            shapes[k] = shapes.get(k, 0) + v
        ctx.close()
    sizes = [a.n_atoms for a, _ in bt.systems]
    return line("batched McWeeny chains: GFLOP/s of the chains replayed as HIP graphs, in "
                "memory-bounded groups", flops / (ms_replay * 1e-3) / 1e9, "GFLOP/s", ms_replay,
                args,
                {"workload": f"{n} water-like systems, {min(sizes)}..{max(sizes)} atoms"
                             f" (log-uniform, seed 4), O 13 / H 5 / 23 for 10%, r<8 A, "
                             f"{steps} McWeeny steps from H scaled to [0,1], eps 1e-7, in "
                             f"{len(groups)} groups sized to device memory",
                 "atoms": int(bt.atoms.n_atoms), "chain_steps": steps, "groups": per_group,
                 "flops_per_chain": flops, "ms_graph_replay": ms_replay,
                 "ms_capture_incl_record": ms_capture, "replays_equal_bitwise": replays_equal,
                 "final_idempotency_error": final_idem.tolist(),
                 "final_trace": final_trace.tolist(),
                 "max_final_idempotency_error": float(np.nanmax(final_idem)),
                 "autotuned_classes": n_classes, "tile_shapes_tuned": shapes})


def run_config4(args):
    """A batch of independent systems, McWeeny chains eager and as a HIP graph."""
    from lamb_amd import device as D
    from lamb_amd import systems as S
    from lamb_amd.matrix import Distribution, LambMatrix
    steps = args.chain_steps or 20
    eps = 1e-7
    t0 = time.perf_counter()
    bt = S.config4(seed=4, n_systems=args.systems, lo=args.lo, hi=args.hi)
    sizes = [a.n_atoms for a, _ in bt.systems]
    print(f"batch of {bt.n_systems} systems, {bt.atoms.n_atoms} atoms ({min(sizes)}..{max(sizes)}),"
          f" built in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    if args.mem_groups:
        return run_config4_groups(args, bt, steps, eps)
    dist = batch_distribution(bt.atoms, Distribution)
    ctx = D.Context(0)
    D.tuning_mode(ctx, args.tune_shapes)
    dd = D.DeviceDistribution(ctx, dist)
    dd.set_groups(bt.group, bt.n_systems)
    dh = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, bt.h))
    flags = D.FP64
    p = D.purification_start(dh)
    t0 = time.perf_counter()
    tr, idem, st = D.mcweeny(p, steps, eps_filter=eps, flags=flags)
    ms_eager = (time.perf_counter() - t0) * 1e3
    print(f"eager chain: {ms_eager:.0f} ms, {st['flops']:.3e} flop", file=sys.stderr, flush=True)
    p0 = D.purification_start(dh)
    t0 = time.perf_counter()
    g = D.ChainGraph(p0, steps, eps_filter=eps, flags=flags)
    ms_capture = (time.perf_counter() - t0) * 1e3
    print(f"graph capture: {ms_capture:.0f} ms", file=sys.stderr, flush=True)
    out = D.DeviceMatrix(dd)
    reps = []
    for _ in range(max(args.steps // 5, 1)):
        tg, ig, ms = g.replay(p0, out)
        reps.append(ms)
        print(f"graph replay: {ms:.0f} ms", file=sys.stderr, flush=True)
    same = bool(np.array_equal(tg, tr) and np.array_equal(ig, idem))
    flops = g.info()["flops"]
    ms_graph = float(np.median(reps))
    occ = np.round(tr[-1]).astype(int)
    g.close()
    out.close()
    bf = None
    if not args.no_bf16:
        # the same chains in the BF16 split form, eager and as a graph; the
        # final iterates are compared with the FP64 ones per system
        fb = D.bf16_split(args.slices)
        pb = D.purification_start(dh)
        t0 = time.perf_counter()
        trb, idb, stb = D.mcweeny(pb, steps, eps_filter=eps, flags=fb)
        ms_eager_b = (time.perf_counter() - t0) * 1e3
        diff = pb.group_diff_norm(p)
        ref = p.group_norm()
        gb = D.ChainGraph(p0, steps, eps_filter=eps, flags=fb)
        outb = D.DeviceMatrix(dd)
        repb = [gb.replay(p0, outb)[2] for _ in range(max(args.steps // 5, 1))]
        ms_graph_b = float(np.median(repb))
        flops_b = gb.info()["flops"]
        bf = {"slices": args.slices, "ms_eager_chain": ms_eager_b, "ms_graph_replay": ms_graph_b,
              "gflops_graph": flops_b / (ms_graph_b * 1e-3) / 1e9,
              "speedup_graph_vs_fp64": ms_graph / ms_graph_b,
              "max_rel_frobenius_diff_final_p_vs_fp64": float(np.max(diff / np.maximum(ref, 1e-300))),
              "same_block_count": pb.info()[0] == p.info()[0],
              "max_trace_diff_final": float(np.max(np.abs(trb[-1] - tr[-1])))}
        print(f"bf16 x{args.slices}: eager {ms_eager_b:.0f} ms, graph {ms_graph_b:.1f} ms, "
              f"final-P diff {bf['max_rel_frobenius_diff_final_p_vs_fp64']:.2e}", file=sys.stderr,
              flush=True)
        for h in (gb, outb, pb):
            h.close()
    for h in (p0, p, dh, dd):
        h.close()
    v, ncls = D.tuning_info(ctx)
    shapes = shape_choices(D, ctx)
    ctx.close()
    return line("batched McWeeny chains: GFLOP/s of the chain replayed as a HIP graph",
                flops / (ms_graph * 1e-3) / 1e9, "GFLOP/s", ms_graph, args,
                {"workload": f"{bt.n_systems} water-like systems, {min(sizes)}..{max(sizes)} atoms"
                             f" (log-uniform, seed 4), O 13 / H 5 / 23 for 10%, r<8 A, "
                             f"{steps} McWeeny steps from H scaled to [0,1], eps 1e-7",
                 "atoms": int(bt.atoms.n_atoms), "chain_steps": steps,
                 "flops_per_chain": flops, "ms_eager_chain": ms_eager,
                 "gflops_eager": flops / (ms_eager * 1e-3) / 1e9, "ms_graph_replay": ms_graph,
                 "ms_capture": ms_capture, "graph_equals_eager_bitwise": same,
                 "max_final_idempotency_error": float(idem[-1].max()),
                 "occupied_counts_first_systems": occ[:8].tolist(),
                 "autotuned_classes": ncls, "tile_shapes_tuned": shapes, "bf16_split": bf})


def run_config5(args):
    """Gradient of L = tr(P_k H) through the adjoint kernels (torch.autograd)."""
    import torch
    from lamb_amd import autograd as AG
    from lamb_amd import device as D
    from lamb_amd import systems as S
    from lamb_amd.matrix import Distribution, LambMatrix
    steps = args.chain_steps or 10
    atoms = S.water_system(args.atoms, seed=5)
    h = S.decay_matrix(atoms, 8.0, 0.4, 6, symmetric=True)
    dist = Distribution(atoms.blk, atoms.pos, atoms.box)
    ctx = D.Context(0)
    dd = D.DeviceDistribution(ctx, dist)
    dh = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, h))
    lo, hi = (float(x[0]) for x in dh.group_gershgorin())
    hv, ph = AG.leaf(dh)
    res = {}
    grads = {}
    legs = [("fp64", D.FP64)]
    if not args.no_bf16:
        legs.append(("bf16_split", D.bf16_split(args.slices)))
    for name, flags in legs:
        # one untimed pass (autotuning, first torch calls)
        w = hv.clone().requires_grad_(True)
        AG.purification_loss(w, ph, atoms.blk, lo, hi, min(steps, 2), flags=flags).backward()
        x = hv.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loss = AG.purification_loss(x, ph, atoms.blk, lo, hi, steps, flags=flags)
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        grads[name] = x.grad.detach()
        res[name] = {"loss": float(loss.detach()), "ms_forward": (t1 - t0) * 1e3,
                     "ms_forward_backward": (t2 - t0) * 1e3}
        print(f"{name}: loss {float(loss.detach()):.10g}, forward {res[name]['ms_forward']:.0f} ms, "
              f"forward+backward {res[name]['ms_forward_backward']:.0f} ms", file=sys.stderr,
              flush=True)
    gdiff = None
    if "bf16_split" in grads:
        gdiff = float(torch.linalg.norm(grads["bf16_split"] - grads["fp64"]) /
                      torch.linalg.norm(grads["fp64"]))
    fd = None
    if args.fd_blocks:
        rows, cols = ph.block_atoms()
        sz = atoms.blk[rows].astype(np.int64) * atoms.blk[cols]
        offs = np.concatenate([[0], np.cumsum(sz)])
        # 32 blocks drawn once; this run checks picks [fd_first, fd_first + fd_blocks)
        picks = np.random.default_rng(7).choice(len(rows), size=max(32, args.fd_blocks),
                                                replace=False)
        worst = 0.0
        with torch.no_grad():
            for n, k in enumerate(picks[args.fd_first:args.fd_first + args.fd_blocks]):
                d = torch.zeros_like(hv)
                v = np.random.default_rng((7, int(k))).normal(size=int(sz[k]))
                d[offs[k]:offs[k + 1]] = torch.from_numpy(v / np.linalg.norm(v)).to(d.device)
                lp = AG.purification_loss(hv + 1e-6 * d, ph, atoms.blk, lo, hi, steps).item()
                lm = AG.purification_loss(hv - 1e-6 * d, ph, atoms.blk, lo, hi, steps).item()
                an = float(torch.dot(grads["fp64"], d))
                e = abs((lp - lm) / 2e-6 - an) / max(abs(an), 1.0)
                worst = max(worst, e)
                print(f"fd block {args.fd_first + n} (block {int(k)}): dL.d {an:.9e}, "
                      f"scaled error {e:.2e}", file=sys.stderr, flush=True)
        fd = worst
    dd.close()
    ctx.close()
    r = res["fp64"]
    return line("gradient of tr(P_k H) through hand-written adjoint kernels: ms per "
                "forward+backward", r["ms_forward_backward"], "ms", r["ms_forward_backward"],
                args, {"workload": f"{atoms.n_atoms}-atom water box, O 13 / H 5, r<8 A, "
                                   f"L = tr(P_{steps} H), {steps} McWeeny steps, checkpoint "
                                   f"every 2", **res,
                       "grad_rel_diff_bf16_vs_fp64": gdiff,
                       "fd_blocks": args.fd_blocks, "fd_first": args.fd_first,
                       "fd_worst_scaled_error": fd},
                higher_is_better=False)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the sector multiply is measured on one GPU's share)", file=sys.stderr)
        sys.exit(2)
    if args.workload != "water6144":
        fn = {"config3": run_config3, "config4": run_config4, "config5": run_config5}
        print(json.dumps(fn[args.workload](args)), flush=True)
        return
    from lamb_amd import device as D
    from lamb_amd import systems as S
    from lamb_amd.matrix import Distribution, LambMatrix

    t_setup = time.perf_counter()
    atoms, a, b, eps = S.config2(args.seed)
    dist = Distribution(atoms.blk, atoms.pos, atoms.box)
    la = LambMatrix.from_global(dist, a)
    lb = LambMatrix.from_global(dist, b)
    ctx = D.Context(0)
    D.tuning_mode(ctx, args.tune_shapes)
    dd = D.DeviceDistribution(ctx, dist)
    # host sectors -> device pools (the boundary's upload and commit), timed
    # to report the PCIe-inclusive rate beside the device-resident one
    t_up = time.perf_counter()
    da = D.DeviceMatrix.from_lamb(dd, la)
    db = D.DeviceMatrix.from_lamb(dd, lb)
    ms_upload = (time.perf_counter() - t_up) * 1e3
    dc = D.DeviceMatrix(dd)
    print(f"setup {time.perf_counter() - t_setup:.1f}s: {atoms.n_atoms} atoms, "
          f"A {a.n_blocks} blocks, B {b.n_blocks} blocks, upload+commit of A and B "
          f"{ms_upload:.1f} ms", file=sys.stderr, flush=True)

    for w in range(args.warmup):
        st = D.multiply(da, db, dc, eps_filter=eps)
        print(f"warmup {w}: {st['ms_total']:.1f} ms (strips+tiles {st['ms_symbolic']:.2f} ms, "
              f"numeric {st['ms_numeric']:.2f} ms, C index {st['ms_index']:.2f} ms)",
              file=sys.stderr, flush=True)
    sts = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        sts.append(D.multiply(da, db, dc, eps_filter=eps))
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    st = sts[-1]
    flops = st["flops"]
    # The block norms of A and B are computed once per operand values and
    # cached across the timed products (a repeated product with unchanged
    # operands). One product on fresh copies of A and B, whose norms are not
    # computed yet, shows what a step pays when the operands change, as in a
    # purification chain.
    fa, fb = D.DeviceMatrix(dd), D.DeviceMatrix(dd)
    fa.copy_from(da)
    fb.copy_from(db)  # the copy waits for the device and drops the cached norms
    t0 = time.perf_counter()
    D.multiply(fa, fb, dc, eps_filter=eps)
    ms_fresh = (time.perf_counter() - t0) * 1e3
    fa.close()
    fb.close()
    elems_a = a.data.size
    elems_b = b.data.size
    nb_c, elems_c = dc.info()
    bytes_alg = 8.0 * (elems_a + elems_b + elems_c)
    ms_num = float(np.mean([s["ms_numeric"] for s in sts]))
    achieved_tf = flops / (ms_num * 1e-3) / 1e12
    variant, _ = D.tuning_info(ctx)
    kname = F64_KERNELS.get(variant, str(variant))
    rowblock = D.tuning_order(ctx)
    form, workgroups, _ = D.tuning_launch(ctx)
    result = {
        "metric": METRIC,
        "value": flops / (ms_step * 1e-3) / 1e9,
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded water-like atom box)",
        "config": {
            "workload": "plain FP64 C=A*B, 2048 H2O (6144 atoms) in a 39.4 A box, O 13 / H 5 "
                        "blocks (dim 47104), r<8 A, U[-1,1]exp(-0.4r), A!=B, eps_filter 1e-7",
            "atoms": int(atoms.n_atoms),
            "eps_filter": eps,
            "seed": args.seed,
            "flops_per_step": flops,
            "kept_triples": st["n_triples_kept"],
            "c_blocks": nb_c,
            "hbm_gb_per_s_step": bytes_alg / (ms_step * 1e-3) / 1e9,
            "hbm_gb_per_s_numeric": bytes_alg / (ms_num * 1e-3) / 1e9,
            "ms_numeric": ms_num,
            "ms_symbolic": float(np.mean([s["ms_symbolic"] for s in sts])),
            "ms_c_index": float(np.mean([s["ms_index"] for s in sts])),
            "tiles": st["n_tiles"],
            "tile_entries": st["n_tile_entries"],
            "ms_upload_ab": ms_upload,
            "operand_norms": "block norms of A and B computed once and cached across the "
                             "timed products (unchanged operands); ms_step_fresh_operands is "
                             "one product on fresh copies, norms included",
            "ms_step_fresh_operands": ms_fresh,
            "gflops_incl_upload": flops / ((ms_step + ms_upload) * 1e-3) / 1e9,
        },
        "roofline": {
            "bound": "mfma",
            "kernel": kname,
            "achieved": achieved_tf,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tf / PEAK_FP64_TFLOPS,
            "traffic": PROFILED_TRAFFIC_BYTES.get((kname, rowblock, form)),
            "tile_order_rowblock": rowblock,
            "launch_form": LAUNCH_FORMS.get(form, str(form)),
            "workgroups": workgroups,
            "algorithmic_bytes": bytes_alg,
        },
    }
    if not args.no_bf16:
        result["bf16_split"] = bf16_leg(D, da, db, dc, dd, eps, args, ms_step, ms_num)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(a, b, eps, args.cpu_rows)
    # Calibration for hardware byte counters: one device copy of A's data
    # pool (a plain read + write of a known byte count) in the same process.
    dcopy = D.DeviceMatrix(dd)
    dcopy.copy_from(da)
    print(f"copy calibration: {8 * a.data.size} bytes read and written", file=sys.stderr,
          flush=True)
    for h in (dcopy, dc, db, da, dd):
        h.close()
    ctx.close()
    result["library"] = library()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
