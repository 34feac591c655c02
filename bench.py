"""Benchmark: time to 1e-6 relative residual of the MSBG FLIP pressure solve
on one MI355X, learned-preconditioned vs plain MG-PCG, plus the smoother's HBM
roofline.

Workload (default, --config c5): the full-scale particle workload -- 1024^3
effective domain, 32^3 base blocks, 4 levels assigned by block distance to
the liquid (fine within one block); liquid = dam-break box x<0.4, y<0.6 plus
a sphere of radius 0.15 at (0.7,0.8,0.5); RHS = divergence of 8 seeded
particles per liquid cell (uniform positions, N(0,0.5) velocities) splatted
to the faces; FP16-stored smoother operands, FP32 Krylov operator, FP64 dots.
The learned coarse correction is trained in the run on pairs from plain
MG-PCG solves of seeded dam-break fields at 256^3 and validated on a held-out
seed.  `value` is the learned solve's time from p = 0 to |r| <= 1e-6 |b| with
the inputs resident in HBM; a step of the timed region (--steps) is one
learned MG-PCG iteration on that system.  --config c1/c2/c4 time the plain
MG-PCG on the smaller workloads instead.

Prints one JSON line (see the repository's benchmark contract).
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "msbg-flip_amd"))

HBM_PEAK_GBS = 8000.0
VAL_SEEDS = (900, 901)   # held-out validation scenes of the coarse form (neither training nor evaluation seeds)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c5", choices=["c1", "c2", "c4", "c5", "c5b16"],
                    help="c5b16: the full-scale particle workload on 16^3 base blocks")
    ap.add_argument("--n", type=int, default=0, help="override the domain size")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--n-pre", type=int, default=1, help="particle workload: pre-smoothing relaxations of the "
                    "V-cycle of both solves compared (the plain solve at the reference's 3/3 is reported beside)")
    ap.add_argument("--n-post", type=int, default=1, help="particle workload: post-smoothing relaxations")
    ap.add_argument("--max-iter", type=int, default=2000)
    ap.add_argument("--cpu-iters", type=int, default=-1,
                    help="PCG iterations in the CPU-baseline sample (-1: 0 for c5 -- the initial V-cycle "
                         "only --, else the whole solve to --tol)")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", action="store_true", help="also run the FP32/FP16/BF16 precision-form runs")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the FLIP caller, particle-to-grid, compact, PDE and batched sections")
    ap.add_argument("--no-fp64-oracle", action="store_true", help="skip the CPU FP64 reference solve")
    ap.add_argument("--p2g-n", type=int, default=1024, help="domain size of the particle-to-grid run")
    ap.add_argument("--batch-max-iter", type=int, default=400)
    ap.add_argument("--learned-pairs", type=int, default=2000)
    ap.add_argument("--learned-steps", type=int, default=2000)
    ap.add_argument("--learned-eval-n", type=int, default=512)
    ap.add_argument("--learned-val-n", type=int, default=1024, help="coarse form: size of the held-out validation "
                    "grid of the particle workload")
    ap.add_argument("--learned-eval-config", default="c4", choices=["c4", "c5", "c5b16"],
                    help="evaluation workload of the learned forms (c4: dam break at --learned-eval-n, seeds 501/502; "
                         "c5/c5b16: the 1024^3 particle workload on 32^3 / 16^3 blocks)")
    ap.add_argument("--learned-forms", default="", help="comma list of fine,coarse: also run the learned forms on "
                    "the dam-break evaluation workload (--learned-eval-config)")
    ap.add_argument("--scales", type=int, default=2, help="coarse form: scales of the network (1x, 2x, 4x block-mean grids)")
    ap.add_argument("--prolongation", type=int, default=1, help="coarse form: 0 constant, 1 trilinear")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch PCG iterations eagerly (no hipGraph) in every section (profiling runs)")
    ap.add_argument("--pmc-calibration", action="store_true",
                    help="only run a block copy of known size plus a few smoother passes (for rocprofv3 --pmc)")
    ap.add_argument("--literal", action="store_true", help="run the V-cycle residual/transfer work even when "
                    "every coarser level has no relaxation block")
    return ap.parse_args()


def make_cfg(msbgx, name, n):
    if name == "c1":
        return msbgx.config_c1(n or 256)
    if name == "c2":
        return msbgx.config_c2(n or 512, 32)
    if name == "c4":
        return msbgx.config_c4(n or 256)
    if name == "c5b16":
        return msbgx.config_c5(n or 1024, bsx=16, levels=4)
    return msbgx.config_c5(n or 1024)


def gpu_run(msbgx, cfg, args, log):
    t0 = time.time()
    g = msbgx.build_grid(cfg)
    opts = (msbgx.OPT_NO_GRAPH if args.no_graph else 0) | (msbgx.OPT_LITERAL_VCYCLE if args.literal else 0)
    if opts:
        g.set_options(opts)
    g.prepare_pressure_solve_flip(scene=cfg.scene)
    log(f"prepare {cfg.name} n={cfg.n}: {time.time() - t0:.1f} s")
    lv, fl = g.block_info(0)
    fluid = int(((fl & msbgx.BLK_NO_FLUID) == 0).sum())
    relax_blocks = len(g.relax_list(0))
    for i in range(args.warmup):
        g.zero_channel(msbgx.CH_PRESSURE, 0)
        it, rr = g.solve_pressure_flip(max_iter=args.max_iter, tol=args.tol)
        log(f"warmup {i}: iters {it} rel {rr:.3e} {g.last_solve_ms():.2f} ms")
    times, iters, rels = [], [], []
    wall0 = time.time()
    for i in range(args.steps):
        g.zero_channel(msbgx.CH_PRESSURE, 0)
        it, rr = g.solve_pressure_flip(max_iter=args.max_iter, tol=args.tol)
        times.append(g.last_solve_ms())
        iters.append(it)
        rels.append(rr)
        log(f"step {i}: iters {it} rel {rr:.3e} {times[-1]:.2f} ms")
    wall = time.time() - wall0
    # smoother roofline: one colour pass of the dominant block size, HIP events
    ms_relax, bytes_relax, cells_relax = g.time_kernel(0, reps=20)
    ms_res, bytes_res, _ = g.time_kernel(1, reps=20)
    ms_upd, bytes_upd, _ = g.time_kernel(2, reps=20)
    return dict(g=g, times=times, iters=iters, rels=rels, wall=wall, fluid_blocks=fluid, relax_blocks=relax_blocks,
                levels=[int(c) for c in __import__("numpy").bincount(lv, minlength=cfg.levels)],
                relax=(ms_relax, bytes_relax, cells_relax), residual=(ms_res, bytes_res), update=(ms_upd, bytes_upd))


def pmc_calibration(msbgx, cfg, log):
    """Workload for the rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs: one copy
    of the MG0 fluid blocks (k_copy_blocks: reads and writes 4 B per cell), ten
    smoother colour passes (both colours in turn) of the block size with the
    most relaxation cells and ten residual launches (k_multiply) over the
    level-0 fluid blocks; the particle workload with FP16 smoother operands."""
    g = msbgx.build_grid(cfg)
    g.prepare_pressure_solve_flip(scene=cfg.scene)
    if cfg.scene.particles_per_cell > 0 and cfg.bsx == 32:
        g.set_precision(msbgx.PREC_FP16)
    g.copy_channel(msbgx.CH_DIVERGENCE, msbgx.CH_R)
    ms, by, cells = g.time_kernel(0, reps=10)
    ms_res, by_res, _ = g.time_kernel(1, reps=10)
    lv, fl = g.block_info(0)
    copy_cells = int(sum((cfg.bsx >> int(l)) ** 3 for l, f in zip(lv, fl) if not f & msbgx.BLK_NO_FLUID))
    log(f"pmc calibration: copy of {copy_cells} cells ({copy_cells * 4} B read, {copy_cells * 4} B written); "
        f"smoother pass: {cells} cells, {by:.0f} algorithmic bytes, {ms:.4f} ms; "
        f"residual: {by_res:.0f} algorithmic bytes, {ms_res:.4f} ms")


def precision_runs(msbgx, g, cfg, args, log):
    """Precision form: FP16/BF16 smoother operands (FP32 compute, FP32 Krylov
    operator, FP64 dots) -- time to 1e-6 and smoother GB/s -- and FP64
    iterative refinement to 1e-10 for every variant, with the error against
    the CPU FP64 solution of the same level-0 system."""
    out = {}
    x_ref = None
    if not args.no_fp64_oracle and not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        o = oracle_lib.build_oracle(cfg, threads=min(args.cpu_threads, os.cpu_count() or 1))
        o.prepare_scene(cfg.scene)
        t0 = time.time()
        it, rel, x_ref = o.solve_fp64(tol=1e-12, max_iter=20000)
        log(f"fp64 oracle: {it} Jacobi-CG iterations, true rel residual {rel:.2e}, {time.time() - t0:.1f} s")
        den = sum(float((v ** 2).sum()) for v in x_ref.values()) ** 0.5
    for name, prec, bpc in (("fp32", msbgx.PREC_FP32, 30), ("fp16", msbgx.PREC_FP16, 22),
                            ("bf16", msbgx.PREC_BF16, 22)):
        g.set_precision(prec)
        times = []
        for i in range(2):
            g.zero_channel(msbgx.CH_PRESSURE, 0)
            it, rr = g.solve_pressure_flip(max_iter=args.max_iter, tol=args.tol)
            times.append(g.last_solve_ms())
        ms_k, _, cells = g.time_kernel(0, reps=20)
        t0 = time.time()
        outer, inner, rel64, xb = g.solve_pressure_refined(tol=1e-10, max_outer=30, inner_tol=1e-5,
                                                          want_x=x_ref is not None)
        t_ref = (time.time() - t0) * 1e3
        rec = {"solve_ms_to_1e-6": round(min(times), 3), "iterations": it, "rel_residual": rr,
               "smoother_GBs": round(bpc * cells / (ms_k * 1e-3) / 1e9, 1), "smoother_bytes_per_cell": bpc,
               "refined_to_1e-10": {"outer": outer, "inner_iterations": inner, "fp64_rel_residual": rel64,
                                    "wall_ms": round(t_ref, 1)}}
        if x_ref is not None:
            num = sum(float(((xb.get(b, 0.0) - v) ** 2).sum()) for b, v in x_ref.items()) ** 0.5
            g.zero_channel(msbgx.CH_PRESSURE, 0)
            g.solve_pressure_flip(max_iter=args.max_iter, tol=args.tol)
            num1 = 0.0
            for b, v in x_ref.items():
                a = g.get_block(msbgx.CH_PRESSURE, 0, b, cfg.bsx)
                num1 += float(((a.astype(float) - v) ** 2).sum())
            rec["error_vs_fp64_oracle"] = {"plain_to_1e-6": num1 ** 0.5 / den, "refined_to_1e-10": num / den}
        log(f"precision {name}: {rec}")
        out[name] = rec
    g.set_precision(msbgx.PREC_FP32)
    return out


def batched_runs(msbgx, args, log):
    """Batched form (C3 workload): 64 independent problems, sizes from
    {64,128,192,256}^3, 16^3 blocks, 2 MG levels, each solved to 1e-6 (at most
    --batch-max-iter iterations) -- one after the other on their own handles,
    then all together as one batched block set after the autotuner has picked
    the launch shape from 40-iteration trial solves."""
    t0 = time.time()
    cfgs = msbgx.config_c3(seed=1234, n_problems=64)
    mi = args.batch_max_iter
    singles = []
    for c in cfgs:
        g = msbgx.build_grid(c)
        g.prepare_pressure_solve_flip(scene=c.scene)
        singles.append(g)
    log(f"batched: prepared {len(cfgs)} problems in {time.time() - t0:.1f} s")
    seq_ms, seq_it = 0.0, []
    for k, g in enumerate(singles):
        g.solve_pressure_flip(max_iter=mi, tol=args.tol)   # warm-up (graph capture)
        g.zero_channel(msbgx.CH_PRESSURE, 0)
        it, _ = g.solve_pressure_flip(max_iter=mi, tol=args.tol)
        seq_ms += g.last_solve_ms()
        seq_it.append(it)
        if k % 8 == 7:
            log(f"batched: sequential solves {k + 1}/{len(singles)}, {seq_ms:.1f} ms so far")
    b = msbgx.BatchGrid.create(singles)
    for g in singles:
        g.destroy()
    chunk, mask = b.autotune(max_iter=40, tol=0.0)
    log(f"batched: autotune chunk {chunk} mask {mask}")
    b.zero_channel(msbgx.CH_PRESSURE, 0)
    it, rr = b.solve(max_iter=mi, tol=args.tol)
    ms = b.last_solve_ms()
    cells = sum(int(c.n) ** 3 for c in cfgs)
    conv = int((it < mi).sum())
    rec = {"problems": len(cfgs), "sizes": [c.n for c in cfgs], "effective_cells": cells,
           "max_iter": mi, "converged_problems": conv,
           "batched_ms": round(ms, 3), "sequential_ms": round(seq_ms, 3),
           "speedup_vs_sequential": round(seq_ms / ms, 2),
           "problems_per_s": round(len(cfgs) / (ms * 1e-3), 1),
           "iterations": [int(v) for v in it], "iterations_equal_to_individual": bool((it == seq_it).all()),
           "autotune": {"graph_chunk": chunk, "smoother_tile_mask": mask}}
    log(f"batched: {ms:.1f} ms batched vs {seq_ms:.1f} ms sequential, {conv}/{len(cfgs)} converged")
    b.destroy()
    return rec


def learned_train(msbgx, args, log, mode="coarse"):
    """Train a learned correction as the learned-preconditioner workload
    prescribes: pairs from plain MG-PCG solves of seeded dam-break fields at
    256^3 (box x<0.4, y<0.6 plus a sphere r~U[0.1,0.2], 16^3 blocks, 3
    levels), every iterate sampled until --learned-pairs iterates (fine form:
    `blocks_per_pair` random level-0 blocks each; coarse form: 16^3 tiles of
    the block-mean grids at each of the net's scales, random tile origins);
    Adam on the GPU.  Coarse form: after each schedule stage the net is
    validated on two held-out seeds of the particle workload at the
    evaluation size (--learned-val-n, VAL_SEEDS: neither training seeds nor
    the evaluation seed; FP16 operands) with the V-cycle of the timed solve
    (--n-pre/--n-post), and the weights with the fewest iterations summed
    over both are kept.
    The training pairs come from solves at the reference's V-cycle (3/3).
    Returns (net, record)."""
    t0 = time.time()
    net = msbgx.LearnedNet(seed=2024, mode=msbgx.NET_COARSE if mode == "coarse" else msbgx.NET_FINE)
    if mode == "coarse":
        net.scales = args.scales
        net.prolongation = args.prolongation
    # the coarse form's validation solves use the V-cycle of the solve it is timed in
    vcyc = {"n_pre": args.n_pre, "n_post": args.n_post} if mode == "coarse" else {}
    pairs, seed = 0, 0
    while pairs < args.learned_pairs:
        seed += 1
        c = msbgx.config_c4(256, seed=1000 + seed)
        g = msbgx.build_grid(c)
        g.prepare_pressure_solve_flip(scene=c.scene)
        pairs += net.collect(g, every=1, max_pairs=args.learned_pairs - pairs, blocks_per_pair=16, seed=seed,
                             tol=1e-7, max_iter=2000)
        g.destroy()
    t_collect = time.time() - t0
    log(f"learned ({mode}): {pairs} pairs from {seed} solves, {net.num_samples} samples, {t_collect:.1f} s")
    net_samples = net.num_samples
    if mode == "fine":
        sched = [(args.learned_steps, 3e-3, 64)]
    else:   # random tile origins make the coarse samples harder: a longer, decaying schedule
        sched = [(2000, 3e-3, 16), (2000, 1e-3, 16), (4000, 1e-3, 32), (4000, 3e-4, 32), (4000, 3e-4, 64),
                 (4000, 1e-4, 64)]
    train_ms, first, last = 0.0, None, None
    gvs = []
    if mode == "coarse":
        for vseed in VAL_SEEDS:
            vc = msbgx.config_c5(args.learned_val_n, seed=vseed)
            gv = msbgx.build_grid(vc)
            gv.prepare_pressure_solve_flip(scene=vc.scene)
            gv.set_precision(msbgx.PREC_FP16)
            gvs.append(gv)
    selection, cum = [], 0
    for k, (steps, lr, bs) in enumerate(sched):
        a, b = net.train(steps=steps, batch=bs, lr=lr, seed=7 + k)
        first = a if first is None else first
        last = b
        train_ms += net.last_train_ms()
        cum += steps
        if gvs:
            its, fbs = [], []
            for gv in gvs:
                gv.zero_channel(msbgx.CH_PRESSURE, 0)
                it_v, rr_v, fb_v = net.solve(gv, max_iter=args.max_iter, tol=args.tol, **vcyc)
                its.append(it_v)
                fbs.append(fb_v)
            selection.append({"steps": cum, "val_iterations": its, "val_fallback": fbs, "weights": net.weights})
            log(f"learned ({mode}): after {cum} steps, validation (particle workload {args.learned_val_n}^3, seeds "
                f"{VAL_SEEDS}): {its} iterations, fallback {fbs}")
    chosen = None
    if selection:
        best = min(range(len(selection)), key=lambda i: (sum(selection[i]["val_iterations"]), -selection[i]["steps"]))
        chosen = selection[best]["steps"]
        net.weights = selection[best]["weights"]
        for gv in gvs:
            gv.destroy()
        log(f"learned ({mode}): keeping the weights after {chosen} steps")
    steps_total = sum(st for st, _, _ in sched)
    log(f"learned ({mode}): {steps_total} Adam steps, loss {first:.4f} -> {last:.4f}, {train_ms:.0f} ms")
    rec = {"form": mode, "workload": "dam break 256^3, 16^3 blocks, 3 levels", "pairs": pairs, "solves": seed,
           "samples": net_samples, "collect_s": round(t_collect, 1), "steps": steps_total,
           "schedule": [{"steps": st, "lr": lr, "batch": bs} for st, lr, bs in sched],
           "loss_first": first, "loss_last": last, "train_gpu_ms": round(train_ms, 1),
           "validation": [{k: v for k, v in e.items() if k != "weights"} for e in selection],
           "kept_weights_after_steps": chosen, "wall_s": round(time.time() - t0, 1)}
    if mode == "coarse":
        rec["pairs_v_cycle"] = "V(3,3) (the reference's nPre/nPost)"
        rec["validation_grids"] = (f"particle workload {args.learned_val_n}^3, seeds {VAL_SEEDS}, "
                                   f"V({args.n_pre},{args.n_post}); the fewest iterations summed over them")
        rec["scales"] = net.scales
        rec["prolongation"] = "trilinear" if net.prolongation == msbgx.PROLONG_LINEAR else "constant"
    return net, rec


def learned_runs(msbgx, args, log, mode="fine"):
    """Learned form on the dam-break evaluation workload: the net of
    learned_train, then plain vs learned MG-PCG to 1e-6 on unseen seeds at
    --learned-eval-n (or the particle workload)."""
    net, train = learned_train(msbgx, args, log, mode)
    evals = []
    if mode == "fine" and args.learned_eval_config == "c5":
        log("learned (fine): the 32^3-block workload needs the coarse form; skipped")
        eval_cfgs = []
    elif args.learned_eval_config == "c4":
        eval_cfgs = [msbgx.config_c4(args.learned_eval_n, seed=s) for s in (501, 502)]
    else:
        eval_cfgs = [make_cfg(msbgx, args.learned_eval_config, 0)]
    for c in eval_cfgs:
        s = c.scene.seed
        g = msbgx.build_grid(c)
        g.prepare_pressure_solve_flip(scene=c.scene)
        warm = args.learned_eval_config == "c4"   # full-scale solves run long enough without a warm-up
        if warm:
            g.solve_pressure_flip(max_iter=args.max_iter, tol=args.tol)
            g.zero_channel(msbgx.CH_PRESSURE, 0)
        log(f"learned ({mode}): plain solve of {c.name} {c.n}^3")
        it_p, rr_p = g.solve_pressure_flip(max_iter=args.max_iter, tol=args.tol)
        ms_p = g.last_solve_ms()
        g.zero_channel(msbgx.CH_PRESSURE, 0)
        if warm:
            net.solve(g, max_iter=args.max_iter, tol=args.tol)
            g.zero_channel(msbgx.CH_PRESSURE, 0)
        log(f"learned ({mode}): learned solve of {c.name} {c.n}^3")
        it_l, rr_l, fb = net.solve(g, max_iter=args.max_iter, tol=args.tol)
        ms_l = g.last_solve_ms()
        rec = {"workload": c.name, "bsx": c.bsx, "seed": s, "n": c.n, "plain_ms": round(ms_p, 2), "plain_iterations": it_p, "plain_rel": rr_p,
               "learned_ms": round(ms_l, 2), "learned_iterations": it_l, "learned_rel": rr_l,
               "fallback_iteration": fb, "speedup": round(ms_p / ms_l, 3)}
        log(f"learned ({mode}): {rec}")
        evals.append(rec)
        g.destroy()
    ms_c, fl_c = net.time_conv(blocks=2048, reps=20)
    conv = {"kernel": "k_nn_conv16<false> (16->16, v_mfma_f32_16x16x32_bf16)", "blocks": 2048,
            "flop_per_launch": fl_c, "flop_counted": "16^3 cells x 27 taps x 16 x 16 channels x 2 per block",
            "launch_ms": round(ms_c, 4), "achieved": round(fl_c / (ms_c * 1e-3) / 1e12, 1),
            "peak": 2500.0, "unit": "TFLOP/s", "frac": round(fl_c / (ms_c * 1e-3) / 1e12 / 2500.0, 4)}
    log(f"learned: conv16 roofline {conv}")
    w = net.weights
    net.destroy()
    return {"form": mode, "train": train, "eval": evals, "conv16_roofline": conv,
            "weights_l2": float((w.astype(float) ** 2).sum() ** 0.5)}


def read_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 --pmc summary (profiles/hbm_traffic.txt), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) >= 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel and \
                        parts[2] == workload:
                    return float(parts[3])
    except OSError:
        pass
    return None


def dense_level0(msbgx, g, ch, n, bsx):
    """Channel `ch` of the level-0 blocks as a dense [z][y][x] array (0 elsewhere)."""
    import numpy as np
    lv, _ = g.block_info(0)
    out = np.zeros((n, n, n), np.uint16 if ch == msbgx.CH_FLAGS else np.float32)
    nbx = n // bsx
    for b in range(len(lv)):
        if lv[b] != 0:
            continue
        bx, by, bz = b % nbx, (b // nbx) % nbx, b // (nbx * nbx)
        out[bz * bsx:(bz + 1) * bsx, by * bsx:(by + 1) * bsx, bx * bsx:(bx + 1) * bsx] = \
            g.get_block(ch, 0, b, bsx).reshape(bsx, bsx, bsx)
    return out, lv


def flip_runs(msbgx, g, cfg, args, log):
    """FLIP caller steps on the device around the solve: divergence right-hand
    side from the scene's MAC velocity, plain MG-PCG to the tolerance, the
    pressure-gradient update (corrected sign) and the divergence left after it;
    HBM roofline of both kernels (HIP events); the CPU oracle's update on the
    same arrays (1 thread)."""
    import numpy as np
    dt = cfg.scene.dt
    g.set_precision(msbgx.PREC_FP32)
    g.flip_set_velocity_scene(cfg.scene)
    g.flip_rhs_from_velocity(dt)
    blocks = g.level0_blocks()
    b0 = np.concatenate([g.get_block(msbgx.CH_DIVERGENCE, 0, int(b), cfg.bsx) for b in blocks]).astype(np.float64)
    g.zero_channel(msbgx.CH_PRESSURE, 0)
    it, rr = g.solve_pressure_flip(max_iter=args.max_iter, tol=args.tol)
    vel = g.flip_get_velocity()
    g.flip_update_velocity(dt, corrected=True)
    g.flip_rhs_from_velocity(dt)
    b1 = np.concatenate([g.get_block(msbgx.CH_DIVERGENCE, 0, int(b), cfg.bsx) for b in blocks]).astype(np.float64)
    red = float(np.linalg.norm(b1) / max(np.linalg.norm(b0), 1e-300))
    log(f"flip: solve {it} iterations rel {rr:.2e}; divergence after the corrected update {red:.2e} of before")
    out = {"solve_iterations": it, "solve_rel_residual": rr, "divergence_after_update": red}
    for which, name in ((3, "velocity_update"), (4, "rhs_from_velocity")):
        ms, by, cells = g.time_kernel(which, reps=20)
        gbs = by / (ms * 1e-3) / 1e9
        out[name] = {"launch_ms": round(ms, 4), "cells": cells, "algorithmic_bytes": int(by),
                     "bytes_per_cell": round(by / max(cells, 1), 2), "GBs": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "Gcells_per_s": round(cells / (ms * 1e-3) / 1e9, 1)}
        log(f"flip {name}: {ms:.4f} ms, {cells} cells, {gbs:.0f} GB/s")
    if not args.no_cpu:
        try:
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            import oracle_lib
            flags, lv = dense_level0(msbgx, g, msbgx.CH_FLAGS, cfg.n, cfg.bsx)
            mass = ((flags & (msbgx.CELL_SOLID | msbgx.CELL_VOID)) == 0).astype(np.float32)
            p, _ = dense_level0(msbgx, g, msbgx.CH_PRESSURE, cfg.n, cfg.bsx)
            beta = [dense_level0(msbgx, g, ch, cfg.n, cfg.bsx)[0]
                    for ch in (msbgx.CH_FACE_U, msbgx.CH_FACE_V, msbgx.CH_FACE_W)]
            t0 = time.time()
            oracle_lib.flip_update_velocity(lv, cfg.bsx, mass, 0.5, p, beta, vel, dt, negate_p=True)
            t_cpu = time.time() - t0
            out["cpu_oracle_update_s"] = round(t_cpu, 3)
            out["cpu_oracle_update_threads"] = 1
            log(f"flip cpu oracle update: {t_cpu:.2f} s (1 thread)")
        except Exception as e:
            log(f"flip cpu oracle failed: {e}")
    return out


def compact_runs(msbgx, g, cfg, args, log):
    """The FLIP caller's compact MIC(0)-PCG (the reference's default solver
    kind) on the GPU as a comparator on the same system, to the same
    tolerance, and the oracle's sequential MIC(0)-PCG on a bounded sample
    (10 iterations, scaled to the GPU's count)."""
    import numpy as np
    g.set_precision(msbgx.PREC_FP32)
    g.prepare_pressure_solve_flip(scene=cfg.scene)      # fresh right-hand side and channels
    its, times = [], []
    for _ in range(2):
        it, rr = g.solve_pressure_compact(max_iter=args.max_iter, tol=args.tol)
        its.append(it)
        times.append(g.last_solve_ms())
    out = {"solver": "MIC(0)-PCG (compact, hyperplane-wavefront triangular solves, one hipGraph per iteration)",
           "iterations": its[-1], "rel_residual": rr, "gpu_ms": round(min(times), 2),
           "ms_per_iteration": round(min(times) / max(its[-1], 1), 3)}
    log(f"compact MIC(0)-PCG: {its[-1]} iterations, rel {rr:.2e}, {min(times):.1f} ms")
    if not args.no_cpu:
        try:
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            import oracle_lib
            flags, lv = dense_level0(msbgx, g, msbgx.CH_FLAGS, cfg.n, cfg.bsx)
            mass = ((flags & (msbgx.CELL_SOLID | msbgx.CELL_VOID)) == 0).astype(np.float32)
            beta = [dense_level0(msbgx, g, ch, cfg.n, cfg.bsx)[0]
                    for ch in (msbgx.CH_FACE_U, msbgx.CH_FACE_V, msbgx.CH_FACE_W)]
            div, _ = dense_level0(msbgx, g, msbgx.CH_DIVERGENCE, cfg.n, cfg.bsx)
            k = 10
            t0 = time.time()
            _, ito, _, _, _ = oracle_lib.compact_pcg(lv, cfg.bsx, mass, 0.5, beta, div, tol=args.tol, max_iter=k)
            t = time.time() - t0
            out["cpu_oracle"] = {"sample_iterations": ito, "sample_s": round(t, 3), "threads": 1,
                                 "estimated_s_for_gpu_iterations": round(t / max(ito, 1) * its[-1], 2)}
            log(f"compact cpu oracle: {ito} iterations in {t:.2f} s (1 thread)")
        except Exception as e:
            log(f"compact cpu oracle failed: {e}")
    return out


def pde_runs(msbgx, args, log):
    """msbg_demo's mean-curvature smoothing (5 iterations, rounded up to 6, 8
    colours, time step 0.05) of a 256^3 render density (a sphere of radius
    0.3 with a tanh profile) over the blocks that hold its interface; the
    oracle on the same input (1 thread)."""
    import numpy as np
    n, bsx = 256, 16
    g = msbgx.MultiresSparseGrid.create(n, n, n, bsx, 3)
    x = (np.arange(n, dtype=np.float32) + 0.5) / n
    r = np.sqrt((x[None, None, :] - 0.5) ** 2 + (x[None, :, None] - 0.5) ** 2 + (x[:, None, None] - 0.45) ** 2)
    rho = np.round(np.clip(0.5 - 0.5 * np.tanh((r - 0.3) * n / 2.5), 0, 1) * 65535).astype(np.uint16)
    nb = n // bsx
    blk = rho.reshape(nb, bsx, nb, bsx, nb, bsx)
    interface = ((blk > 0) & (blk < 65535)).any(axis=(1, 3, 5))        # [bz][by][bx]
    active = np.nonzero(interface.reshape(-1))[0].astype(np.int32)      # bid = bx + nb (by + nb bz)
    g.pde_mean_curvature(rho, active, 5, 0.05)                          # warm-up
    out, ms = g.pde_mean_curvature(rho, active, 5, 0.05)
    vox = len(active) * bsx ** 3
    rec = {"workload": "sphere r=0.3 render density 256^3 (uint16), 16^3 blocks, interface blocks listed",
           "active_blocks": int(len(active)), "iterations": 6, "gpu_ms": round(ms, 3),
           "Gvoxel_updates_per_s": round(vox * 6 / (ms * 1e-3) / 1e9, 2),
           "GBs_algorithmic": round(vox * 6 * 4 / (ms * 1e-3) / 1e9, 1)}
    log(f"pde: {len(active)} blocks, 6 iterations, {ms:.2f} ms")
    if not args.no_cpu:
        try:
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            import oracle_lib
            t0 = time.time()
            ref = oracle_lib.pde_mean_curvature(np.zeros(nb ** 3, np.int32), bsx, rho, active, 5, 0.05)
            rec["cpu_oracle_s"] = round(time.time() - t0, 3)
            rec["cpu_oracle_threads"] = 1
            rec["bit_identical_to_oracle"] = bool(np.array_equal(ref, out))
            log(f"pde cpu oracle: {rec['cpu_oracle_s']:.2f} s (1 thread), identical: {rec['bit_identical_to_oracle']}")
        except Exception as e:
            log(f"pde cpu oracle failed: {e}")
    g.destroy()
    return rec


def p2g_runs(msbgx, args, log):
    """Particle-to-grid transfer of the full-scale particle scene (dam break +
    sphere, 8 particles per liquid cell) on the GPU, and GPU vs the CPU oracle
    on the same scene at 256^3 (the oracle is serial)."""
    out = {}
    for n in sorted({args.p2g_n, 256}, reverse=True):
        cfg = msbgx.config_c5(n, bsx=16, levels=4)
        g = msbgx.build_grid(cfg)
        t = msbgx.ParticleToGrid(g)
        t0 = time.time()
        npart = t.run_scene(cfg.scene)
        wall = time.time() - t0
        t.run_scene(cfg.scene)                   # second run: steady state
        ms, msg, _ = t.last_times()
        act = len(t.active_blocks())
        rec = {"workload": f"dam break + sphere {n}^3, 8 particles per liquid cell, 16^3 blocks",
               "particles": npart, "gpu_ms": round(ms, 3), "gather_ms": round(msg, 3),
               "binning_ms": round(ms - msg, 3), "Gparticles_per_s": round(npart / (ms * 1e-3) / 1e9, 3),
               "active_blocks": act, "first_call_wall_s": round(wall, 2)}
        log(f"p2g {n}^3: {npart} particles, {ms:.2f} ms ({msg:.2f} ms gather), {act} active blocks")
        if n == 256 and not args.no_cpu:
            try:
                sys.path.insert(0, os.path.join(ROOT, "tests"))
                import oracle_lib
                lv, _ = g.block_info(0)
                hp = oracle_lib.scene_particles(cfg.scene)
                t0 = time.time()
                oracle_lib.p2g(lv, n, n, n, cfg.bsx, hp)
                rec["cpu_oracle_s"] = round(time.time() - t0, 3)
                rec["cpu_oracle_threads"] = 1
                log(f"p2g cpu oracle {n}^3: {rec['cpu_oracle_s']:.2f} s (1 thread)")
            except Exception as e:
                log(f"p2g cpu oracle failed: {e}")
        out[f"n{n}"] = rec
        del t, g
    return out


def cpu_baseline(cfg, args, gpu_iters, log):
    """CPU oracle (C++ restatement of the reference solver) timed on the host
    cores on a bounded sample of the same solve: `cpu_iters` PCG iterations
    (0: the initial V-cycle alone, which is one iteration's work), scaled to
    the GPU's plain iteration count.  For the particle workload the oracle is
    prepared from the same scene with the velocity-divergence right-hand side
    instead of the particle splat: topology, level map, cell flags and face
    weights are those of the workload, and the work per iteration does not
    depend on the right-hand side's values, while the oracle's serial
    particle transfer would take tens of minutes at 1024^3."""
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    outdir = tempfile.mkdtemp(prefix="msbg_oracle_")
    so = os.path.join(outdir, "libmsbg_oracle.so")
    try:
        subprocess.run(["g++", "-O3", "-march=native", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
                        "-std=c++17", "-shared", "-o", so, os.path.join(ROOT, "oracle", "msbg_oracle.cpp"),
                        os.path.join(ROOT, "oracle", "flip_oracle.cpp")],
                       check=True, capture_output=True)
        os.environ["MSBG_ORACLE_SO"] = so
    except Exception as e:  # fall back to the prebuilt portable build
        log(f"native oracle build failed ({e}); using oracle/libmsbg_oracle.so")
    import oracle_lib
    threads = min(args.cpu_threads, os.cpu_count() or 1)
    particles = cfg.scene.particles_per_cell > 0
    scene = type(cfg.scene).from_buffer_copy(cfg.scene)
    if particles:
        scene.particles_per_cell = 0
    big = cfg.n >= 512
    o = oracle_lib.build_oracle(cfg, threads=threads)
    if args.config == "c5":
        o.set_smoother_precision(1)   # FP16 operands, as the GPU solve
    t0 = time.time()
    o.prepare_scene(scene)
    t_prep = time.time() - t0
    k = args.cpu_iters if args.cpu_iters >= 0 else (0 if big else max(1, gpu_iters))
    vc = {"n_pre": args.n_pre, "n_post": args.n_post} if args.config == "c5" else {}   # the GPU plain solve's V-cycle
    t0 = time.time()
    it, rr, _ = o.solve(max_iter=k, tol=args.tol, **vc)
    t_k = time.time() - t0
    del o
    # the solve runs one preconditioner application before the loop: count it as one iteration
    per_iter = t_k / (k + 1)
    whole = it == gpu_iters and k > 0
    est = t_k if whole else per_iter * (gpu_iters + 1)
    log(f"cpu oracle: prepare {t_prep:.1f} s, {it} iterations (+1 initial V-cycle) {t_k:.1f} s on {threads} threads, "
        f"{per_iter*1e3:.1f} ms/iter; for the GPU's {gpu_iters} iterations: {est:.1f} s")
    t1 = None
    if not big:
        o1 = oracle_lib.build_oracle(cfg, threads=1)
        o1.prepare_scene(scene)
        t0 = time.time()
        o1.solve(max_iter=1, tol=args.tol)
        t1 = (time.time() - t0) / 2
        log(f"cpu oracle, 1 thread: {t1*1e3:.0f} ms per iteration")
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    rhs = ("; right-hand side: velocity divergence in place of the particle splat (same topology, levels, flags "
           "and weights)" if particles else "")
    return {"value": round(est * 1e3, 1), "unit": "ms", "cores": threads, "kind": "port",
            "one_thread_ms_per_iteration": round(t1 * 1e3, 1) if t1 is not None else None,
            "sample": (f"oracle/msbg_oracle.cpp (C++ restatement of the reference solver, -O3 -march=native, OpenMP "
                       f"over the blocks of a colour) on {threads} threads of {model}, {cfg.name} {cfg.n}^3"
                       + (f", FP16 smoother operands, V({args.n_pre},{args.n_post})" if args.config == "c5" else "")
                       + ": "
                       + (f"the whole solve to {args.tol:g} ({it} PCG iterations, the GPU's count) timed"
                          if whole else
                          f"{it} PCG iterations + the initial V-cycle timed = {t_k:.1f} s, i.e. {per_iter:.2f} s per "
                          f"iteration, scaled to the GPU plain solve's {gpu_iters} iterations")
                       + f"; setup {t_prep:.1f} s excluded" + rhs)}


def c5_headline(msbgx, args, log):
    """The headline: plain vs learned MG-PCG to --tol on the particle workload
    (FP16 smoother operands), both with the V-cycle V(--n-pre, --n-post) so
    that the ratio is the network's own gain; the plain MG-PCG with the
    reference's V(3,3) is timed beside them.  Then the timed region of
    --steps learned iterations and the smoother's roofline on the same
    grid."""
    import numpy as np
    cfg = make_cfg(msbgx, "c5", args.n)
    t0 = time.time()
    g = msbgx.build_grid(cfg)
    if args.no_graph:
        g.set_options(msbgx.OPT_NO_GRAPH)
    g.prepare_pressure_solve_flip(scene=cfg.scene)
    g.set_precision(msbgx.PREC_FP16)
    lv, fl = g.block_info(0)
    nb0 = len(g.level0_blocks())
    log(f"prepare {cfg.name} n={cfg.n}: {time.time() - t0:.1f} s, {nb0} PCG blocks "
        f"({nb0 * cfg.bsx ** 3} cells), {len(g.relax_list(0))} relaxation blocks at MG0")
    net, train = learned_train(msbgx, args, log, "coarse")
    vc = {"n_pre": args.n_pre, "n_post": args.n_post}
    vname = f"V({args.n_pre},{args.n_post})"
    # plain MG-PCG: graph capture on a short solve, then the whole solve; the
    # compared V-cycle, then the reference's V(3,3)
    plain = {}
    for name, kw in ((vname, vc), ("V(3,3)", {"n_pre": 3, "n_post": 3})):
        if name in plain:
            continue
        g.solve_pressure_flip(max_iter=3, tol=0.0, **kw)
        g.zero_channel(msbgx.CH_PRESSURE, 0)
        it, rr = g.solve_pressure_flip(max_iter=args.max_iter, tol=args.tol, **kw)
        ms = g.last_solve_ms()
        plain[name] = (it, rr, ms)
        log(f"plain MG-PCG {name}: {it} iterations, rel {rr:.3e}, {ms:.1f} ms ({ms / max(it, 1):.2f} ms/iteration)")
    it_p, rr_p, ms_p = plain[vname]
    it_r, rr_r, ms_r = plain["V(3,3)"]
    # timed region: --warmup untimed, then exactly --steps learned MG-PCG iterations
    g.zero_channel(msbgx.CH_PRESSURE, 0)
    net.solve(g, max_iter=max(args.warmup, 1), tol=0.0, **vc)
    g.zero_channel(msbgx.CH_PRESSURE, 0)
    w0 = time.perf_counter()
    k_it, _, _ = net.solve(g, max_iter=args.steps, tol=0.0, **vc)   # returns after the device is idle
    wall_ms = (time.perf_counter() - w0) * 1e3
    steps_ms = g.last_solve_ms()
    log(f"timed: {k_it} learned iterations in {steps_ms:.1f} ms (HIP events), {wall_ms:.1f} ms wall")
    # the learned solve to the tolerance
    g.zero_channel(msbgx.CH_PRESSURE, 0)
    it_l, rr_l, fb = net.solve(g, max_iter=args.max_iter, tol=args.tol, **vc)
    ms_l = g.last_solve_ms()
    log(f"learned MG-PCG {vname}: {it_l} iterations, rel {rr_l:.3e}, fallback {fb}, {ms_l:.1f} ms; "
        f"speedup {ms_p / ms_l:.2f}x over plain {vname}, {ms_r / ms_l:.2f}x over plain V(3,3)")
    ms_relax, bytes_relax, cells_relax = g.time_kernel(0, reps=20)
    ms_res, bytes_res, _ = g.time_kernel(1, reps=20)
    ms_upd, bytes_upd, _ = g.time_kernel(2, reps=20)
    achieved = bytes_relax / (ms_relax * 1e-3) / 1e9
    log(f"smoother: {ms_relax:.4f} ms per colour pass, {cells_relax} cells, {achieved:.0f} GB/s")
    ms_c, fl_c = net.time_conv(blocks=2048, reps=20)
    conv = {"kernel": "k_nn_conv16<false> (16->16 forward layer, v_mfma_f32_16x16x32_bf16)", "blocks": 2048,
            "launch_ms": round(ms_c, 4), "flop_per_launch": fl_c,
            "flop_counted": "16^3 cells x 27 taps x 16 x 16 channels x 2 per block",
            "achieved": round(fl_c / (ms_c * 1e-3) / 1e12, 1), "peak": 2500.0, "unit": "TFLOP/s",
            "frac": round(fl_c / (ms_c * 1e-3) / 1e12 / 2500.0, 4)}
    log(f"learned: conv16 {conv}")
    ws = net.weights
    net.destroy()
    rec = {
        "metric": "time to 1e-6 rel. residual (learned vs plain MG-PCG, 1 GPU); smoother HBM GB/s",
        "value": round(ms_l, 1),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(steps_ms / max(k_it, 1), 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP16-stored smoother operands, FP64 dots)",
        "data": "synthetic (seeded scene; particle-to-grid velocity divergence, 8 particles per liquid cell)",
        "config": {
            "workload": f"{cfg.name} {cfg.n}^3, {cfg.bsx}^3 base blocks, {cfg.levels} levels, FP16 smoother "
                        f"operands; learned coarse correction ({train['scales']} scales, {train['prolongation']} "
                        f"prolongation) vs plain MG-PCG, both with the V-cycle {vname}, to {args.tol:g}",
            "v_cycle": {"n_pre": args.n_pre, "n_post": args.n_post, "n_coarse": 24},
            "step": "one learned MG-PCG iteration on this system (value: the whole learned solve)",
            "blocks_per_level": [int(c) for c in np.bincount(lv, minlength=cfg.levels)],
            "pcg_blocks": nb0,
            "pcg_cells": nb0 * cfg.bsx ** 3,
            "relax_blocks_mg0": len(g.relax_list(0)),
            "tol": args.tol,
            "parallelism": "1 GPU",
        },
        "learned_vs_plain": {
            "v_cycle": vname,
            "plain_ms": round(ms_p, 1), "plain_iterations": it_p, "plain_rel_residual": rr_p,
            "plain_ms_per_iteration": round(ms_p / max(it_p, 1), 3),
            "plain_reference_v_cycle": {"v_cycle": "V(3,3) (FlipPressureSolveParams defaults, src/msbg.h:837-844)",
                                        "ms": round(ms_r, 1), "iterations": it_r, "rel_residual": rr_r,
                                        "learned_speedup": round(ms_r / ms_l, 3)},
            "learned_ms": round(ms_l, 1), "learned_iterations": it_l, "learned_rel_residual": rr_l,
            "learned_fallback_iteration": fb, "learned_ms_per_iteration": round(steps_ms / max(k_it, 1), 3),
            "speedup": round(ms_p / ms_l, 3), "timed_region_wall_ms": round(wall_ms, 1),
            "train": train, "weights_l2": float((ws.astype(float) ** 2).sum() ** 0.5),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_relax_c32h<FP16> (one red-black colour pass over 32^3 blocks: 4 local sweeps + Jacobi)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": read_traffic("k_relax_c32h<1>", cfg.name),
            "launch_ms": round(ms_relax, 4),
            "cells_per_launch": cells_relax,
            "bytes_per_cell": 22,
        },
        "other_kernels": {
            "residual_k_multiply_z32_GBs": round(bytes_res / (ms_res * 1e-3) / 1e9, 1),
            "pcg_update_GBs": round(bytes_upd / (ms_upd * 1e-3) / 1e9, 1),
            "mfma_conv16": conv,
        },
    }
    return rec, g, it_p


def plain_line(msbgx, cfg, args, log):
    """--config c1/c2/c4/c5b16: plain MG-PCG time to --tol (FP32 storage)."""
    r = gpu_run(msbgx, cfg, args, log)
    ms = sorted(r["times"])[len(r["times"]) // 2]
    ms_relax, bytes_relax, cells_relax = r["relax"]
    achieved = bytes_relax / (ms_relax * 1e-3) / 1e9
    line = {
        "metric": "time to 1e-6 rel. residual (plain MG-PCG, 1 GPU); smoother HBM GB/s",
        "value": round(ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sum(r["times"]) / len(r["times"]), 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 dots)",
        "data": "synthetic (seeded scene, uniform [-1,1] MAC velocity divergence)",
        "config": {
            "workload": f"{cfg.name} {cfg.n}^3, {cfg.bsx}^3 base blocks, {cfg.levels} levels, FP32 storage",
            "blocks_per_level": r["levels"],
            "fluid_blocks_mg0": r["fluid_blocks"],
            "relax_blocks_mg0": r["relax_blocks"],
            "iterations": r["iters"][0],
            "final_rel_residual": r["rels"][0],
            "tol": args.tol,
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_relax_col<{cfg.bsx}> (one red-black colour pass, 5 local sweeps)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": read_traffic(f"k_relax_col<{cfg.bsx}>", cfg.name),
            "launch_ms": round(ms_relax, 4),
            "cells_per_launch": cells_relax,
            "bytes_per_cell": 30,
        },
        "other_kernels": {
            "residual_GBs": round(r["residual"][1] / (r["residual"][0] * 1e-3) / 1e9, 1),
            "pcg_update_GBs": round(r["update"][1] / (r["update"][0] * 1e-3) / 1e9, 1),
        },
    }
    return line, r["g"], r["iters"][0]


def main():
    args = parse()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on exactly one GPU; multi-GPU runs are not part of it "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if args.no_graph:
        os.environ["MSBGX_NO_GRAPH"] = "1"   # every handle created below launches eagerly
    import msbgx
    cfg = make_cfg(msbgx, args.config, args.n)
    if args.pmc_calibration:
        pmc_calibration(msbgx, cfg, log)
        return
    if args.config == "c5":
        line, g, gpu_iters = c5_headline(msbgx, args, log)
    else:
        line, g, gpu_iters = plain_line(msbgx, cfg, args, log)
    if args.precision:
        line["precision_form"] = precision_runs(msbgx, g, cfg, args, log)
    g.destroy()
    if not args.no_extras:
        c1 = msbgx.config_c1(256)
        g1 = msbgx.build_grid(c1)
        g1.prepare_pressure_solve_flip(scene=c1.scene)
        line["flip_caller"] = flip_runs(msbgx, g1, c1, args, log)
        line["particle_to_grid"] = p2g_runs(msbgx, args, log)
        line["compact_mic0_pcg"] = compact_runs(msbgx, g1, c1, args, log)
        line["pde_demo"] = pde_runs(msbgx, args, log)
        line["batched_form"] = batched_runs(msbgx, args, log)
        g1.destroy()
    if args.learned_forms:
        for form in args.learned_forms.split(","):
            line[f"learned_{form}_form_dambreak"] = learned_runs(msbgx, args, log, form)
    if not args.no_cpu:
        try:
            line["cpu_baseline"] = cpu_baseline(cfg, args, gpu_iters, log)
        except Exception as e:
            log(f"cpu baseline failed: {e}")
            line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
