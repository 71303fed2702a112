#!/usr/bin/env python3
"""Benchmark: lozenge q-volume chain on the a = b = c = 1000 hexagon, one MI355X.

Workload: LozengeTiler::AlmostHexDomain(1000, 0) (tiling side N = 2007, 1,342,683
generators, three 2007 x 669 uint8 colour classes), started from the empty room (the
minimal tiling), the reference's q-rule (qlozengekernel.cl) with q = 0.999, the
reference's colour schedule (minstd_rand0 + uniform_int_distribution(0, 2)), TinyMT32 with
the default parameter set per generator.  Synthetic: no data is read.

One bench "step" = one rt_chain_run of --chain-steps chain steps (each chain step flips one
colour class: 1,342,683 flip attempts, padding included as in the reference).  The timed
region is bracketed by stream synchronizations; `value` is flip attempts per second of
wall-clock time with the state resident in HBM.  The roofline block uses HIP events recorded
on the library's own stream around the same region.

    python bench.py [--steps K] [--warmup W] [--chain-steps S] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "flip attempts/sec + HBM GB/s fraction; sweeps-to-equilibrium warm vs plain"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--chain-steps", type=int, default=8192)
    ap.add_argument("--hex", type=int, default=1000)
    ap.add_argument("--q", type=float, default=0.999)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--cpu-steps", type=int, default=400)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile-config", type=int, default=-1)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (to scale HBM byte counters)")
    ap.add_argument("--workload", default="c2",
                    choices=["c2", "c1", "c3", "c4", "c5", "tridimer", "closure", "gemm"],
                    help="c2 (default): lozenge hexagon 1000; c1: domino Aztec diamond 128, 10^5 steps; "
                         "c4: coupling from the past on an Aztec diamond (--aztec); tridimer: triangle-dimer cone")
    ap.add_argument("--aztec", type=int, default=256, help="Aztec diamond order of the c4 workload")
    ap.add_argument("--p12", nargs=4, default=None, help="c4: weighted heat-bath probabilities (else p = 1/2)")
    ap.add_argument("--pair-steps", type=int, default=0,
                    help="c4: only time this many coupled steps at --aztec (no coalescence)")
    ap.add_argument("--placement-n", type=int, default=16, help="c4: Aztec order of the placement check")
    ap.add_argument("--placement-samples", type=int, default=200, help="c4: exact samples for the placement check")
    ap.add_argument("--cone", type=int, default=1000, help="cone side of the tridimer workload")
    ap.add_argument("--check", action="store_true", help="c1: also compare the final tiling with the oracle")
    ap.add_argument("--domains", type=int, default=64, help="c3: number of domains")
    ap.add_argument("--sides", type=int, nargs=2, default=[64, 512], help="c3: domain side range")
    ap.add_argument("--domain-steps", type=int, default=1000, help="c3: chain steps per domain")
    ap.add_argument("--autotune", action="store_true", help="c3: measure the batch's tile configurations first")
    ap.add_argument("--batch-configs", type=int, nargs=4, default=None, metavar=("DOM", "LOZ", "RT", "TD"),
                    help="c3: tile configuration per model (domino, lozenge, rect-triangle, triangle-dimer; -1 "
                         "automatic), e.g. what an earlier --autotune chose")
    ap.add_argument("--no-solo", action="store_true",
                    help="c3: skip the one-run-per-domain pass (and the comparison with it): only the batched run")
    ap.add_argument("--kinds", nargs="+", default=["aztec", "notched", "tridimer", "recttri"],
                    help="c3: domain kinds of the mix (aztec, notched, tridimer, recttri; also lozenge: "
                         "almost-hexagons under lozengekernel.cl's rule)")
    ap.add_argument("--canvas", type=int, default=64, help="c5: largest vertex grid side (training canvas)")
    ap.add_argument("--train-domains", type=int, default=96, help="c5: training domains")
    ap.add_argument("--samples-per-domain", type=int, default=16, help="c5: chain samples averaged per domain")
    ap.add_argument("--eq-steps", type=int, default=30000, help="c5: chain steps per training sample")
    ap.add_argument("--epochs", type=int, default=200, help="c5: training epochs")
    ap.add_argument("--horizon", type=int, default=20000, help="c5: least chain steps of each evaluation trajectory")
    ap.add_argument("--horizon-n2", type=float, default=16.0,
                    help="c5: most chain steps of an evaluation trajectory in units of N^2 (it ends earlier once "
                         "the chains from the minimal and maximal tilings agree, and runs as long again)")
    ap.add_argument("--held", type=int, default=4, help="c5: held-out domains")
    ap.add_argument("--train-max-side", type=int, default=0, help="c5: largest training tiling side (0: 2 x canvas)")
    ap.add_argument("--held-big", type=int, default=0, help="c5: tiling side of one extra large held-out Aztec diamond")
    ap.add_argument("--big-horizon-n2", type=float, default=3.0, help="c5: horizon cap of the large one, in N^2")
    ap.add_argument("--big-chains", type=int, default=4, help="c5: chains per start on held-out domains above 520")
    ap.add_argument("--held-sides", type=int, nargs=2, default=[96, 250], help="c5: held-out tiling side range")
    ap.add_argument("--eval-chains", type=int, default=16, help="c5: chains per start in the evaluation")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_each_way": n * 4, "dispatches": 3}), flush=True)
        return 0

    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (independent chains shard "
              "across GPUs without communication)", file=sys.stderr)
        return 2

    import randomtilings_amd as rt
    from randomtilings_amd import statics
    from randomtilings_amd.params import default_table

    if args.workload != "c2":
        return {"c1": bench_c1, "c3": bench_c3, "c4": bench_c4, "c5": bench_c5, "tridimer": bench_tridimer,
                "closure": bench_closure, "gemm": bench_gemm}[args.workload](args)

    d = statics.call("lozenge.AlmostHexDomain", args=[args.hex, 0])
    t0 = statics.call("lozenge.MinTiling", d)
    N = int(round(np.sqrt(t0.size)))
    G = N * (N // 3)
    params = default_table(G)

    ctx = rt.Context(0)
    ctx.load_params(params)
    ch = ctx.chain(rt.RT_LOZENGE, rt.RT_LOZENGE_QREF, t0, [args.q])
    if args.tile_config >= 0:
        ch.set_path(rt.RT_PATH_PERSISTENT, args.tile_config)
    else:
        ch.autotune(16384)  # on a scratch clone; the measured chain's trajectory is unchanged
    ch.seed(args.seed)
    S = args.chain_steps
    print(f"# bench: lozenge hexagon {args.hex} (N={N}), q={args.q}, {S} chain steps per bench step, "
          f"{args.warmup} warmup + {args.steps} timed", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        ch.run(S)
    ch.sync()

    ctx.sync()
    ctx.mark(0)
    w0 = time.perf_counter()
    for _ in range(args.steps):
        ch.run(S)
    ctx.mark(1)
    ch.sync()
    w1 = time.perf_counter()
    gpu_ms = ctx.elapsed_ms(0, 1)
    wall = w1 - w0
    attempts = ch.attempts_per_step
    # vertices of the tiled region (nonzero codes) in one colour class, what a step visits
    in_domain = int(np.count_nonzero(t0)) // 3
    total_attempts = args.steps * S * attempts
    value = total_attempts / wall

    # roofline: algorithmic bytes = read + write of the packed state per chain step
    state_bytes = ch.state_bytes
    launches_per_run = max(1, ch.last_launches)
    n_launch = args.steps * launches_per_run
    steps_per_launch = S / launches_per_run
    launch_s = gpu_ms / 1e3 / n_launch
    bytes_per_launch = 2.0 * state_bytes * steps_per_launch
    achieved = bytes_per_launch / launch_s / 1e9

    # HBM traffic is measured by rocprofv3 (FETCH_SIZE / WRITE_SIZE) in a separate profiling run
    # of this same workload; it is attached only when this run used the profiled kernel
    # configuration on the profiled hexagon, else null.
    cfg_used = list(ch.tile_configs()[ch.last_config]) if ch.last_config >= 0 else None
    traffic, traffic_src = None, None
    tf = ROOT / "profiles" / "traffic_lozenge_hex1000.json"
    if tf.exists():
        try:
            tj = json.loads(tf.read_text())
            if tj.get("hexagon") == args.hex and tj.get("tile_config") == cfg_used:
                traffic = float(tj["bytes_per_chain_step"]) * steps_per_launch
                traffic_src = f"rocprofv3 FETCH_SIZE/WRITE_SIZE run of this workload and kernel ({tf.name})"
        except Exception:
            traffic = None

    # Executed work: the reference rotates every work-item of the N x N/3 class arrays; the
    # packed sweep does not run tiles whose cells can never change (outside the hexagon).
    executed = attempts
    if cfg_used is not None and rt.tile_config_packed(rt.RT_LOZENGE, ch.last_config):
        tr, tc, th = cfg_used[:3]
        m, ok = rt.lozenge_skip_map(t0, tr, tc, th)
        if ok and ch.last_grid[1] < ch.last_grid[0]:
            W = N // 3
            executed = 0
            for (ti, tj_), run in np.ndenumerate(m):
                if run:
                    executed += (min(N, (ti + 1) * tr) - ti * tr) * (min(W, (tj_ + 1) * tc) - tj_ * tc)
    exec_frac = executed / attempts

    cpu = None
    if not args.no_cpu and args.cpu_steps > 0:
        import oracle

        # all cores the process may use (the oracle's per-kernel loops over a colour class run on
        # threads; the result does not depend on it), and one core
        cores = oracle.default_threads()
        oracle.set_threads(cores)
        cpu_all = oracle.lozenge_time(t0, 4 * args.cpu_steps, args.seed, params, rule=rt.RT_LOZENGE_QREF, q=args.q)
        oracle.set_threads(1)
        cpu_one = oracle.lozenge_time(t0, args.cpu_steps, args.seed, params, rule=rt.RT_LOZENGE_QREF, q=args.q)
        oracle.set_threads(cores)
        # parity on the same sample: the GPU walk of the same steps must match the oracle bit for bit
        g = ctx.lozenge_walk(t0, args.cpu_steps, args.seed, rule=rt.RT_LOZENGE_QREF, q=args.q)
        o = oracle.lozenge_walk(t0, args.cpu_steps, args.seed, params, rule=rt.RT_LOZENGE_QREF, q=args.q)
        cpu = {
            "value": 4 * args.cpu_steps * attempts / cpu_all,
            "unit": "flip attempts/s",
            "cores": cores,
            "kind": "port",
            "single_core_value": args.cpu_steps * attempts / cpu_one,
            "sample": f"oracle (sequential C++ restatement, per-kernel loops on {cores} threads) LozengeTiler Walk "
                      f"of {4 * args.cpu_steps} steps on the same N={N} hexagon, q-rule q={args.q}, host split/merge "
                      f"included; single_core_value: {args.cpu_steps} steps on one thread; GPU result of "
                      f"{args.cpu_steps} steps bit-identical: {bool(np.array_equal(g, o))}",
        }

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "flip attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (deterministic hexagon domain, TinyMT default parameter table)",
        "config": {
            "workload": f"lozenge q-volume chain, hexagon a=b=c={args.hex} (AlmostHexDomain({args.hex},0), N={N}), "
                        f"q={args.q} reference q-rule, empty-room start, {S} chain steps per bench step",
            "chain_steps_per_step": S,
            "attempts_per_chain_step": attempts,
            "sweeps_per_step": S / 3.0,
            "parallelism": "single GPU",
            "tile_config": ch.tile_configs()[ch.last_config] if ch.last_config >= 0 else "stepwise",
            # tiles of the sweep's grid and how many ran: tiles whose cells are all zero (outside
            # the hexagon) never change and are left out; the attempts above still count every
            # Rotate work-item, as the reference does
            "tiles": ch.last_grid[0],
            "tiles_run": ch.last_grid[1],
            "in_domain_sites_per_chain_step": in_domain,
            "in_domain_attempts_per_s": in_domain * args.steps * S / wall,
            # the work the GPU executed: positions of the tiles that ran
            "executed_attempts_per_chain_step": executed,
            "executed_attempts_per_s": executed * args.steps * S / wall,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            # achieved counts algorithmic bytes: a read and a write of the whole packed state
            # per chain step (every Rotate work-item, as the reference counts them);
            # frac_executed counts only the state of the tiles that ran
            "basis": "algorithmic bytes, 2 x packed state per chain step",
            "frac_executed": achieved * exec_frac / HBM_PEAK_GBS,
            "traffic_source": traffic_src,
            # what the profile of this kernel found binding (the state is LDS-resident; HBM sees
            # only the edge-band exchange): VALU issue and LDS/barrier latency
            "limiter": "valu issue + latency (SQ counters, profiles/lozenge_hex1000_qref_persist.md)",
        },
        "cpu_baseline": cpu,
        "gpu_kernel_ms_per_chain_step": gpu_ms / (args.steps * S),
    }
    print(json.dumps(line), flush=True)
    ch.close()
    ctx.close()
    return 0


def _line(metric, value, unit, args, ms_per_step, dtype, workload, extra, higher=True):
    return {
        "metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": higher, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype, "data": "synthetic (deterministic domain, TinyMT default parameter table)",
        "config": dict({"workload": workload, "parallelism": "single GPU"}, **extra),
    }


def bench_c1(args) -> int:
    """Domino tilings of the Aztec diamond of order 128 (tiling side 262), dominokernelCFTP.cl's
    rule, minimal start, 10^5 chain steps per bench step; latency-bound (the 67 KB state lives
    in LDS), so the figure of merit is chain steps per second."""
    import randomtilings_amd as rt
    from randomtilings_amd import statics
    from randomtilings_amd.params import default_table

    d = statics.call("domino.AztecDiamond", args=[128])
    t0 = statics.call("domino.MinTiling", d)
    N = int(round(np.sqrt(t0.size)))
    ctx = rt.Context(0)
    ctx.load_params(default_table(t0.size // 2))
    ch = ctx.chain(rt.RT_DOMINO, rt.RT_DOMINO_CFTP, t0)
    ch.autotune(4096)
    S = 100000
    ch.seed(args.seed)
    for _ in range(args.warmup):
        ch.run(S)
    ch.sync()
    if args.check:
        ch.seed(args.seed)
    ctx.sync()
    ctx.mark(0)
    w0 = time.perf_counter()
    for _ in range(args.steps):
        ch.run(S)
    ctx.mark(1)
    ch.sync()
    wall = time.perf_counter() - w0
    gpu_ms = ctx.elapsed_ms(0, 1)
    extra = {"aztec_order": 128, "tiling_side": N, "chain_steps_per_step": S,
             "attempts_per_chain_step": ch.attempts_per_step,
             "flip_attempts_per_s": ch.attempts_per_step * S * args.steps / wall,
             "gpu_us_per_chain_step": gpu_ms * 1e3 / (S * args.steps),
             "tile_config": ch.tile_configs()[ch.last_config] if ch.last_config >= 0 else "stepwise"}
    if args.check:
        import oracle
        g = ch.read()
        o = oracle.domino_walk(t0, S * args.steps, args.seed, default_table(t0.size // 2), rule=1)
        extra["bit_identical_to_oracle"] = bool(np.array_equal(g, o))
    line = _line("chain steps/s (domino Aztec diamond n=128, CFTP rule)", S * args.steps / wall, "steps/s", args,
                 wall * 1e3 / args.steps, "u8", "domino Aztec diamond n=128, dominokernelCFTP.cl rule, minimal start, "
                 "10^5 chain steps per bench step", extra)
    if not args.no_cpu:
        import oracle
        cs = 20000
        w0 = time.perf_counter()
        oracle.domino_walk(t0, cs, args.seed, default_table(t0.size // 2), rule=1)
        cpu_s = time.perf_counter() - w0
        line["cpu_baseline"] = {"value": cs / cpu_s, "unit": "steps/s", "cores": oracle.default_threads(), "kind": "port",
                                "sample": f"oracle (C++ restatement, per-kernel loops on {oracle.default_threads()} threads) DominoTiler Walk of {cs} steps on the "
                                          "same Aztec diamond, host split/merge included"}
    print(json.dumps(line), flush=True)
    ch.close()
    ctx.close()
    return 0


def bench_c4(args) -> int:
    """Coupling from the past on the Aztec diamond of order --aztec, doubling schedule: wall
    time to coalescence of the coupled chain (one launch sequence per restart, device coupling
    test), under the uniform rule (dominokernelCFTP.cl) or, with --p12, the weighted heat-bath.
    Then --placement-samples exact samples of the order --placement-n diamond under the same
    rule, checked against the exact placement probabilities (Kasteleyn): z-scores of every
    domino position."""
    import randomtilings_amd as rt
    from randomtilings_amd import placement as PL
    from randomtilings_amd import statics
    from randomtilings_amd.params import default_table

    n = args.aztec
    p12 = None if args.p12 is None else [float(x) for x in args.p12]
    rule = rt.RT_DOMINO_CFTP if p12 is None else rt.RT_DOMINO_WEIGHTED
    d = statics.call("domino.AztecDiamond", args=[n])
    tmax, tmin = statics.call("domino.MaxTiling", d), statics.call("domino.MinTiling", d)
    ctx = rt.Context(0)
    ctx.load_params(default_table(tmax.size // 2))
    if args.pair_steps > 0:
        # time per coupled step at this order (no coalescence): the coupled pair chain
        # (RT_DOMINO_PAIR) on its autotuned configuration, one epoch of --pair-steps steps
        ch = ctx.chain(rt.RT_DOMINO_PAIR, rule, tmax | (tmin << 8), p12)
        cfg = ch.autotune(256)
        ch.seed(args.seed, mode=rt.RT_SEED_EPOCH)
        ch.run(max(1, args.pair_steps // 8))
        ch.sync()
        ctx.sync()
        ctx.mark(0)
        w0 = time.perf_counter()
        ch.run(args.pair_steps)
        ctx.mark(1)
        ch.sync()
        wall = time.perf_counter() - w0
        gpu_ms = ctx.elapsed_ms(0, 1)
        att = 2 * ch.attempts_per_step  # both chains
        extra = {"aztec_order": n, "tiling_side": int(round(np.sqrt(tmax.size))),
                 "rule": "dominokernelCFTP.cl (p = 1/2)" if p12 is None else f"weighted heat-bath p12={p12}",
                 "coupled_steps": args.pair_steps, "us_per_coupled_step": gpu_ms * 1e3 / args.pair_steps,
                 "tile_config": list(ch.tile_configs()[cfg]), "packed": rt.tile_config_packed(rt.RT_DOMINO, cfg),
                 "flip_attempts_per_s_both_chains": att * args.pair_steps / (gpu_ms / 1e3),
                 "state_bytes": ch.state_bytes}
        args.steps, args.warmup = 1, 0
        line = _line(f"microseconds per coupled step (CFTP pair, domino Aztec diamond n={n})",
                     gpu_ms * 1e3 / args.pair_steps, "us", args, wall * 1e3, "u8x2",
                     f"coupled max/min domino chains, Aztec diamond n={n}, {args.pair_steps} steps", extra,
                     higher=False)
        print(json.dumps(line), flush=True)
        ch.close()
        ctx.close()
        return 0
    times, steps, restarts = [], [], []
    for k in range(args.warmup + args.steps):
        w0 = time.perf_counter()
        def said(r, n, c, w0=w0, k=k):  # after every epoch and replay: the sample's progress
            what = "running" if c is None else ("coalesced" if c else "not coalesced")
            print(f"# sample {k}: replay {r}, {n} steps, {what}, {time.perf_counter() - w0:.1f} s", file=sys.stderr,
                  flush=True)

        out, tot, r = ctx.domino_cftp(tmax, tmin, args.seed + k, rule=rule, p12=p12, mode=rt.RT_CFTP_DOUBLING,
                                      first_steps=1024, max_restarts=30, progress=said)
        print(f"# sample {k}: {time.perf_counter() - w0:.2f} s, {r} restarts", file=sys.stderr, flush=True)
        if k >= args.warmup:
            times.append(time.perf_counter() - w0)
            steps.append(tot)
            restarts.append(r)
    mean = float(np.mean(times))
    executed = [sum(1024 * 2 ** i for i in range(r + 1)) for r in restarts]  # every replay's steps
    extra = {"aztec_order": n, "rule": "dominokernelCFTP.cl (p = 1/2)" if p12 is None else f"weighted heat-bath p12={p12}",
             "schedule": "doubling from 1024 steps", "final_replay_steps": steps,
             "restarts": restarts, "coupled_steps_executed_per_sample": executed,
             "chain_attempts_per_step": 2 * (tmax.size // 2), "samples": args.steps}
    if args.placement_samples > 0:
        m = args.placement_n
        dp = statics.call("domino.AztecDiamond", args=[m])
        pmax, pmin = statics.call("domino.MaxTiling", dp), statics.call("domino.MinTiling", dp)
        w0 = time.perf_counter()
        samples = [ctx.domino_cftp(pmax, pmin, 7919 * s + 3, rule=rule, p12=p12, mode=rt.RT_CFTP_DOUBLING)[0]
                   for s in range(args.placement_samples)]
        sample_s = time.perf_counter() - w0
        z, live = PL.zscores(samples, dp, p12 if p12 is not None else (0.5, 0.5, 0.5, 0.5))
        extra["placement_check"] = {"aztec_order": m, "samples": args.placement_samples, "positions": live,
                                    "max_abs_z": float(np.max(np.abs(z))), "mean_z2": float(np.mean(z ** 2)),
                                    "sampling_seconds": sample_s}
    line = _line(f"seconds to coalescence (CFTP, domino Aztec diamond n={n})", mean, "s", args, mean * 1e3,
                 "u8x2", f"CFTP domino Aztec diamond n={n}, coupled max/min chains", extra, higher=False)
    if not args.no_cpu:
        # the oracle's walk on one chain for a bounded sample; the reference's driver replays
        # both chains separately, so its time to coalescence projects as 2 x steps / rate
        import oracle
        cs = 500
        w0 = time.perf_counter()
        oracle.domino_walk(tmax, cs, args.seed, default_table(tmax.size // 2), rule=1 if p12 is None else 2, p12=p12)
        rate = cs / (time.perf_counter() - w0)
        line["cpu_baseline"] = {"value": 2 * float(np.mean(executed)) / rate, "unit": "s", "cores": oracle.default_threads(), "kind": "port",
                                "sample": f"oracle (C++ restatement, per-kernel loops on {oracle.default_threads()} threads) DominoTiler Walk of {cs} steps on the "
                                          f"maximal tiling ({rate:.0f} steps/s), projected over the "
                                          f"{float(np.mean(executed)):.0f} steps of both chains' replays"}
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


def bench_tridimer(args) -> int:
    """Triangle-dimer chain on IceCreamCone(--cone, 0) (the next tileable side: the cone's
    triangle count must be even): flip attempts per second of the one-launch-per-step tiled
    kernel (all three move types, reference schedule)."""
    import randomtilings_amd as rt
    from randomtilings_amd import statics
    from randomtilings_amd.params import default_table

    cone = args.cone
    while ((cone + 1) * (cone + 2) // 2) % 2:  # an odd number of triangles has no dimer cover
        cone += 1
    args.cone = cone
    t0 = statics.call("tridimer.IceCreamCone", args=[cone, 0])
    N = int(round(np.sqrt(t0.size)))
    ctx = rt.Context(0)
    ctx.load_params(default_table((N // 2) ** 2))
    ch = ctx.chain(rt.RT_TRIDIMER, 0, t0)
    ch.autotune(64)
    ch.seed(args.seed)
    S = 2000
    for _ in range(args.warmup):
        ch.run(S)
    ch.sync()
    ctx.sync()
    ctx.mark(0)
    w0 = time.perf_counter()
    for _ in range(args.steps):
        ch.run(S)
    ctx.mark(1)
    ch.sync()
    wall = time.perf_counter() - w0
    gpu_ms = ctx.elapsed_ms(0, 1)
    att = ch.attempts_per_step
    extra = {"cone_side": args.cone, "tiling_side": N, "move_lattice": N // 2, "chain_steps_per_step": S,
             "attempts_per_chain_step": att, "gpu_us_per_chain_step": gpu_ms * 1e3 / (S * args.steps),
             "state_bytes": ch.state_bytes, "tile_config": ch.last_config}
    line = _line("flip attempts/sec (triangle dimers)", att * S * args.steps / wall, "flip attempts/s", args,
                 wall * 1e3 / args.steps, "u16", f"triangle-dimer IceCreamCone({args.cone},0), reference move schedule",
                 extra)
    if not args.no_cpu:
        import oracle
        cs = 20
        oracle.set_threads(1)  # the triangle-dimer walk is sequential; its table set-up too now
        w0 = time.perf_counter()
        oracle.tridimer_walk(t0, cs, args.seed, default_table((N // 2) ** 2))
        cpu_s = time.perf_counter() - w0
        oracle.set_threads(oracle.default_threads())
        line["cpu_baseline"] = {"value": att * cs / cpu_s, "unit": "flip attempts/s", "cores": 1, "kind": "port",
                                "sample": f"oracle (sequential C++ restatement) TriangleDimerTiler Walk of {cs} steps "
                                          "on the same cone, host state derivation included"}
    print(json.dumps(line), flush=True)
    ch.close()
    ctx.close()
    return 0


def bench_c3(args) -> int:
    """Batched plain chains: --domains domains drawn from --seed with sides uniform in --sides,
    cycling Aztec diamonds, domino rectangles with random staircase notches (tileability
    checked by the device max/min construction), triangle-dimer hexagons and rectangle-triangle
    hexagons (or the --kinds given); every chain runs --domain-steps steps.  One batched run (blocked launches per
    model, HIP-graph replay) against one rt_chain_run per domain, on identical chains."""
    import randomtilings_amd as rt
    from randomtilings_amd import statics
    from randomtilings_amd.params import default_table

    args.steps, args.warmup = 1, 0  # the line describes the single timed batched run
    from randomtilings_amd import domains

    lo, hi = args.sides
    ctx = rt.Context(0)
    specs = domains.mix(ctx, args.domains, (lo, hi), args.seed, tuple(args.kinds))
    gmax = max(int(t.size) for _, _, t, _ in specs)
    ctx.load_params(default_table(gmax))
    steps = args.domain_steps
    print(f"# c3: {len(specs)} domains built (largest tiling {gmax} vertices)", file=sys.stderr, flush=True)

    def make():
        chs = [ctx.chain(m, r, t) for m, r, t, _ in specs]
        for k, ch in enumerate(chs):
            ch.seed(args.seed + k)
        return chs

    solo, solo_s = [], None
    if not args.no_solo:
        solo = make()
        ctx.sync()
        w0 = time.perf_counter()
        for k, ch in enumerate(solo):
            ch.run(steps)
            if k % 64 == 63:  # progress (the stream runs the chains one after another either way)
                ctx.sync()
                print(f"# c3: {k + 1} domains run one by one, {time.perf_counter() - w0:.1f} s", file=sys.stderr,
                      flush=True)
        ctx.sync()
        solo_s = time.perf_counter() - w0
        print(f"# c3: one run per domain {solo_s:.1f} s", file=sys.stderr, flush=True)
    batched = make()
    b = rt.Batch(ctx)
    for ch in batched:
        b.add(ch)
    tuned = None
    if args.autotune:
        # every applicable configuration of each model timed on the batch itself (its chains
        # are put back as they were), the fastest kept for the run below
        w0 = time.perf_counter()
        tuned = b.autotune(64)
        print(f"# c3: batch autotune {tuned} in {time.perf_counter() - w0:.1f} s", file=sys.stderr, flush=True)
    elif args.batch_configs:
        for m, cfg in enumerate(args.batch_configs):
            b.set_config(m, cfg)
    ctx.sync()
    w0 = time.perf_counter()
    b.run(steps, True)
    b.sync()
    batch_s = time.perf_counter() - w0
    print(f"# c3: batched {batch_s:.1f} s", file=sys.stderr, flush=True)
    att = sum(ch.attempts_per_step for ch in batched) * steps
    same = None if args.no_solo else True
    for k, (a, c) in enumerate(zip(solo, batched)):
        try:
            same = same and bool(np.array_equal(a.read(), c.read()))
        except rt.RTError:
            print(f"# c3: domain {k} ({specs[k][3]}, tiling side {int(round(np.sqrt(specs[k][2].size)))}, "
                  f"config {a.last_config}) failed", file=sys.stderr, flush=True)
            raise
    kinds = {}
    for (_, _, t, kind), ch in zip(specs, batched):
        kinds.setdefault(kind, [0, 0])
        kinds[kind][0] += 1
        kinds[kind][1] += ch.attempts_per_step
    extra = {"domains": args.domains, "sides": [lo, hi], "steps_per_domain": steps,
             "attempts_per_step_all_domains": att // steps, "per_kind_domains_attempts": kinds,
             "batch_seconds": batch_s, "one_run_per_domain_seconds": solo_s,
             "one_run_per_domain_flips_per_s": att / solo_s if solo_s else None, "batch_launches": b.last_launches,
             "batch_graphs": b.last_graphs, "bit_identical_batch_vs_individual": same,
             "batch_configs_autotuned": tuned, "batch_configs_set": args.batch_configs}
    what = {"aztec": "Aztec diamonds", "notched": "notched domino rectangles", "tridimer": "triangle-dimer hexagons",
            "recttri": "rect-triangle hexagons", "lozenge": "lozenge almost-hexagons"}
    line = _line("aggregate flip attempts/s (batched domains)", att / batch_s, "flip attempts/s", args,
                 batch_s * 1e3, "u8/i32/u16",
                 f"{args.domains} domains, sides U[{lo},{hi}], kinds {'/'.join(args.kinds)} "
                 f"({', '.join(what[k] for k in args.kinds)}), {steps} steps each",
                 extra)
    if not args.no_cpu:
        # the oracle on one domain of each kind (the largest), a bounded number of steps each
        import oracle
        walk = {rt.RT_DOMINO: lambda t, n, sd, p: oracle.domino_walk(t, n, sd, p, rule=1),
                rt.RT_TRIDIMER: lambda t, n, sd, p: oracle.tridimer_walk(t, n, sd, p),
                rt.RT_RECTTRI: lambda t, n, sd, p: oracle.recttri_walk(t, n, sd, p),
                rt.RT_LOZENGE: lambda t, n, sd, p: oracle.lozenge_walk(t, n, sd, p, rule=0)}
        cpu_att, cpu_s, picked = 0, 0.0, []
        for kind in sorted({sp[3] for sp in specs}):
            k = max((i for i in range(len(specs)) if specs[i][3] == kind), key=lambda i: specs[i][2].size)
            m, _, t, _ = specs[k]
            cs = 20
            w0 = time.perf_counter()
            walk[m](t, cs, args.seed + k, default_table(gmax))
            cpu_s += time.perf_counter() - w0
            cpu_att += batched[k].attempts_per_step * cs
            picked.append(f"{kind} side {int(round(np.sqrt(t.size)))}")
        line["cpu_baseline"] = {"value": cpu_att / cpu_s, "unit": "flip attempts/s", "cores": oracle.default_threads(), "kind": "port",
                                "sample": f"oracle (C++ restatement, per-kernel loops on {oracle.default_threads()} threads; the triangle-dimer walk runs on one) Walk of 20 steps on the largest domain "
                                          "of each kind: " + ", ".join(picked)}
    print(json.dumps(line), flush=True)
    b.close()
    for ch in solo + batched:
        ch.close()
    ctx.close()
    return 0


def bench_c5(args) -> int:
    """Learned warm start on 2-periodic weighted domino chains.  Training: --train-domains
    random domains (Aztec diamonds and notched squares fitting a --canvas vertex grid), each
    with edge weights a, b ~ U[0.3, 1] (weighted heat-bath, p12 = 1/2, 1/2, a^2/(a^2+b^2),
    b^2/(a^2+b^2)); the target of each is the mean height fraction of --samples-per-domain
    chains run --eq-steps steps, alternately from the minimal and the maximal tiling.  The U-Net (4 levels, 32 base channels,
    bf16 matrix cores) is trained on the GPU.  Evaluation on --held held-out weighted domains
    with tiling sides in --held-sides (predicted on features sampled down to the canvas):
    chain steps to equilibrium from the minimal tiling (plain), from the midpoint
    (h_min + h_max) / 2 projected and mean-matched exactly as the prediction is (midpoint), and
    from the projected prediction (learned), on two observables: the mean interior height
    fraction (which the projection is matched to) and the mean |h - h_bar| (h_bar the long-run
    mean field; not targeted).  The line's value is plain / learned steps on the second."""
    import randomtilings_amd as rt
    from randomtilings_amd import warmstart as ws
    from randomtilings_amd.params import default_table

    rng = np.random.default_rng(args.seed)
    ctx = rt.Context(0)
    lo_side, hi_side = args.held_sides
    ctx.load_params(default_table(max(2 * args.canvas, args.train_max_side, hi_side + 8, args.held_big + 8) ** 2))
    t0 = time.perf_counter()
    # training domains up to --train-max-side (default twice the canvas): the larger ones are
    # seen sampled down, as the held-out domains are
    tmax_side = args.train_max_side or 2 * args.canvas
    doms = [ws.Domain(ctx, d, ws.periodic_weights(rng)) for _, d in ws.domain_bank(rng, args.train_domains, tmax_side)]
    samples = ws.equilibrium_samples(ctx, doms, args.eq_steps, args.samples_per_domain, args.seed)
    X, T, M = ws.build_dataset(ctx, doms, samples, args.canvas, average=True)
    data_s = time.perf_counter() - t0
    net = rt.UNet(ctx, 4, 32, ws.IN_CH, seed=args.seed)
    t0 = time.perf_counter()
    losses = ws.train(net, X, T, M, args.epochs, 16, 2e-3, rng)
    train_s = time.perf_counter() - t0
    hrng = np.random.default_rng(args.seed + 99)
    held = [ws.Domain(ctx, d, ws.periodic_weights(hrng))
            for _, d in ws.domain_bank(hrng, args.held, hi_side, min_side=lo_side)]
    if args.held_big:  # one large Aztec diamond, with a shorter horizon cap
        held.append(ws.Domain(ctx, ws.aztec(max(4, (args.held_big - 6) // 2)), ws.periodic_weights(hrng)))
    results = []
    for k, dom in enumerate(held):
        n2 = args.big_horizon_n2 if (args.held_big and k == len(held) - 1) else args.horizon_n2
        horizon = int(max(args.horizon, n2 * dom.N * dom.N))
        every = max(100, horizon // 400)
        horizon = (horizon // every) * every
        print(f"# held-out domain {k + 1}/{len(held)}: N = {dom.N}, p12 = {dom.p12}, horizon {horizon}",
              file=sys.stderr, flush=True)
        ws.warm_tiling(ctx, net, dom, args.canvas)  # warm-up of the inference path
        w0 = time.perf_counter()
        tw = ws.warm_tiling(ctx, net, dom, args.canvas)  # U-Net forward, projection, tiling read-off
        prep_s = time.perf_counter() - w0
        tm = ws.midpoint_tiling(ctx, dom)
        chains = args.eval_chains if dom.N <= 520 else args.big_chains
        w0 = time.perf_counter()
        def said(p, step, total, k=k):
            print(f"# held-out domain {k + 1}: pass {p}, {step}/{total} steps", file=sys.stderr, flush=True)

        cmp = ws.compare_starts(ctx, dom, {"plain": dom.tmin, "midpoint": tm, "learned": tw}, horizon, every,
                                77 + k, chains, progress=said)
        eval_s = time.perf_counter() - w0
        # wall clock of one chain on this domain: seconds per chain step
        ch = dom.chain(ctx, dom.tmin)
        ch.seed(5)
        ch.run(2000)
        ch.sync()
        w0 = time.perf_counter()
        ch.run(20000)
        ch.sync()
        step_s = (time.perf_counter() - w0) / 20000
        ch.close()
        cmp.update({"N": dom.N, "p12": dom.p12, "horizon": horizon, "every": every, "chains_per_start": chains,
                    "sampled_down_by": max(1, -(-dom.N // args.canvas)),
                    "seconds_per_chain_step": step_s, "learned_prep_seconds": prep_s, "eval_seconds": eval_s})
        print(f"# {json.dumps(cmp)}", file=sys.stderr, flush=True)
        results.append(cmp)

    # domains whose brackets never met within the horizon have no trustworthy equilibrium value:
    # they are reported but left out of the ratios
    scored = [r for r in results if r["settled"]] or results

    def ratio(a, b, obs):
        # a start already in the band reads 0 steps; readings are `every` steps apart, so the
        # denominator is floored at that resolution and the ratio is a lower bound
        num = np.mean([r[a][obs] for r in scored])
        den = max(np.mean([r["every"] for r in scored]), np.mean([r[b][obs] for r in scored]))
        return float(num / den)

    summary = {f"{a}_over_{b}_{obs}": ratio(a, b, f"steps_{obs}")
               for a in ("plain", "midpoint") for b in ("learned",) for obs in ("height", "deviation")}
    summary["plain_over_midpoint_deviation"] = ratio("plain", "midpoint", "steps_deviation")
    summary["scored_domains"] = len(scored)
    wall = float(np.mean([r["plain"]["steps_deviation"] * r["seconds_per_chain_step"] for r in scored]) /
                 np.mean([r["learned_prep_seconds"] + r["learned"]["steps_deviation"] * r["seconds_per_chain_step"]
                          for r in scored]))
    extra = {"canvas": args.canvas, "train_domains": args.train_domains, "train_examples": int(len(X)),
             "weights": "2-periodic, a, b ~ U[0.3, 1], p12 = (1/2, 1/2, a^2/(a^2+b^2), b^2/(a^2+b^2))",
             "chain_samples_averaged_per_domain": args.samples_per_domain, "eq_steps_per_sample": args.eq_steps,
             "unet": "4 levels x 32 base channels, bf16 MFMA, fp32 accumulate", "params": net.param_count,
             "epochs": args.epochs, "final_loss": float(np.mean(losses[-10:])), "first_loss": float(losses[0]),
             "data_seconds": data_s, "train_seconds": train_s, "held_out": results, **summary,
             "wall_clock_plain_over_learned_deviation": wall}
    line = _line("sweeps-to-equilibrium, plain / learned warm start (domino, weighted, |h - h_bar| observable)",
                 summary["plain_over_learned_deviation"], "x", args, train_s * 1e3, "bf16/fp32",
                 f"warm start from the projected U-Net prediction vs the minimal tiling and the midpoint, held-out "
                 f"sides {lo_side}-{hi_side}", extra)
    print(json.dumps(line), flush=True)
    net.close()
    ctx.close()
    return 0


def bench_closure(args) -> int:
    """Maximal tiling of the a=b=c=--hex hexagon (N = 2007 at 1000) built on the device
    (host boundary walk, device closure, device read-off) against the host construction."""
    import randomtilings_amd as rt
    from randomtilings_amd import statics

    d = statics.call("lozenge.AlmostHexDomain", args=[args.hex, 0])
    ctx = rt.Context(0)
    for _ in range(args.warmup):
        ctx.extremal_tiling(rt.RT_LOZENGE, d, True)
    times = []
    for _ in range(args.steps):
        w0 = time.perf_counter()
        t, _ = ctx.extremal_tiling(rt.RT_LOZENGE, d, True)
        times.append(time.perf_counter() - w0)
    w0 = time.perf_counter()
    ht = statics.call("lozenge.MaxTiling", d)
    host_s = time.perf_counter() - w0
    mean = float(np.mean(times))
    N = int(round(np.sqrt(t.size)))
    extra = {"hexagon": args.hex, "tiling_side": N, "host_construction_s": host_s,
             "identical_to_host": bool(np.array_equal(t, ht))}
    line = _line("seconds per maximal tiling (device closure)", mean, "s", args, mean * 1e3, "i32",
                 f"lozenge MaxTiling of AlmostHexDomain({args.hex},0)", extra, higher=False)
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


def bench_gemm(args) -> int:
    """The U-Net's matrix-core products on device-resident bf16 operands: the convolution
    products of the 4-level, 32-channel network on a batch of 16 64 x 64 images (forward,
    weight gradient, input gradient at every level), and a large square product; TFLOP/s
    against the dense bf16 peak (2.5 PFLOP/s)."""
    import randomtilings_amd as rt

    ctx = rt.Context(0)
    shapes = []
    for lvl, cin, cout in [(0, 32, 32), (1, 64, 64), (2, 128, 128), (3, 256, 256)]:
        P = 16 * (64 >> lvl) ** 2
        shapes += [(f"fwd L{lvl}", P, cout, 9 * cin, 0), (f"dW L{lvl}", 9 * cin, cout, P, 1),
                   (f"dX L{lvl}", P, 9 * cin, cout, 2)]
    shapes.append(("square 8192", 8192, 8192, 8192, 0))
    rows = []
    for name, M, N, K, lay in shapes:
        ms = rt.nn_gemm_time(ctx, M, N, K, lay, reps=max(2, args.steps))
        tf = 2.0 * M * N * K / (ms * 1e-3) / 1e12
        rows.append({"product": name, "M": M, "N": N, "K": K, "layout": lay, "ms": ms, "tflops": tf})
        print(f"# gemm {name}: {M}x{N}x{K} layout {lay}: {ms:.3f} ms, {tf:.1f} TFLOP/s", file=sys.stderr, flush=True)
    big = rows[-1]
    peak = 2500.0
    line = _line("bf16 matrix-core product throughput (U-Net kernel)", big["tflops"], "TFLOP/s", args, big["ms"],
                 "bf16/fp32", "square 8192 product and the U-Net's convolution products",
                 {"products": rows})
    line["roofline"] = {"bound": "mfma", "achieved": big["tflops"], "peak": peak, "unit": "TFLOP/s",
                        "frac": big["tflops"] / peak, "traffic": None}
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
