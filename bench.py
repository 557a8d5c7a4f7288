"""Benchmark: two-flavour (Nf=2) Wilson-fermion HMC trajectories per second on a
1024x1024 lattice on one MI355X, mixed-precision MD solves beside the plain
FP64 form on identical random streams, plus the roofline position of the
dominant kernel and the CPU oracle on the host cores.

Workload (one "step" = one HMC trajectory): compact U(1) gauge action + Nf=2
Wilson fermions, beta = 4.0, m0 = 0.2, leapfrog ns = 10, eps = 0.1, CG on
(gamma5 Dw)^2 to |r|^2 < tol |b|^2, start from a hot configuration (link
angles uniform in [-pi, pi), seed 1234) thermalised by 200 pure-gauge OMF4
trajectories.  The acceptance step checks the last MD solve in FP64 (true
residual below tol, FP64 iterations if not) and evaluates the pseudofermion
action as 2 Re X^dag F - |Q X|^2, in both precisions.

    python bench.py [--steps K] [--warmup W] [--L 1024] [--tol 1e-24]

Prints ONE JSON line (rank 0).  Only N = 1 GPU is supported.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "dav-qed2d_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "HMC traj/s at 1024² Nf=2 (mixed vs FP64, same ΔH); Dw GB/s vs HBM peak"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per lattice site of each kernel class (each logical field
# access counted once; FP64 complex = 16 B, FP32 complex = 8 B, 2 spin
# components or 2 link directions per site).
BYTES_PER_SITE = {
    # p = r + beta p_old (reads r, p_old), t = Q p (reads U), writes p, t; the
    # FP32 site-pair K1 reads its U(1) links packed, one float each (8 B/site),
    # and every other one (even k) applies the two pending x updates
    # x += a p_{k-2} + a p_old (x read+write, p_{k-2} read): 72 and 120 B/site,
    # 96 on average over the alternating launches
    "cg_k1_f64": 32 * 5, "cg_k1_f32": (16 * 4 + 8 + 16 * 7 + 8) // 2,
    # Ap = Q t (+ mu^2 p) (reads t, U), r -= a Ap (r read+write); FP64 also
    # x += a p (reads p, x read+write); FP32: packed links
    "cg_k2_f64": 32 * 7, "cg_k2_f32": 16 * 3 + 8,
    # so = Q si: si, U read, so written
    "q_f64": 96,
    # mom += eps (F_gauge + F_fermion(X, Q X)): U, X, Y read, mom read+write
    "force_mom": 128,
    # U <- exp(i eps p) U: U read+write, p read
    "updt_fld": 80,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--rocprof-summary":
        return rocprof_summary(sys.argv[2:])
    if len(sys.argv) > 1 and sys.argv[1] == "--cpu-worker":
        return cpu_worker(sys.argv[2], sys.argv[3], sys.argv[4])
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--L", type=int, default=1024)
    ap.add_argument("--beta", type=float, default=4.0)
    ap.add_argument("--m0", type=float, default=0.2)
    ap.add_argument("--ns", type=int, default=10)
    ap.add_argument("--eps", type=float, default=0.1)
    ap.add_argument("--tol", type=float, default=1e-24)
    ap.add_argument("--therm", type=int, default=200)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile", type=int, default=0, help="FP64 tile rows (0: library default)")
    ap.add_argument("--tile32", type=int, default=0, help="FP32 tile rows (0: library default)")
    ap.add_argument("--no-dw", action="store_true")
    ap.add_argument("--dw-only", action="store_true",
                    help="only the STREAM and gamm5Dw apply measurements (one JSON line)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--accept-traj", type=int, default=100,
                    help="trajectories of the acceptance comparison (0: skip)")
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--leg", choices=["ensemble", "rhmc", "tolerance"], default=None,
                    help="run one side measurement alone and print its JSON (for profiling)")
    ap.add_argument("--ens-seq", type=int, default=3,
                    help="trajectories of the ensemble's one-chain-at-a-time comparison "
                         "(0 skips it)")
    ap.add_argument("--rhmc-traj", type=int, default=20,
                    help="trajectories of the RHMC 128^2 Markov chain, per force precision")
    ap.add_argument("--tiles", type=str, default="",
                    help="q64,q32,cg64,cg32[,cg64_k2,cg32_k2] tile heights (skips the "
                         "autotuner)")
    ap.add_argument("--cpu-full", action="store_true",
                    help="also time the 8192^2 and RHMC trajectories at one thread (minutes)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this package has no multi-GPU path "
              "(use --gpus 1).", file=sys.stderr)
        sys.exit(2)

    # one-thread CPU oracle workers start before anything touches the GPU
    pool = CpuPool(cpu_jobs(args)) if not args.no_cpu else None

    import numpy as np
    import torch

    import qed2d as Q

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.dw_only:
        sprm = Q.LattParm((64, 64), 4.0)
        stream = Q.hbm_stream(sprm, 1 << 30)
        sprm.close()
        print(json.dumps({"hbm_stream_GBs": {k: round(v, 1) for k, v in stream.items()},
                          "dw_apply": dw_apply(Q, torch, args, stream)}), flush=True)
        return
    if args.leg == "ensemble":
        print(json.dumps(ensemble(Q, torch, nseq=args.ens_seq)), flush=True)
        return
    if args.leg == "tolerance":
        print(json.dumps(tolerance_scan(Q, torch)), flush=True)
        return
    if args.leg == "rhmc":
        r = rhmc_config(Q, torch, args, None)
        r.pop("_gpu", None)
        log("rhmc: " + json.dumps(r))
        print(json.dumps(r), flush=True)
        return
    L = args.L
    prm = Q.LattParm((L, L), args.beta)
    if args.tile:
        Q.lib.qed2d_ctx_set_tile(prm.ctx(), args.tile)
    if args.tile32:
        Q.lib.qed2d_ctx_set_tile_f32(prm.ctx(), args.tile32)
    seed = 1234

    # ---- start configuration: hot start + pure-gauge thermalisation --------
    U = Q.spinor(prm)
    Q.hot_start(U, prm, seed)
    Q.seed(seed, chain=1)
    t0 = time.time()
    for _ in range(args.therm):
        Q.HMC(U, 0.1, 10, [], prm, None)
    torch.cuda.synchronize()
    log(f"thermalised {args.therm} pure-gauge trajectories in {time.time() - t0:.2f}s, "
        f"plaquette {Q.Plaquette(U, prm):.6f}")
    U0 = U.clone()
    if args.tiles:
        tv = [int(v) for v in args.tiles.split(",")]
        Q.set_tiles(prm, *tv[:4])
        if len(tv) == 6:
            Q.set_tiles(prm, cg64_k2=tv[4], cg32_k2=tv[5])
    elif not (args.no_autotune or args.tile or args.tile32):
        t0 = time.time()
        tiles = Q.autotune(prm, 160)
        log(f"autotuned tiles {tiles} in {time.time() - t0:.2f}s")
    tiles = Q.get_tiles(prm)

    def run(precision, label, profile=False, steps=None):
        steps = steps or args.steps
        U.copy_(U0)
        Q.seed(seed, chain=0)
        acc = []
        infos = []
        for _ in range(args.warmup):
            Q.HMC(U, args.m0, args.eps, args.ns, acc, 100000, args.tol, prm, None,
                  precision=precision, acc_tol=args.tol)
        torch.cuda.synchronize()
        if profile:
            Q.profile_enable(prm, True, every=1)
        # (a one-element fill before and after the timed trajectories marks
        # their window in a kernel trace: bench.py --rocprof-summary --windows)
        mark = torch.empty(1, dtype=torch.int16, device="cuda")
        mark.fill_(0)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            infos.append(Q.HMC(U, args.m0, args.eps, args.ns, acc, 100000, args.tol, prm, None,
                               precision=precision, acc_tol=args.tol))
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        mark.fill_(1)
        prof = None
        if profile:
            prof = Q.profile_read(prm)
            Q.profile_enable(prm, False)
        it = sum(i.cg_iters for i in infos) / len(infos)
        log(f"{label}: {steps} traj in {dt:.3f}s = {steps / dt:.3f} traj/s, "
            f"{it:.0f} CG it/traj, acc {sum(i.accepted for i in infos)}/{len(infos)}")
        if prof:
            for k, (n, ms) in prof.items():
                if n:
                    log(f"   {k:16s} {n:7d} launches {ms:9.2f} ms  avg {1000 * ms / n:8.2f} us")
        return dt, infos, prof

    dt_mx, inf_mx, _ = run(Q.PREC_MIXED, "mixed")
    res = {}
    if not args.no_fp64:
        dt_64, inf_64, _ = run(Q.PREC_F64, "fp64")
        dmax = max(abs(a.dH - b.dH) for a, b in zip(inf_mx, inf_64))
        res["fp64_traj_per_s"] = round(args.steps / dt_64, 4)
        res["speedup_vs_fp64"] = round(dt_64 / dt_mx, 4)
        res["max_abs_dH_diff_vs_fp64"] = dmax
        res["accept_identical_vs_fp64"] = [a.accepted for a in inf_mx] == [b.accepted for b in inf_64]
        res["cg_iters_per_traj_fp64"] = sum(i.cg_iters for i in inf_64) / len(inf_64)
    res["cg_iters_per_traj_mixed"] = sum(i.cg_iters for i in inf_mx) / len(inf_mx)
    res["tiles"] = tiles
    res["dH_mixed"] = [round(i.dH, 10) for i in inf_mx]
    res["acceptance"] = sum(i.accepted for i in inf_mx) / len(inf_mx)

    # ---- roofline of the dominant kernel: a profiled pass of the same
    # workload, HIP events around every launch on the context's stream.  The
    # CG kernels' time is divided by the iterations that did work (launches
    # that only found the solve finished or paused are charged to them).
    _, inf_p, prof_mx = run(Q.PREC_MIXED, "mixed (profiled)", profile=True,
                            steps=min(args.steps, 10))
    V = L * L
    real_it = sum(i.cg_iters for i in inf_p)
    units = {k: (real_it if k.startswith("cg_") else n) for k, (n, ms) in prof_mx.items()}
    dom = max((k for k in prof_mx if k in BYTES_PER_SITE and prof_mx[k][0]),
              key=lambda k: prof_mx[k][1])
    n, ms = prof_mx[dom]
    avg_s = ms / units[dom] / 1000.0
    achieved = BYTES_PER_SITE[dom] * V / avg_s / 1e9
    traffic = pmc_traffic(dom, L)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": dom, "avg_us": round(avg_s * 1e6, 2), "launches_timed": n,
            "bytes_per_site": BYTES_PER_SITE[dom], "sites": V,
            "note": "1024^2 working set fits the 256 MB Infinity Cache: cache-assisted"}
    kern = {}
    for k, (n, ms) in prof_mx.items():
        # (the FP64 CG classes of a mixed run are the acceptance checks of the
        # last MD solve, which mostly find it converged: no per-iteration figure)
        if n and k in BYTES_PER_SITE and not (k.startswith("cg_") and k.endswith("_f64")):
            a = BYTES_PER_SITE[k] * V / (ms / units[k] / 1000.0) / 1e9
            kern[k] = {"avg_us": round(1000 * ms / units[k], 2), "GBs": round(a, 1),
                       "launches": n}
    res["kernels_mixed"] = kern
    stream = Q.hbm_stream(prm, 1 << 30)
    res["hbm_stream_GBs"] = {k: round(v, 1) for k, v in stream.items()}
    log(f"STREAM (3 x 1 GiB): " + ", ".join(f"{k} {v:.0f} GB/s" for k, v in stream.items()))
    roof["achievable_copy_GBs"] = round(stream["copy"], 1)
    if args.accept_traj > 0:
        res["acceptance_comparison"] = acceptance_comparison(Q, torch, args, prm, U, U0, seed)
    if not args.no_ensemble:
        res["ensemble_64_chains"] = ensemble(Q, torch, nseq=args.ens_seq)
    # (the CPU workers start here: nothing timed before runs beside them; the
    # apply bandwidths below are HIP-event times of the kernels themselves)
    if pool is not None:
        pool.submit("nf2_1024", U=U0.cpu().numpy(), m0=args.m0, eps=args.eps, ns=args.ns,
                    tol=args.tol, beta=args.beta, seed=seed)
    if not args.no_dw:
        res["dw_apply"] = dw_apply(Q, torch, args, stream)
    if not args.no_configs:
        res["configs"] = config_measurements(Q, torch, pool, args)

    # ---- CPU oracle on the host cores: whole trajectories, never scaled -----
    cpu = None
    if pool is not None:
        del U, U0
        prm.close()
        torch.cuda.empty_cache()
        cpu = cpu_baselines(pool, res.get("configs"), args)

    ms_step = 1000.0 * dt_mx / args.steps
    out = {
        "metric": METRIC,
        "value": round(args.steps / dt_mx, 4),
        "unit": "traj/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64 (FP32 CG iterations, FP64 reliable updates, forces, H)",
        "data": "synthetic",
        "config": {"workload": f"Nf=2 HMC {L}x{L} beta={args.beta} m0={args.m0} leapfrog "
                               f"ns={args.ns} eps={args.eps} CG tol={args.tol} (squared "
                               f"relative); hot start seed {seed} + {args.therm} pure-gauge "
                               f"OMF4 trajectories",
                   "lattice": [L, L], "precision": "mixed", "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    out.update(res)
    ac = res.get("acceptance_comparison")
    if res.get("acceptance") == 0.0:
        out["acceptance_note"] = (
            "at eps 0.1, ns 10 the 1024^2 energy violation is about 500 and every trajectory is "
            "rejected, so the identical accept/reject sequence of the headline is trivial; "
            "acceptance_comparison runs the same start as a Markov chain at tau = 1 in 50 steps")
    if ac:
        out["accepting_chain"] = {k: ac[k] for k in ("acceptance", "speedup_vs_fp64",
                                                     "accept_identical_vs_fp64",
                                                     "max_abs_dH_diff_vs_fp64", "trajectories")}
    print(json.dumps(out), flush=True)


def dw_apply(Q, torch, args, stream=None):
    """gamm5Dw apply bandwidth at the shapes the HBM claim is made on
    (1024^2 x 16 RHS, 4096^2, 8192^2; 1024^2 x 1 RHS is cache-assisted), FP64
    and FP32, timed with HIP events on the context's stream; every tile
    height in TILES is timed and the best kept."""
    out = {}
    shapes = [(1024, 1, "f64"), (1024, 16, "f64"), (4096, 1, "f64"), (8192, 1, "f64"),
              (1024, 1, "f32"), (1024, 16, "f32"), (4096, 1, "f32"), (8192, 1, "f32")]
    for L, k, dt in shapes:
        prm = Q.LattParm((L, L), 4.0)
        cdt = torch.complex128 if dt == "f64" else torch.complex64
        U = torch.exp(1j * torch.rand((2, L, L), dtype=torch.float64, device="cuda") * 6.283)
        U = U.to(cdt)
        shape = (k, 2, L, L) if k > 1 else (2, L, L)
        si = torch.randn(shape, dtype=cdt, device="cuda")
        so = torch.empty_like(si)

        def apply():
            if k > 1:
                Q.gamm5Dw_mrhs(so, U, si, 0.1, prm)
            elif dt == "f64":
                Q.gamm5Dw(so, U, si, 0.1, prm)
            else:
                Q.gamm5Dw_f32(so, U, si, 0.1, prm)
        Q.lib.qed2d_ctx_set_sync(prm.ctx(), 0)
        reps = 50 if L <= 1024 else 20
        per_tile = {}
        for rows in DW_TILES:
            if dt == "f64":
                Q.set_tiles(prm, q64=rows)
            else:
                Q.set_tiles(prm, q32=rows)
            for _ in range(3):
                apply()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record()
            for _ in range(reps):
                apply()
            e1.record()
            torch.cuda.synchronize()
            per_tile[rows] = e0.elapsed_time(e1) * 1000.0 / reps
        rows = min(per_tile, key=per_tile.get)
        us = per_tile[rows]
        b = (16 if dt == "f64" else 8)
        bytes_site = 2 * b + 4 * b * k          # U (2 links) + k x (read + write 2 components)
        gbs = bytes_site * L * L / (us * 1e-6) / 1e9
        key = f"{L}x{L}x{k}rhs_{dt}"
        out[key] = {"us": round(us, 2), "GBs": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 3),
                    "bytes_per_site": bytes_site, "tile": rows,
                    "us_per_tile": {str(r): round(v, 2) for r, v in per_tile.items()}}
        if stream:
            # against the measured copy of the same access pattern: one array
            # (single RHS) or 16 arrays at once (16 RHS)
            ref = stream["copy16"] if k > 1 else stream["copy"]
            out[key]["frac_of_measured_copy"] = round(gbs / ref, 3)
        log(f"dw {key}: {us:.1f} us, {gbs:.0f} GB/s (tile {rows})")
        del U, si, so
        prm.close()
        torch.cuda.empty_cache()
    return out


DW_TILES = (2, 4, 8)


def ensemble(Q, torch, ntraj=10, nseq=3):
    """64 independent Nf=2 chains of uneven sizes (16:16:16:8:8 chains of
    16^2, 32^2, 64^2, 128^2, 256^2; beta cycling over {2, 4, 6}, m0 over
    {0.1, 0.2}, seed 1000 + chain), leapfrog ns 10 eps 0.1, CG tol 1e-16:
    chain-trajectories per second of the batched entry point (mixed and FP64)
    beside the same chains run one after another through the single-lattice
    entry point (mixed)."""
    sizes = [(16, 16)] * 16 + [(32, 32)] * 16 + [(64, 64)] * 16 + [(128, 128)] * 8 + \
            [(256, 256)] * 8
    n = len(sizes)
    betas = [(2.0, 4.0, 6.0)[c % 3] for c in range(n)]
    am0s = [(0.1, 0.2)[(c // 3) % 2] for c in range(n)]
    seeds = [1000 + c for c in range(n)]
    B = Q.Batch(sizes, betas)
    B.hot_start(seeds)
    for t in range(2):
        B.HMC(am0s, 0.1, 10, 100000, 1e-16, seeds, t, precision=Q.PREC_MIXED)
    U0 = [B.get_links(c).clone() for c in range(n)]
    t0 = time.perf_counter()
    tiles = B.autotune(60)
    log(f"ensemble: autotuned tiles {tiles} in {time.perf_counter() - t0:.2f}s")
    out = {"chains": n, "sites": B.sites, "trajectories_timed": ntraj,
           "sequential_trajectories_timed": nseq, "tiles": tiles}
    mark = torch.empty(1, dtype=torch.int16, device="cuda")  # (trace windows, as run())
    dhs = {}
    # the same trajectories with each CG chunk queued as one HIP graph launch
    # (windows 2 and 3 of a trace): same kernels, so the same bits
    for prec, key, graphs in ((Q.PREC_MIXED, "mixed", False), (Q.PREC_F64, "fp64", False),
                              (Q.PREC_MIXED, "graph_mixed", True),
                              (Q.PREC_F64, "graph_fp64", True)):
        B.set_graphs(graphs)
        for c in range(n):
            B.set_links(c, U0[c])
        if graphs:  # (capture every chunk shape once, outside the timed region)
            B.HMC(am0s, 0.1, 10, 100000, 1e-16, seeds, 2, precision=prec)
            for c in range(n):
                B.set_links(c, U0[c])
        mark.fill_(0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dh = []
        for t in range(ntraj):
            dh += [i.dH for i in B.HMC(am0s, 0.1, 10, 100000, 1e-16, seeds, 2 + t,
                                       precision=prec)]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        mark.fill_(1)
        dhs[key] = dh
        out[f"{key}_chain_traj_per_s"] = round(n * ntraj / dt, 2)
        log(f"ensemble {key}: {n} chains x {ntraj} traj in {dt:.3f}s")
    B.set_graphs(False)
    out["graph_same_dH"] = (dhs["graph_mixed"] == dhs["mixed"] and
                            dhs["graph_fp64"] == dhs["fp64"])
    if nseq <= 0:
        B.close()
        return out
    # the same chains one at a time (single-lattice entry point, mixed)
    prms = [Q.LattParm(sz, b) for sz, b in zip(sizes, betas)]
    Us = [u.clone() for u in U0]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in range(nseq):
        for c in range(n):
            Q.seed(seeds[c], chain=c, traj=2 + t)
            Q.HMC(Us[c], am0s[c], 0.1, 10, [], 100000, 1e-16, prms[c], None,
                  precision=Q.PREC_MIXED)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out["sequential_mixed_chain_traj_per_s"] = round(n * nseq / dt, 2)
    out["batched_vs_sequential"] = round(out["mixed_chain_traj_per_s"] /
                                         out["sequential_mixed_chain_traj_per_s"], 2)
    log(f"ensemble sequential: {n} chains x {nseq} traj in {dt:.3f}s")
    for p in prms:
        p.close()
    B.close()
    return out


def tolerance_scan(Q, torch):
    """|dH_mixed - dH_fp64| against the CG tolerance (squared relative
    residual) on the same start and random streams: the 1024^2 accepting chain
    (leapfrog ns 50 eps 0.02, 3 trajectories), the 64-chain ensemble (one
    trajectory, the largest difference over the chains) and Nf=2 4096^2 (OMF4
    ns 2 eps 0.05, one trajectory).  Shows the tolerance at which the
    difference falls below 1e-9."""
    out = {}
    L = 1024
    prm = Q.LattParm((L, L), 4.0)
    U = Q.spinor(prm)
    Q.hot_start(U, prm, 1234)
    Q.seed(1234, chain=1)
    for _ in range(200):
        Q.HMC(U, 0.1, 10, [], prm, None)
    U0 = U.clone()
    r = {}
    for tol in (1e-16, 1e-18, 1e-20, 1e-22, 1e-24):
        dh = {}
        for prec in (Q.PREC_MIXED, Q.PREC_F64):
            U.copy_(U0)
            Q.seed(1234, chain=5)
            dh[prec] = [Q.HMC(U, 0.2, 0.02, 50, [], 100000, tol, prm, None, precision=prec,
                              acc_tol=tol).dH for _ in range(3)]
        r[f"{tol:.0e}"] = max(abs(a - b) for a, b in zip(dh[Q.PREC_MIXED], dh[Q.PREC_F64]))
        log(f"tolerance 1024^2 tol {tol:.0e}: max |ddH| {r[f'{tol:.0e}']:.2e}")
    out["nf2_1024_leapfrog_ns50_eps0.02"] = r
    del U, U0
    prm.close()
    # the ensemble
    sizes = [(16, 16)] * 16 + [(32, 32)] * 16 + [(64, 64)] * 16 + [(128, 128)] * 8 + \
            [(256, 256)] * 8
    n = len(sizes)
    betas = [(2.0, 4.0, 6.0)[c % 3] for c in range(n)]
    am0s = [(0.1, 0.2)[(c // 3) % 2] for c in range(n)]
    seeds = [1000 + c for c in range(n)]
    B = Q.Batch(sizes, betas)
    B.hot_start(seeds)
    for t in range(2):
        B.HMC(am0s, 0.1, 10, 100000, 1e-16, seeds, t, precision=Q.PREC_MIXED)
    E0 = [B.get_links(c).clone() for c in range(n)]
    r = {}
    for tol in (1e-16, 1e-20, 1e-24):
        dh = {}
        for prec in (Q.PREC_MIXED, Q.PREC_F64):
            for c in range(n):
                B.set_links(c, E0[c])
            dh[prec] = [i.dH for i in B.HMC(am0s, 0.1, 10, 100000, tol, seeds, 2,
                                            precision=prec, acc_tol=tol)]
        r[f"{tol:.0e}"] = max(abs(a - b) for a, b in zip(dh[Q.PREC_MIXED], dh[Q.PREC_F64]))
        log(f"tolerance ensemble tol {tol:.0e}: max |ddH| {r[f'{tol:.0e}']:.2e}")
    out["ensemble_64_chains"] = r
    B.close()
    torch.cuda.empty_cache()
    # Nf=2 4096^2
    L = 4096
    prm = Q.LattParm((L, L), 11.25)
    U = Q.cold_links(prm)
    Q.seed(55, chain=2)
    for _ in range(3):
        Q.HMC(U, 0.025, 20, [], prm, None)
    U0 = U.clone()
    r = {}
    for tol in (1e-20, 1e-22, 1e-24):
        dh = {}
        for prec in (Q.PREC_MIXED, Q.PREC_F64):
            U.copy_(U0)
            Q.seed(55)
            dh[prec] = Q.HMC(U, 0.05, 0.05, 2, [], 100000, tol, prm, None, integrator=Q.OMF4(),
                             precision=prec, acc_tol=tol).dH
        r[f"{tol:.0e}"] = abs(dh[Q.PREC_MIXED] - dh[Q.PREC_F64])
        log(f"tolerance 4096^2 tol {tol:.0e}: |ddH| {r[f'{tol:.0e}']:.2e}")
    out["nf2_4096_omf4_ns2_eps0.05"] = r
    del U, U0
    prm.close()
    return out


def rhmc_config(Q, torch, args, pool):
    """RHMC 128^2, beta 4, m0 0.1, 16 poles on power-method bounds, 4
    pseudofermions, OMF4 ns 12 eps 0.083: a Markov chain in each force
    precision (the CPU job, when a pool is given, times its first trajectory
    on the oracle)."""
    import numpy as np

    def sync_time(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, r

    prm = Q.LattParm((128, 128), 4.0)
    U = Q.cold_links(prm)
    Q.seed(77)
    for _ in range(20):
        Q.HMC(U, 0.1, 10, [], prm, None)
    b0 = Q.spinor(prm)
    Q.fill_pseudofermion(b0, prm, 77, 3, 0, 0)
    lmin, lmax = Q.power_method(U, 0.1, prm, iter=5000, b0=b0)
    ra, rb = np.sqrt(lmin * 0.99), np.sqrt(lmax * 1.01)
    rp = Q.get_rhmc_params(16, ra, rb)
    Uc = U.clone()
    traj0 = Q.rng_state()[2]
    # a Markov chain of args.rhmc_traj trajectories from the same start and
    # random streams with the force's solves in FP64 and in mixed precision
    # (FP32 multi-shift, then every pole by the batched mixed-precision CG to
    # the same FP64 tolerance; heatbath and acceptance FP64 in both)
    nt = args.rhmc_traj
    c3 = {"trajectories": nt, "cg_tol": 1e-16, "delta": rp.delta}
    runs = {}
    for prec, key in ((Q.PREC_F64, "fp64"), (Q.PREC_MIXED, "mixed")):
        U.copy_(Uc)
        Q.seed(77)
        dt, infos = sync_time(lambda: [
            Q.HMC(U, [0.1] * 4, 0.083, 12, [], 100000, 1e-16, prm, None, [rp] * 4,
                  integrator=Q.OMF4(), precision=prec, traj=traj0 + i) for i in range(nt)])
        runs[key] = infos
        dh = np.array([i.dH for i in infos])
        c3[key] = {"traj_per_s": round(nt / dt, 3),
                   "acceptance": round(float(np.mean([i.accepted for i in infos])), 4),
                   "exp_minus_dH_mean": round(float(np.mean(np.exp(-dh))), 4),
                   "exp_minus_dH_stderr": round(float(np.std(np.exp(-dh)) / np.sqrt(nt)), 4),
                   "accepted": [i.accepted for i in infos],
                   "dH": [round(float(x), 8) for x in dh]}
    c3["gpu_traj_per_s"] = c3["mixed"]["traj_per_s"]  # (the config's FP32 force solves)
    c3["mixed_vs_fp64"] = round(c3["mixed"]["traj_per_s"] / c3["fp64"]["traj_per_s"], 3)
    c3["accept_identical_vs_fp64"] = c3["mixed"]["accepted"] == c3["fp64"]["accepted"]
    c3["max_abs_dH_diff_vs_fp64"] = max(abs(a.dH - b.dH) for a, b in
                                         zip(runs["mixed"], runs["fp64"]))
    c3["_gpu"] = {"acc": [runs["fp64"][0].accepted], "dH": [runs["fp64"][0].dH]}
    if pool is not None:
        pool.submit("rhmc_128", U=Uc.cpu().numpy(), r_a=float(ra), r_b=float(rb), traj=traj0)
    prm.close()
    return c3


def config_measurements(Q, torch, pool, args):
    """The other workloads of the path, each timed on the GPU; their inputs go
    to the CPU oracle jobs (cpu_baselines), which time the same work."""
    import numpy as np
    out = {}

    def sync_time(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, r

    # Nf=2 24x24, beta 4, m0 0.2, hot start seed 1234, leapfrog ns 10 eps 0.1,
    # CG tol 1e-12, 20 trajectories, FP64
    prm = Q.LattParm((24, 24), 4.0)
    U = Q.spinor(prm)
    Q.hot_start(U, prm, 1234)
    U0 = U.cpu().numpy()
    Q.seed(1234)
    dt, infos = sync_time(lambda: [Q.HMC(U, 0.2, 0.1, 10, [], 100000, 1e-12, prm, None)
                                   for _ in range(20)])
    out["nf2_24x24_20traj"] = {"gpu_traj_per_s": round(20 / dt, 2),
                               "accepted": sum(i.accepted for i in infos),
                               "_gpu": {"acc": [i.accepted for i in infos],
                                        "dH": [i.dH for i in infos]}}
    if pool is not None:
        pool.submit("nf2_24", U=U0)
    prm.close()

    # 1024^2, links from 200 pure-gauge OMF4 trajectories at beta 4 (cold start,
    # seed 1234), 16 Gaussian RHS, m0 0.1, CG to 1e-12: FP64 and mixed
    L = 1024
    prm = Q.LattParm((L, L), 4.0)
    U = Q.cold_links(prm)
    Q.seed(1234)
    for _ in range(200):
        Q.HMC(U, 0.1, 10, [], prm, None)
    si = torch.zeros((16, 2, L, L), dtype=torch.complex128, device="cuda")
    for j in range(16):
        Q.fill_pseudofermion(si[j], prm, 1234, 5, 0, j)
    bs = [si[j] for j in range(16)]
    if pool is not None:
        pool.submit("cg16_1024", U=U.cpu().numpy(), b=bs[0].cpu().numpy())
    c2 = {}
    for prec, key in ((Q.PREC_F64, "fp64"), (Q.PREC_MIXED, "mixed")):
        xs = [Q.spinor(prm) for _ in range(16)]
        Q.CG(xs[0], U, bs[0], 0.1, 100000, 1e-12, Q.gamm5Dw_sqr_msq, prm, None, precision=prec)
        dt, its = sync_time(lambda: [Q.CG(x, U, b, 0.1, 100000, 1e-12, Q.gamm5Dw_sqr_msq, prm,
                                          None, precision=prec) for x, b in zip(xs, bs)])
        bps = 384 if prec == Q.PREC_F64 else 192
        # the 16 solves as one set (CG_mrhs: one launch per kernel, shared links);
        # checked bit for bit against one-at-a-time solves, in mixed precision
        # both with reliable updates decided by the host (the same schedule)
        so = torch.zeros_like(si)
        Q.CG_mrhs(so, U, si, 0.1, 100000, 1e-12, prm, None, precision=prec)
        dts, itm = sync_time(lambda: Q.CG_mrhs(so, U, si, 0.1, 100000, 1e-12, prm, None,
                                               precision=prec))
        ref, itr = xs, its
        if prec == Q.PREC_MIXED:
            Q.set_reliable_mode(prm, False)
            ref = [Q.spinor(prm) for _ in range(16)]
            itr = [Q.CG(x, U, b, 0.1, 100000, 1e-12, Q.gamm5Dw_sqr_msq, prm, None,
                        precision=prec) for x, b in zip(ref, bs)]
            so = torch.zeros_like(si)
            itm = Q.CG_mrhs(so, U, si, 0.1, 100000, 1e-12, prm, None, precision=prec)
            Q.set_reliable_mode(prm, True)
        c2[key] = {"seconds_16_rhs": round(dt, 4), "iterations": its,
                   "GBs_cg_kernels": round(bps * L * L * sum(its) / dt / 1e9, 1),
                   "seconds_16_rhs_one_set": round(dts, 4),
                   "one_set_equals_sequential": bool(itm == itr and all(
                       torch.equal(so[j], ref[j]) for j in range(16)))}
    c2["mixed_vs_fp64"] = round(c2["fp64"]["seconds_16_rhs"] / c2["mixed"]["seconds_16_rhs"], 3)
    c2["mixed_vs_fp64_one_set"] = round(c2["fp64"]["seconds_16_rhs_one_set"] /
                                        c2["mixed"]["seconds_16_rhs_one_set"], 3)
    out["cg_16rhs_1024x1024"] = c2
    del bs, si, U
    prm.close()

    out["rhmc_128x128_16poles_4pf"] = rhmc_config(Q, torch, args, pool)

    # master field: pure gauge 8192^2, beta 11.25, OMF4 ns 20 eps 0.025
    L = 8192
    prm = Q.LattParm((L, L), 11.25)
    U = Q.spinor(prm)
    Q.hot_start(U, prm, 4321)
    Q.seed(4321)
    Q.HMC(U, 0.025, 20, [], prm, None)
    if pool is not None:
        pool.submit("mf_8192", U=U.cpu().numpy(), traj=Q.rng_state()[2])
    dt, infos = sync_time(lambda: [Q.HMC(U, 0.025, 20, [], prm, None) for _ in range(2)])
    out["pure_gauge_8192x8192"] = {
        "gpu_traj_per_s": round(2 / dt, 3), "Q": round(Q.Qtop(U, prm), 6),
        "plaquette": Q.Plaquette(U, prm),
        "_gpu": {"acc": [infos[0].accepted], "dH": [infos[0].dH]}}
    del U
    prm.close()
    torch.cuda.empty_cache()

    # Nf=2 4096^2, beta 11.25, m0 0.05, links after 3 pure-gauge trajectories
    # from a cold start; one OMF4 trajectory ns 2 eps 0.05, CG tol 1e-20 (a
    # leapfrog trajectory of the same length violates the energy by ~4e4)
    L = 4096
    prm = Q.LattParm((L, L), 11.25)
    U = Q.cold_links(prm)
    Q.seed(55, chain=2)
    for _ in range(3):
        Q.HMC(U, 0.025, 20, [], prm, None)
    U0 = U.clone()
    c6 = {}
    for prec, key in ((Q.PREC_MIXED, "mixed"), (Q.PREC_F64, "fp64")):
        U.copy_(U0)
        Q.seed(55)
        dt, info = sync_time(lambda: Q.HMC(U, 0.05, 0.05, 2, [], 100000, 1e-20, prm, None,
                                           integrator=Q.OMF4(), precision=prec, acc_tol=1e-20))
        c6[key] = {"seconds": round(dt, 3), "cg_iters": info.cg_iters, "dH": info.dH,
                   "accepted": info.accepted}
    c6["mixed_vs_fp64"] = round(c6["fp64"]["seconds"] / c6["mixed"]["seconds"], 3)
    c6["abs_dH_diff"] = abs(c6["mixed"]["dH"] - c6["fp64"]["dH"])
    out["nf2_4096x4096_m0.05"] = c6
    del U, U0
    prm.close()
    torch.cuda.empty_cache()
    for k, v in out.items():
        log(f"config {k}: " + json.dumps({a: b for a, b in v.items() if a != "_gpu"}))
    return out


def pmc_traffic(kernel, L):
    """HBM bytes per launch of `kernel` on an L x L lattice, from the committed
    rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE per dispatch, each
    scaled by what the counter reads for a copy kernel of known size), or None
    when that summary does not cover the kernel at this size."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    e = d.get("kernels", {}).get(kernel)
    if not e or e.get("lattice") != [L, L]:
        return None
    return e.get("hbm_bytes_per_launch")


def acceptance_comparison(Q, torch, args, prm, U, U0, seed):
    """The same 1024^2 start and physics with leapfrog tau = 1 in 50 steps
    (eps 0.02).  At the headline's eps 0.1, ns 10 the energy violation is about
    500 and every trajectory is rejected, so identical accept/reject sequences
    say little there; here |dH| is O(1) and about half the trajectories are
    accepted.  Mixed and FP64 run the same random streams as a Markov chain;
    reports both rates, whether the accept/reject sequences are identical and
    the largest |dH_mixed - dH_fp64|."""
    n, ns, eps = args.accept_traj, 50, 0.02
    out = {"integrator": "leapfrog", "ns": ns, "eps": eps, "trajectories": n}
    seqs, dhs = {}, {}
    for prec, key in ((Q.PREC_MIXED, "mixed"), (Q.PREC_F64, "fp64")):
        U.copy_(U0)
        Q.seed(seed, chain=5)
        torch.cuda.synchronize()
        t = time.perf_counter()
        inf = [Q.HMC(U, args.m0, eps, ns, [], 100000, args.tol, prm, None, precision=prec,
                     acc_tol=args.tol) for _ in range(n)]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        out[key + "_traj_per_s"] = round(n / dt, 4)
        seqs[key] = [i.accepted for i in inf]
        dhs[key] = [i.dH for i in inf]
    out["speedup_vs_fp64"] = round(out["mixed_traj_per_s"] / out["fp64_traj_per_s"], 4)
    out["acceptance"] = sum(seqs["mixed"]) / n
    out["accept_identical_vs_fp64"] = seqs["mixed"] == seqs["fp64"]
    out["max_abs_dH_diff_vs_fp64"] = max(abs(a - b) for a, b in zip(dhs["mixed"], dhs["fp64"]))
    out["max_abs_dH"] = max(abs(a) for a in dhs["fp64"])
    log(f"acceptance comparison (leapfrog ns {ns} eps {eps}, {n} traj): mixed "
        f"{out['mixed_traj_per_s']} traj/s, fp64 {out['fp64_traj_per_s']} traj/s, acceptance "
        f"{out['acceptance']:.2f}, identical {out['accept_identical_vs_fp64']}, max ddH "
        f"{out['max_abs_dH_diff_vs_fp64']:.2e}")
    return out




# ---------------------------------------------------------------------------
# CPU oracle baselines.  Every workload is timed as whole trajectories (or a
# whole solve) of the FP64 CPU oracle, never scaled: at one thread in worker
# processes started before the GPU is initialised (one core each), which run
# while the GPU legs do, and at all host threads in this process once the GPU
# legs are done.  The GPU legs hand each job its inputs through files.
# ---------------------------------------------------------------------------
def _cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


class CpuPool:
    """Oracle workers: (job, threads) pairs, each pinned to cores of its own;
    once the first job is submitted the parent keeps to the remaining cores
    until collect()."""

    def __init__(self, jobs):
        import subprocess
        import tempfile
        self.dir = tempfile.mkdtemp(prefix="qed2d_cpu_")
        self.cores = sorted(os.sched_getaffinity(0))
        self.procs = {}
        free = list(self.cores)
        for name, threads in jobs:
            threads = max(1, min(threads, len(free) - 2))
            mine, free = free[-threads:], free[:-threads]
            key = f"{name}@{threads}"
            env = dict(os.environ, OMP_NUM_THREADS=str(threads))
            with open(os.path.join(self.dir, key + ".log"), "w") as lf:
                self.procs[key] = subprocess.Popen(
                    [sys.executable, os.path.abspath(__file__), "--cpu-worker", key, self.dir,
                     ",".join(str(c) for c in mine)], env=env, stdout=lf,
                    stderr=subprocess.STDOUT)
        self.free = set(free)
        self.inputs = {}

    def submit(self, name, **kw):
        import numpy as np
        # from the first job on, the GPU legs keep off the workers' cores
        os.sched_setaffinity(0, self.free)
        self.inputs[name] = kw
        if not any(k.split("@")[0] == name for k in self.procs):
            return
        arrays = {k: v for k, v in kw.items() if isinstance(v, np.ndarray)}
        params = {k: v for k, v in kw.items() if not isinstance(v, np.ndarray)}
        base = os.path.join(self.dir, name)
        np.savez(base + ".npz", **arrays)
        with open(base + ".json.tmp", "w") as f:
            json.dump(params, f)
        os.replace(base + ".json.tmp", base + ".in.json")

    def collect(self, timeout):
        """{job: {threads: result}} of the workers that finished."""
        res = {}
        os.sched_setaffinity(0, set(self.cores))
        t_end = time.time() + timeout
        for key, p in self.procs.items():
            try:
                p.wait(timeout=max(1.0, t_end - time.time()))
            except Exception:
                p.kill()
                p.wait()
                log(f"cpu worker {key} did not finish within {timeout:.0f}s")
                continue
            path = os.path.join(self.dir, key + ".out.json")
            if os.path.exists(path):
                with open(path) as f:
                    r = json.load(f)
                res.setdefault(key.split("@")[0], {})[r["threads"]] = r
            else:
                log(f"cpu worker {key}: " + open(os.path.join(self.dir, key + ".log")).read())
        return res


def cpu_worker(key, workdir, cores):
    """An oracle worker (job@threads): waits for its job's inputs, runs it on
    its cores, writes the result; gives up if the parent process goes away."""
    import numpy as np
    os.sched_setaffinity(0, {int(c) for c in cores.split(",")})
    name = key.split("@")[0]
    parent = os.getppid()
    base = os.path.join(workdir, name)
    t0 = time.time()
    while not os.path.exists(base + ".in.json"):
        if os.getppid() != parent or time.time() - t0 > 3600:
            return
        time.sleep(0.2)
    with open(base + ".in.json") as f:
        params = json.load(f)
    with np.load(base + ".npz") as z:
        arrays = {k: z[k] for k in z.files}
    from oracle import qed2d_oracle as O
    r = run_cpu_job(name, dict(params, **arrays))
    r["threads"] = O.omp_threads()
    out = os.path.join(workdir, key)
    with open(out + ".out.json.tmp", "w") as f:
        json.dump(r, f)
    os.replace(out + ".out.json.tmp", out + ".out.json")


def run_cpu_job(name, kw):
    """Times one oracle job on the calling process's threads."""
    import numpy as np
    from oracle import qed2d_oracle as O
    U = np.array(kw["U"], copy=True)
    t = time.perf_counter()
    if name == "nf2_1024":
        L = U.shape[-1]
        p = O.latt(L, L, kw["beta"])
        hi, hf, a, it, rc = O.hmc_nf2(U, kw["m0"], kw["eps"], kw["ns"], 100000, kw["tol"], p,
                                      kw["seed"], 0)
        r = {"trajectories": 1, "cg_iters": it, "dH": hf - hi, "accepted": a}
    elif name == "nf2_24":
        p = O.latt(24, 24, 4.0)
        ref = [O.hmc_nf2(U, 0.2, 0.1, 10, 100000, 1e-12, p, 1234, t) for t in range(20)]
        r = {"trajectories": 20, "acc": [x[2] for x in ref], "dH": [x[1] - x[0] for x in ref]}
    elif name == "cg16_1024":
        _, it, st = O.cg(U, kw["b"], 0.1, 100000, 1e-12)
        r = {"solves": 1, "cg_iters": it}
    elif name == "rhmc_128":
        z = O.zolotarev(16, kw["r_a"], kw["r_b"])
        p = O.latt(128, 128, 4.0)
        hi, hf, a, rc = O.hmc_rhmc(U, [0.1] * 4, [z] * 4, 0.083, 12, 100000, 1e-16, p, 77,
                                   kw["traj"], integrator=1)
        r = {"trajectories": 1, "acc": [a], "dH": [hf - hi]}
    elif name == "mf_8192":
        L = U.shape[-1]
        p = O.latt(L, L, 11.25)
        hi, hf, a = O.hmc_pure(U, 0.025, 20, p, 4321, kw["traj"])
        r = {"trajectories": 1, "acc": [a], "dH": [hf - hi]}
    else:
        raise ValueError(name)
    r["seconds"] = time.perf_counter() - t
    return r


def cpu_jobs(args):
    """The oracle workers, started before the GPU is initialised and fed once
    the timed headline legs are done: the cheap jobs at one thread, and the
    RHMC trajectory on most of the remaining cores, while the GPU's other legs
    run; the rest run at all threads afterwards (cpu_baselines).  The 8192^2
    and RHMC trajectories at one thread (minutes each) only with --cpu-full."""
    ncores = len(os.sched_getaffinity(0))
    jobs = [("nf2_1024", 1)]
    if not args.no_configs:
        jobs += [("nf2_24", 1), ("cg16_1024", 1)]
        if args.cpu_full:
            jobs += [("mf_8192", 1), ("rhmc_128", 1)]
        jobs.append(("rhmc_128", max(1, ncores - len(jobs) - 2)))
    return jobs


def cpu_baselines(pool, configs, args):
    """The headline's cpu_baseline (one whole 1024^2 trajectory of the same
    workload at all threads and at one thread) and each config's CPU fields,
    with the host's CPU model; no figure is scaled."""
    from oracle import qed2d_oracle as O
    threads = O.omp_threads()
    done = pool.collect(timeout=900)
    model = _cpu_model()
    for name, kw in pool.inputs.items():
        have = done.setdefault(name, {})
        if not any(t > 1 for t in have):
            r = run_cpu_job(name, kw)
            r["threads"] = threads
            have[threads] = r
        log(f"cpu oracle {name}: " + ", ".join(f"{t} threads {r['seconds']:.2f}s"
                                               for t, r in sorted(have.items())))

    def fields(name, unit="trajectories"):
        have = done[name]
        many = max(have)

        def one(r):
            return {"threads": r["threads"], "seconds": round(r["seconds"], 3),
                    "per_s": round(r.get(unit, 1) / r["seconds"], 6)}
        return {"cpu_model": model, "host_threads": len(pool.cores),
                unit + "_timed": have[many].get(unit, 1), "scale_factor": None,
                "all_threads": one(have[many]) if many > 1 else None,
                "one_thread": one(have[1]) if 1 in have else None}

    if configs:
        for key, name, unit in (("nf2_24x24_20traj", "nf2_24", "trajectories"),
                                ("cg_16rhs_1024x1024", "cg16_1024", "solves"),
                                ("rhmc_128x128_16poles_4pf", "rhmc_128", "trajectories"),
                                ("pure_gauge_8192x8192", "mf_8192", "trajectories")):
            c = configs.get(key)
            if c is None or name not in done:
                continue
            c["cpu"] = fields(name, unit)
            g = c.pop("_gpu", None)
            r = done[name][max(done[name])]
            if g and "acc" in r:
                n = len(g["acc"])
                c["cpu_accept_identical"] = g["acc"] == r["acc"][:n]
                c["cpu_max_abs_ddH"] = max(abs(a - b) for a, b in zip(g["dH"], r["dH"][:n]))
            if name == "cg16_1024":
                c["cpu"]["cg_iters"] = r["cg_iters"]
            if 1 not in done[name]:
                c["cpu"]["one_thread_note"] = "timed with bench.py --cpu-full"
    base = fields("nf2_1024")
    h = done["nf2_1024"][max(done["nf2_1024"])]
    return dict({"value": round(1.0 / h["seconds"], 6), "unit": "traj/s",
                 "cores": h["threads"], "kind": "port",
                 "sample": f"one whole Nf=2 HMC trajectory of the headline workload "
                           f"({args.L}x{args.L}, the same start links, leapfrog ns {args.ns} "
                           f"eps {args.eps}, FP64 CG to tol {args.tol}: {h['cg_iters']} "
                           f"iterations) on the FP64 C oracle (OpenMP) at {h['threads']} "
                           f"threads, not scaled; one_thread: the same trajectory on one "
                           f"core"}, **base)


# ---------------------------------------------------------------------------
# rocprofv3 summaries: python bench.py --rocprof-summary KT.db FETCH.db WRITE.db
#   [--L 1024] [--json profiles/pmc_traffic.json]
# KT.db from `rocprofv3 --kernel-trace --stats`, FETCH.db / WRITE.db from
# `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs of the same bench
# command.  Prints a markdown summary; with --json also writes the per-launch
# HBM traffic of the kernel classes that pmc_traffic() reads.  The counters
# are scaled by what they read for the STREAM copy kernel (1 GiB read, 1 GiB
# written per dispatch) in the same runs.
# ---------------------------------------------------------------------------
KERNEL_CLASS = [("k_cg1v", "cg_k1_f32"), ("k_cg2v", "cg_k2_f32"), ("k_cg1<float", "cg_k1_f32"),
                ("k_cg2<float", "cg_k2_f32"), ("k_cg1<double", "cg_k1_f64"),
                ("k_cg2<double", "cg_k2_f64"), ("k_q_apply<double", "q_f64"),
                ("k_force_mom", "force_mom"), ("k_updt_fld", "updt_fld")]


def _short(name):
    import re
    name = re.sub(r"\(.*", "", name)
    return name.replace("void ", "").replace("qed2d::", "")


def rocprof_summary(argv):
    import sqlite3
    import numpy as np
    ap = argparse.ArgumentParser(prog="bench.py --rocprof-summary")
    ap.add_argument("kt")
    ap.add_argument("fetch", nargs="?", default="")
    ap.add_argument("write", nargs="?", default="")
    ap.add_argument("--L", type=int, default=1024)
    ap.add_argument("--json", default="")
    ap.add_argument("--by-grid", action="store_true",
                    help="one row per (kernel, grid size): shapes that share a kernel")
    ap.add_argument("--windows", type=int, default=0,
                    help="trajectories per timed run: GPU busy and idle time per trajectory "
                         "inside the windows the timed runs mark with one-element fills")
    a = ap.parse_args(argv)
    con = sqlite3.connect(a.kt)
    by = {}

    def key(n, g):
        return f"{_short(n)} [grid {g}]" if a.by_grid else _short(n)
    for n, d, g in con.execute("select name, duration, grid_x from kernels"):
        by.setdefault(key(n, g), []).append(d / 1000.0)
    rows = sorted(((n, np.array(d)) for n, d in by.items()), key=lambda r: -r[1].sum())
    total = sum(d.sum() for _, d in rows)
    print("| kernel | calls | total us | mean us | median us (calls > 5 us) | % time |")
    print("|---|---|---|---|---|---|")
    for n, d in rows[:40 if a.by_grid else 16]:
        w = d[d > 5.0]
        print(f"| {n} | {len(d)} | {d.sum():.0f} | {d.mean():.2f} | "
              f"{np.median(w) if len(w) else 0.0:.2f} | {100 * d.sum() / total:.1f} |")

    if a.windows:
        import re
        ks = list(con.execute("select name, start, end from kernels order by start"))
        # (int16 fills: nothing else in the run fills a 16-bit tensor)
        marks = [i for i, (n, _, _) in enumerate(ks) if re.search(r"FillFunctor<short>", n)]
        print()
        print(f"Timed windows (between the fills that mark them; {a.windows} trajectories each):")
        print()
        print("| window | kernels | span ms | busy ms | idle ms | idle ms per trajectory |")
        print("|---|---|---|---|---|---|")
        per = []
        for w, (i0, i1) in enumerate(zip(marks[0::2], marks[1::2])):
            inner = ks[i0 + 1:i1]
            if not inner:
                continue
            span = (ks[i1][1] - ks[i0][2]) / 1e6
            busy = sum(e - s0 for _, s0, e in inner) / 1e6
            print(f"| {w} | {len(inner)} | {span:.2f} | {busy:.2f} | {span - busy:.2f} | "
                  f"{(span - busy) / a.windows:.3f} |")
            by_k = {}
            for n, s0, e in inner:
                t = by_k.setdefault(_short(n), [0, 0.0])
                t[0] += 1
                t[1] += (e - s0) / 1e3
            per.append((w, by_k))
        for w, by_k in per:
            print()
            print(f"Window {w}, per trajectory: launches and GPU time by kernel")
            print()
            print("| kernel | launches | us | mean us |")
            print("|---|---|---|---|")
            for n, (cnt, us) in sorted(by_k.items(), key=lambda kv: -kv[1][1])[:14]:
                print(f"| {n} | {cnt / a.windows:.1f} | {us / a.windows:.0f} | {us / cnt:.2f} |")

    if not (a.fetch and a.write):
        return

    def counter(db, cname):
        c = sqlite3.connect(db)
        out = {}
        for n, v, g in c.execute("select kernel_name, value, grid_size_x from "
                                 "counters_collection where counter_name=?", (cname,)):
            out.setdefault(key(n, g), []).append(v)
        return {n: np.array(v) for n, v in out.items()}
    F, W = counter(a.fetch, "FETCH_SIZE"), counter(a.write, "WRITE_SIZE")
    gib_kib = float(1 << 20)
    cp = [n for n in F if n.startswith("k_stream_copy")]
    if not cp or cp[0] not in W:
        sys.exit("the counter runs hold no k_stream_copy dispatch to calibrate against")
    fcal = gib_kib / np.median(F[cp[0]])
    wcal = gib_kib / np.median(W[cp[0]])
    print()
    print(f"Calibration (STREAM copy, 1 GiB read + 1 GiB written per dispatch): FETCH_SIZE x "
          f"{fcal:.3f}, WRITE_SIZE x {wcal:.3f}")
    print()
    print("| kernel | dispatches | FETCH_SIZE KiB (median, dispatches > 1 MiB) | "
          "WRITE_SIZE KiB (median) | HBM MB per launch (scaled) |")
    print("|---|---|---|---|---|")
    # a class's traffic: the mean over its kernels' medians, weighted by their
    # working dispatches (instances of one class, such as the FP32 K1 of even
    # and of odd k, alternate)
    acc = {}
    for n in sorted(set(F) | set(W), key=lambda n: -np.sum(F.get(n, [0]))):
        f = F.get(n, np.zeros(1))
        w = W.get(n, np.zeros(1))
        fb, wb = f[f > 1024], w[w > 1024]
        fm = float(np.median(fb) if len(fb) else np.median(f))
        wm = float(np.median(wb) if len(wb) else np.median(w))
        mb = (fm * fcal + wm * wcal) * 1024 / 1e6
        print(f"| {n} | {len(f)} | {fm:.0f} | {wm:.0f} | {mb:.1f} |")
        for pat, cls in KERNEL_CLASS:
            if n.startswith(pat):
                # (a class is the first pattern's kernels: k_cg1v before k_cg1<float)
                if cls in acc and not acc[cls]["pat"] == pat:
                    break
                e = acc.setdefault(cls, {"pat": pat, "names": [], "nw": 0, "f": 0.0, "w": 0.0})
                nw = max(len(fb), 1)
                e["names"].append(n)
                e["nw"] += nw
                e["f"] += fm * nw
                e["w"] += wm * nw
                break
    traffic = {}
    for cls, e in acc.items():
        fm, wm = e["f"] / e["nw"], e["w"] / e["nw"]
        traffic[cls] = {"kernel": " + ".join(e["names"]), "lattice": [a.L, a.L],
                        "fetch_kib": fm, "write_kib": wm, "fetch_scale": fcal,
                        "write_scale": wcal,
                        "hbm_bytes_per_launch": int(round((fm * fcal + wm * wcal) * 1024))}
    if len(acc):
        print()
        print("| class | kernels | HBM MB per launch (dispatch-weighted) |")
        print("|---|---|---|")
        for cls, t in traffic.items():
            print(f"| {cls} | {t['kernel']} | {t['hbm_bytes_per_launch'] / 1e6:.1f} |")
    if a.json:
        with open(a.json, "w") as fo:
            json.dump({"source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of bench.py",
                       "kernels": traffic}, fo, indent=1)


if __name__ == "__main__":
    main()
