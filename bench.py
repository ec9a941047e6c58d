#!/usr/bin/env python
"""Benchmark of the ATV continuous max-flow on one GPU.

Workload (one step): the grid search of the reference experiment on one
128^3 helix phantom (tube radius 6, helix radius 32, pitch 48, noise sigma
1.5): 25 nt_weightings x 17 nt_decays = 425 ATV solves (a = 10^-1.5) of 500
iterations each, run as one batched solve on device-resident inputs. The
shared inputs (frames from Frangi -> adjustment -> GVF -> completion,
capacities, |grad I|^2) are prepared once, outside the timed region.

Reported: whole-job voxel-iterations/s of the batched grid; every cell's
fast Dice against the noise-free mask (counted on the device); the same
solves run one at a time (device-resident, and through the numpy drop-in
API); a 256^3 single solve (the bandwidth-scale plain workload); serving
latency (p50/p99) and throughput under Poisson arrivals of mixed-size cuts;
the HBM roofline of the max-flow iteration (84 algorithmic bytes per
voxel-iteration); and the CPU oracle timed on a bounded sample on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BYTES_PER_VOXEL_ITER = 84.0      # algorithmic bytes of one max-flow iteration
HBM_PEAK_GBS = 8000.0            # MI355X


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed(torch, fn, reps=1):
    """Run fn reps times between HIP events on the current stream; returns seconds."""
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--iters", type=int, default=500, help="max-flow iterations per solve")
    ap.add_argument("--solves", type=int, default=425, help="grid cells per step (<= 425)")
    ap.add_argument("--plain-solves", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-256", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=150)
    ap.add_argument("--serve-qps", type=float, default=8.0)
    ap.add_argument("--serve-requests", type=int, default=40)
    ap.add_argument("--no-serve", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one GPU and this build has no "
              "multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    from miccai2016_amd import phantom as ph
    from miccai2016_amd.device import default_context
    from miccai2016_amd.pipeline import PipelineConfig, cut_parameters, prepare
    from miccai2016_amd.processing.continuous_cut_anisotropic_gpu import (
        CmfCutQAnisotropicGpu, initial_u)

    dev = default_context(0)
    cfg = PipelineConfig()
    cfg.cut_iters = args.iters
    spec = ph.GRID_128
    log(f"phantom {spec.shape} sigma={spec.sigma}")
    img, truth = ph.make_helix(spec)
    t0 = time.time()
    prep = prepare(img, spec.sigma, cfg, dev)
    torch.cuda.synchronize()
    t_prep = time.time() - t0
    log(f"prepare (Frangi, adjustment, GVF, frames, capacities): {t_prep:.2f} s")

    w_all, d_all = ph.grid_pairs()
    S = max(1, min(args.solves, len(w_all)))
    w, d = w_all[:S].astype(np.float32), d_all[:S].astype(np.float32)
    eigs = (1, cfg.a_atv, cfg.a_atv)
    n = img.size
    cutter = CmfCutQAnisotropicGpu(cut_parameters(cfg), dev)
    q_d = dev.upload(prep.quats).reshape(-1)
    gms_d = dev.upload(prep.gms).reshape(-1)
    cs_d, ct_d = dev.upload(prep.cs).reshape(-1), dev.upload(prep.ct).reshape(-1)
    u0 = dev.upload(initial_u(img)).reshape(1, -1)
    u = torch.empty((S, n), dtype=torch.float32, device=dev.device)

    def step():
        u.copy_(u0.expand(S, n))
        cutter.execute_batched_device(u, S, w, d, q_d, gms_d, cs_d, ct_d, img.shape, eigs)

    for k in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"warmup step {k + 1}/{args.warmup} done")
    wall0 = time.time()
    secs = 0.0
    for k in range(args.steps):
        secs += timed(torch, step)
        log(f"timed step {k + 1}/{args.steps}: {secs / (k + 1):.3f} s/step")
    wall = time.time() - wall0
    vox_iters = float(S) * n * args.iters
    value = vox_iters * args.steps / secs
    ms_step = secs / args.steps * 1e3
    log(f"batched grid: {value:.4e} voxel-iters/s ({ms_step:.1f} ms/step, wall {wall:.2f} s)")

    # ---- quality of every grid cell against the noise-free mask, on the device
    from miccai2016_amd.evaluation import DeviceDice
    grid_dice = DeviceDice(truth, dev)(u, S)
    best = int(np.nanargmax(grid_dice))
    log(f"grid: best fast Dice {grid_dice[best]:.4f} at w={w_all[best]:.4g}, decay={d_all[best]:.4g}")

    # ---- the same solves one at a time (device-resident) and through the API
    P = max(1, min(args.plain_solves, S))
    pick = np.linspace(0, S - 1, P).astype(int)
    ub = u.detach().cpu().numpy()
    plain_cut = CmfCutQAnisotropicGpu(cut_parameters(cfg), dev)
    alphas = [ph.nonterminal_weight(prep.gms, w_all[j], d_all[j]).astype(np.float32) for j in pick]
    eig_d = q_d
    ud = dev.empty(n)
    t_dev = 0.0
    dices = []
    for j, a in zip(pick, alphas):
        a_d = dev.upload(a).reshape(-1)

        def one():
            ud.copy_(u0.reshape(-1))
            plain_cut.execute_device(ud, eig_d, a_d, cs_d, ct_d, const_sqevals=np.sqrt(eigs),
                                     shape=img.shape)
        t_dev += timed(torch, one)
        up = ud.detach().cpu().numpy()
        dices.append(ph.dice_between(ub[j], up))
    per_dev = t_dev / P
    t_api = 0.0
    napi = min(P, 3)
    for j, a in zip(pick[:napi], alphas[:napi]):
        t1 = time.time()
        plain_cut.execute(img, prep.quats, eigs, a, prep.cs, prep.ct, verbose=False)
        t_api += time.time() - t1
    per_api = t_api / napi
    plain_rate = n * args.iters / per_dev
    log(f"plain device-resident: {per_dev * 1e3:.2f} ms/solve; numpy API: {per_api * 1e3:.2f} "
        f"ms/solve; min Dice batched vs plain {min(dices):.5f}")

    # ---- 256^3 single solve (bandwidth scale): the 256^3 helix (tube 8, helix 64,
    # pitch 96, noise 1.0) generated on the device with its centre-line frames
    big = None
    if not args.no_256:
        import ctypes as C
        from miccai2016_amd import _lib
        sp256 = ph.PLAIN_256
        shape = sp256.shape
        nb = int(np.prod(shape))
        L, h, st = dev.lib, dev.handle, dev.stream
        img_b, qb = dev.empty(nb), dev.empty(3 * nb)
        cs_b, ct_b, al_b, ub0 = dev.empty(nb), dev.empty(nb), dev.empty(nb), dev.empty(nb)
        dev.check(L.mh_helix_phantom(h, _lib.dims3(shape), sp256.r_tube, sp256.r_helix,
                                     sp256.pitch, sp256.sigma, C.c_uint64(42), dev.ptr(img_b),
                                     None, dev.ptr(qb), st), "mh_helix_phantom")
        dev.check(L.mh_complete_frames(h, dev.ptr(qb), dev.ptr(qb), nb, st), "mh_complete_frames")
        dev.check(L.mh_terminal_capacities(h, dev.ptr(img_b), nb, sp256.sigma, 1e-9,
                                           dev.ptr(cs_b), dev.ptr(ct_b), st), "capacities")
        dev.check(L.mh_grad_mag_sq(h, dev.ptr(img_b), _lib.dims3(shape), dev.ptr(al_b), st),
                  "mh_grad_mag_sq")
        dev.check(L.mh_nonterminal_weight(h, dev.ptr(al_b), nb, 10 ** -1.625, 10 ** 1.5,
                                          dev.ptr(al_b), st), "mh_nonterminal_weight")
        wk = dev.empty(int(L.mh_initial_u_work_floats(nb)))
        dev.check(L.mh_initial_u(h, dev.ptr(img_b), nb, dev.ptr(ub0), dev.ptr(wk), st),
                  "mh_initial_u")
        ub_ = torch.empty(nb, device=dev.device)
        p256 = cut_parameters(cfg)
        p256.num_iter = 1000
        c256 = CmfCutQAnisotropicGpu(p256, dev)

        def solve256():
            ub_.copy_(ub0)
            c256.execute_device(ub_, qb, al_b, cs_b, ct_b, const_sqevals=np.sqrt(eigs), shape=shape)
        solve256()
        t256 = timed(torch, solve256)
        rate256 = nb * 1000 / t256
        big = {"voxel_iters_per_s": rate256, "ms_per_solve": t256 * 1e3,
               "achieved_GBs": rate256 * BYTES_PER_VOXEL_ITER / 1e9}
        log(f"256^3 x 1000 iterations: {t256:.3f} s -> {rate256:.4e} voxel-iters/s, "
            f"{big['achieved_GBs']:.0f} GB/s algorithmic")

    # ---- serving: Poisson arrivals of synthetic helix cuts of mixed sizes
    serve = None
    if not args.no_serve:
        from miccai2016_amd.serving import MaxFlowServer, poisson_load, synthetic_request
        del u, ub
        torch.cuda.empty_cache()
        sp = cut_parameters(cfg)
        srv = MaxFlowServer(sp, chunk=50).start()
        log(f"serving {args.serve_requests} requests at {args.serve_qps} QPS "
            f"(budget {srv.budget / 2 ** 30:.1f} GiB)")
        tickets = poisson_load(srv, synthetic_request, args.serve_qps, args.serve_requests,
                               seed=11)
        srv.stop(drain=True, timeout=600)
        st = srv.stats()
        errors = [t.error for t in tickets if t.error is not None]
        dice = [t.dice for t in tickets if t.dice is not None]
        serve = {"qps": args.serve_qps, "requests": args.serve_requests,
                 "completed": st["completed"], "errors": len(errors),
                 "p50_s": st.get("p50_s"), "p99_s": st.get("p99_s"), "mean_s": st.get("mean_s"),
                 "voxel_iters_per_s": st.get("voxel_iters_per_s"),
                 "peak_resident_jobs": srv.peak_resident,
                 "peak_reserved_GiB": srv.peak_reserved / 2 ** 30,
                 "budget_GiB": srv.budget / 2 ** 30, "chunk": srv.chunk,
                 "median_dice_vs_truth": float(np.median(dice)) if dice else None,
                 "workload": "side log-U[64,320], aspect <= 2:1, 300-1000 iterations, "
                             "helix/noise as the ragged set, (w, decay) from the grid"}
        log(f"serving: p50 {serve['p50_s']:.3f} s, p99 {serve['p99_s']:.3f} s, "
            f"{serve['voxel_iters_per_s']:.3e} voxel-iters/s, peak {srv.peak_resident} resident")

    # ---- CPU oracle on a bounded sample of the same workload
    cpu = None
    if not args.no_cpu:
        from oracle import ref32
        threads = len(os.sched_getaffinity(0))
        ref32.set_num_threads(threads)
        a0 = alphas[0]
        t1 = time.time()
        ref32.cmf_solve(img, prep.quats, eigs, a0, prep.cs, prep.ct, steps=cfg.steps, cc=cfg.cc,
                        num_iter=args.cpu_iters, update=cfg.cut_update, error_bound=0)
        tc = time.time() - t1
        cpu = {"value": n * args.cpu_iters / tc, "unit": "voxel-iters/s", "cores": threads,
               "kind": "port",
               "sample": f"1 of the 425 solves, {args.cpu_iters} of 500 iterations, 128^3, "
                         f"C/OpenMP restatement of the reference kernels ({tc:.1f} s)"}
        log(f"cpu oracle: {cpu['value']:.4e} voxel-iters/s on {threads} threads")

    achieved = value * BYTES_PER_VOXEL_ITER / 1e9
    line = {
        "metric": "ATV max-flow voxel-iters/s, batched vs plain at Dice>=0.999; p50/p99 latency",
        "value": value,
        "unit": "voxel-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic helix phantom (seeded)",
        "config": {"workload": f"batched grid search, {S} ATV solves x {args.iters} iterations on "
                               f"one 128^3 helix phantom (sigma 1.5)",
                   "solves": S, "iterations": args.iters, "volume": list(img.shape)},
        "plain": {"voxel_iters_per_s": plain_rate, "ms_per_solve_device": per_dev * 1e3,
                  "ms_per_solve_numpy_api": per_api * 1e3,
                  "sequential_425_s_device": per_dev * 425,
                  "sequential_425_s_numpy_api": per_api * 425,
                  "batched_speedup_vs_device": value / plain_rate,
                  "batched_speedup_vs_numpy_api": value / (n * args.iters / per_api),
                  "min_dice_batched_vs_plain": float(min(dices))},
        "grid_quality": {"best_fast_dice": float(grid_dice[best]),
                         "best_weighting": float(w_all[best]), "best_decay": float(d_all[best])},
        "plain_256": big,
        "serving": serve,
        "prepare_s": t_prep,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "max-flow iteration (p-step + u-step), 84 B per voxel-iteration"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
