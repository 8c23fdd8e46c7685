#!/usr/bin/env python3
"""Benchmark of the MI355X gravity kick: particle-kicks/s on one GPU.

Workload (default): 256^3 = 16,777,216 particles, 70% in 2000 NFW halos
(masses log-uniform over 3 decades, concentration 5-15, truncated at r200) and
30% uniform, seed 42; theta = 0.55, Ewald on, expansion order 7, softening
(1/80)(693/256) N^(-1/3), GM = 1/N; first kick of the top level (kick_params
of src/driver.cpp:481-508 with tau = 0).  The leaf bucket is autotuned on the
resident box (a timed kick at each of 64, 96, 128, 160, 192; the reference's
default is 128) unless --bucket names one.

A step is one kick of every particle of the resident box: the walk, all
interactions (M2L, Ewald, P2L, P2P, M2P, L2P) and the kick update.  Inputs are
resident in HBM when the timed region starts; the tree is built once before.

Prints one JSON line.  `roofline` describes the dominant kernel (P2P, the
k_pp launches of one kick: 21 flops per source-sink pair, the reference's own
weight, src/gravity.cu:343) from HIP events recorded on the kick's stream;
`accuracy` is the rms and 99th-percentile relative force error of the kicked
box against the FP64 direct Ewald sum on a seeded sample of sinks;
`cpu_baseline` times the reference's CPU SIMD kick path (kick() and the
cpu_gravity_* kernels of src/kick.cpp and src/gravity.cpp, restated in AVX2
FP32 with OpenMP threads) on a bounded sample of the same box on this host,
and compares its forces on the sample with the GPU kick's.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 figure
# the reference's flop weight of one P2P pair (src/gravity.cu:343)
W_PP = 21.0


def make_box(args):
    from cosmictiger_amd import workloads
    n = args.n
    if args.workload == "uniform":
        x, y, z = workloads.uniform_box(n, args.seed)
    else:
        x, y, z = workloads.nfw_box(n, args.seed, clustered_fraction=0.7, nhalos=2000)
    return x, y, z


def direct_reference(box, x, y, z, h, n, nsinks, seed):
    """FP64 direct Ewald forces (device, ct_direct_ewald) on a seeded sample of
    sinks: (sink indices, reference) for accuracy()."""
    sinks = np.random.default_rng(seed).choice(n, min(nsinks, n), replace=False)
    return sinks, box.direct_ewald(x[sinks], y[sinks], z[sinks], h)


def accuracy(box, x, y, z, h, n, nsinks, seed, ref=None):
    """Relative force error of the last kick against the FP64 direct Ewald sum
    on a seeded sample of sinks."""
    out = box.read()
    inv = np.empty(n, dtype=np.int64)
    inv[out["id"]] = np.arange(n)
    sinks, ref = ref if ref is not None else direct_reference(box, x, y, z, h, n, nsinks, seed)
    gm = 1.0 / n
    g = np.stack([out["gx"][inv][sinks], out["gy"][inv][sinks], out["gz"][inv][sinks]], 1).astype(np.float64)
    gr = gm * np.stack([ref["gx"], ref["gy"], ref["gz"]], 1)
    nr = np.linalg.norm(gr, axis=1)
    err = np.linalg.norm(g - gr, axis=1) / nr
    merr = np.abs(np.linalg.norm(g, axis=1) - nr) / nr  # the reference's |g| metric (src/analytic.cpp:88)
    return {"sinks": int(len(sinks)), "rms": float(np.sqrt(np.mean(err ** 2))), "p50": float(np.percentile(err, 50)),
            "p99": float(np.percentile(err, 99)), "max": float(err.max()), "mag_p50": float(np.percentile(merr, 50)),
            "mag_p99": float(np.percentile(merr, 99)), "oracle": "FP64 direct Ewald (alpha=9, |h|^2<=220), device"}


def precision_sweep(ct, box, x, y, z, h, n, args):
    """The same box kicked in the three arithmetic modes at three opening
    angles; one JSON line with the throughput and force errors of each."""
    ref = direct_reference(box, x, y, z, h, n, args.accuracy_sinks, args.seed + 1)
    rows = []
    modes = [("fp32_compensated", ct.PREC_F32C), ("fp64", ct.PREC_F64), ("fp32_uncompensated", ct.PREC_F32U)]
    for theta in (0.4, 0.55, 0.7):
        box.tree_build(theta, args.bucket, h)
        p = ct.default_kick_params(theta, h, 1.0 / n)
        for name, mode in modes:
            box.kick(p, precision=mode)
            t1 = time.perf_counter()
            box.kick(p, precision=mode)
            dt = time.perf_counter() - t1
            acc = accuracy(box, x, y, z, h, n, args.accuracy_sinks, args.seed + 1, ref=ref)
            rows.append({"theta": theta, "mode": name, "value": n / dt, "ms_per_step": 1e3 * dt, "rms": acc["rms"],
                         "p50": acc["p50"], "p99": acc["p99"], "max": acc["max"]})
            print("sweep: theta %.2f %-18s %8.2f ms  rms %.3e  p99 %.3e" % (theta, name, 1e3 * dt, acc["rms"], acc["p99"]),
                  file=sys.stderr, flush=True)
    print(json.dumps({"metric": "precision sweep: rms / p99 relative force error vs FP64 direct Ewald", "unit": "particle-kicks/s",
                      "n_gpus": 1, "data": "synthetic", "config": {"workload": "clustered box %d particles" % n, "n": n,
                                                                    "bucket": args.bucket, "sinks": int(len(ref[0]))},
                      "sweep": rows}), flush=True)


def ensemble(ct, args):
    """Independent boxes (N log-uniform in [32^3, 128^3], clustered fraction
    U[0, 0.8], NFW halos as in the single box, max(8, N/4096) halos per box)
    kicked together, against the same boxes kicked one per launch."""
    from cosmictiger_amd import autotune, workloads
    ct.init(0)
    theta = args.theta
    specs = autotune.ensemble_specs(args.ensemble_boxes)
    t0 = time.perf_counter()
    boxes, hs, ps = [], [], []
    for n, frac, seed in specs:
        x, y, z = workloads.nfw_box(n, seed, clustered_fraction=frac, nhalos=max(8, n // 4096))
        b = ct.Box(x, y, z)
        h = ct.softening(n)
        boxes.append(b)
        hs.append(h)
        ps.append(ct.default_kick_params(theta, h, 1.0 / n))
    ntot = sum(n for n, _, _ in specs)
    print("ensemble: %d boxes, %d particles, setup %.1fs" % (len(boxes), ntot, time.perf_counter() - t0), file=sys.stderr, flush=True)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    best, tune = autotune.tune_bucket(ct, boxes, [theta] * len(boxes), hs, ps, log=log)
    steps = max(2, args.steps)
    # one box per launch sequence (each box's kick is its own set of launches)
    for b, p in zip(boxes, ps):
        b.kick(p)
    t1 = time.perf_counter()
    for _ in range(steps):
        for b, p in zip(boxes, ps):
            b.kick(p)
    t_single = (time.perf_counter() - t1) / steps
    batch = ct.Batch(boxes)
    batch.kick(ps)
    t1 = time.perf_counter()
    for _ in range(steps):
        batch.kick(ps)
    t_eager = (time.perf_counter() - t1) / steps
    batch.kick(ps, graph=True)
    batch.kick(ps, graph=True)
    lat = []
    for _ in range(steps):
        t2 = time.perf_counter()
        batch.kick(ps, graph=True)
        lat.append(time.perf_counter() - t2)
    t_graph = float(np.mean(lat))
    st = batch.stats()
    batch.close()
    out = {"metric": "ensemble particle-kicks/s (64 boxes in one batched launch sequence vs one box per launch)",
           "value": ntot / t_graph, "unit": "particle-kicks/s", "n_gpus": 1, "steps": steps, "higher_is_better": True,
           "data": "synthetic",
           "config": {"workload": "%d boxes, N log-uniform in [32^3, 128^3] (total %d), clustered fraction U[0, 0.8], theta %.2f, "
                                  "autotuned bucket %d" % (len(boxes), ntot, theta, best), "boxes": len(boxes), "n_total": ntot,
                      "theta": theta, "bucket": best},
           "one_box_per_launch": {"value": ntot / t_single, "ms_per_step": 1e3 * t_single},
           "batched_eager": {"value": ntot / t_eager, "ms_per_step": 1e3 * t_eager},
           "batched_graph": {"value": ntot / t_graph, "ms_per_step": 1e3 * t_graph,
                             "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99))}},
           "speedup_vs_one_box_per_launch": t_single / t_graph,
           "autotune_ms_per_batched_kick": {str(k): 1e3 * v for k, v in tune.items()},
           "events": {k: st[k] for k in ("pp_pairs", "cc", "ewald_cc", "cp_parts", "pc_pairs", "nodes", "sinks", "checks")}}
    for b in boxes:
        b.close()
    print(json.dumps(out), flush=True)


def serving(ct, args):
    """Poisson-arriving requests of 1-10 kick+drift steps on boxes with N
    log-uniform in [16^3, 192^3]: continuous batching against one box per
    launch, at 0.5x, 1x and 2x the saturation arrival rate of the batched
    server (measured first with all requests arriving at once)."""
    from cosmictiger_amd import serve
    ct.init(0)
    ex = serve.GpuExecutor(ct, theta=args.theta, bucket=args.bucket, dt=1e-3)
    budget = 64 * 2 ** 30
    nreq = args.serve_requests

    # the requests' particles are the clients' data: made once, before the runs
    base = serve.poisson_requests(nreq, 1.0, seed=77)
    for r in base:
        serve.make_particles(r)

    def run(rate, batched):
        reqs = serve.poisson_requests(nreq, rate, seed=77)
        for r, b in zip(reqs, base):
            r.particles = b.particles
        srv = serve.Server(ex, hbm_budget=budget, host_threads=min(16, os.cpu_count() or 1), batched=batched)
        for k in ex.timing:
            ex.timing[k] = 0
        res = srv.run(reqs)
        res["max_batch"] = srv.max_batch
        res["executor_time"] = dict(ex.timing)
        srv.close()
        return res

    sat = run(1e9, True)
    rate_sat = nreq / sat["wall_s"]
    rows = []
    for f in (0.5, 1.0, 2.0):
        rate = f * rate_sat
        b = run(rate, True)
        s1 = run(rate, False)
        rows.append({"rate_per_s": rate, "rate_x_saturation": f, "batched": b, "one_box_per_launch": s1})
        print("serve: rate %.2f/s batched %.3e p-steps/s p50 %.0f ms p99 %.0f ms | one box per launch %.3e p50 %.0f p99 %.0f"
              % (rate, b["throughput"], b["latency_ms"]["p50"], b["latency_ms"]["p99"], s1["throughput"], s1["latency_ms"]["p50"],
                 s1["latency_ms"]["p99"]), file=sys.stderr, flush=True)
    ex.close()
    print(json.dumps({"metric": "serving: particle-steps/s and p50/p99 step latency (kick + drift + tree build per step)",
                      "unit": "particle-steps/s", "n_gpus": 1, "data": "synthetic",
                      "config": {"workload": "%d Poisson requests of 1-10 steps, N log-uniform in [16^3, 192^3], clustered U[0,0.8], "
                                             "theta %.2f, bucket %d, HBM budget %d GiB" % (nreq, args.theta, args.bucket, budget >> 30)},
                      "saturation": {"rate_per_s": rate_sat, "run": sat}, "rates": rows}), flush=True)


def config1(ct, args):
    """The plain method: 64^3 uniform particles (seed 42), one kick at theta
    0.7, bucket 128, Ewald on; the GPU kick and the CPU SIMD kick (the
    reference's CPU path) of the whole box, both checked against the FP64
    direct Ewald sum on 4096 seeded sinks."""
    from cosmictiger_amd import workloads
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ct_oracle as oc
    n = 64 ** 3
    theta, bucket = 0.7, 128
    x, y, z = workloads.uniform_box(n, 42)
    h = ct.softening(n)
    ct.init(0)
    box = ct.Box(x, y, z)
    box.tree_build(theta, bucket, h)
    p = ct.default_kick_params(theta, h, 1.0 / n)
    box.kick(p)
    steps = 10
    t0 = time.perf_counter()
    for _ in range(steps):
        box.kick(p)
    gpu_s = (time.perf_counter() - t0) / steps
    ref = direct_reference(box, x, y, z, h, n, 4096, 7)
    acc_gpu = accuracy(box, x, y, z, h, n, 4096, 7, ref=ref)
    nthreads = args.cpu_threads or min(os.cpu_count() or 1, 16)
    tr = oc.OracleTree(x, y, z, bucket, theta, h, nthreads=nthreads)
    po = oc.default_params(theta, h, 1.0 / n)
    t0 = time.perf_counter()
    f, _ = oc.cpu_simd_kick(tr, po, np.zeros((n, 3), np.float32), np.zeros(n, np.int8), nthreads=nthreads)
    cpu_s = time.perf_counter() - t0
    inv = np.empty(n, np.int64)
    inv[tr.id] = np.arange(n)
    sinks, rf = ref
    g = np.stack([f["gx"][inv][sinks], f["gy"][inv][sinks], f["gz"][inv][sinks]], 1).astype(np.float64)
    gr = (1.0 / n) * np.stack([rf["gx"], rf["gy"], rf["gz"]], 1)
    e = np.linalg.norm(g - gr, axis=1) / np.linalg.norm(gr, axis=1)
    box.close()
    print(json.dumps({"metric": "config 1 (plain method): particle-kicks/s and force error vs FP64 direct Ewald",
                      "unit": "particle-kicks/s", "n_gpus": 1, "data": "synthetic",
                      "config": {"workload": "uniform box 262144 particles, seed 42, theta 0.7, bucket 128, order 7, Ewald"},
                      "gpu": {"value": n / gpu_s, "ms_per_kick": 1e3 * gpu_s, "accuracy": acc_gpu},
                      "cpu_simd": {"value": n / cpu_s, "s_per_kick": cpu_s, "cores": nthreads,
                                   "accuracy": {"sinks": int(len(sinks)), "rms": float(np.sqrt(np.mean(e ** 2))),
                                                "p99": float(np.percentile(e, 99))}}}), flush=True)


def drift_check(ct, box, args):
    """The fixed32 drift of every particle (each rung present, dt = rung_dt
    [rung] t0 with t0 = 1/64, a = 1; cuda_drift, src/drift.cu:36-154): device
    time and HBM rate at 37 B per particle, and the positions compared bit for
    bit with the oracle's drift of the same input."""
    d0 = box.read()
    rungs = sorted(set(np.unique(d0["rung"]).tolist()))
    t0 = 1.0 / 64
    ms = 0.0
    for r in rungs:
        td = time.perf_counter()
        box.drift(r, 1.0, t0 * 2.0 ** -r)
        ms += 1e3 * (time.perf_counter() - td)
    d1 = box.read()
    n = len(d0["x"])
    out = {"rungs": rungs, "ms": ms, "bytes_per_particle": 37, "GB_per_s": 37.0 * n / (ms * 1e-3) / 1e9,
           "hbm_frac": 37.0 * n / (ms * 1e-3) / 8e12}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import ct_oracle as oc
        x, y, z = d0["x"], d0["y"], d0["z"]
        for r in rungs:
            x, y, z = oc.drift(x, y, z, d0["vel"], d0["rung"], r, 1.0, t0 * 2.0 ** -r)
        out["bit_exact_vs_oracle"] = bool(np.array_equal(x, d1["x"]) and np.array_equal(y, d1["y"]) and np.array_equal(z, d1["z"]))
    out["lightcone"] = lightcone_check(ct, box, args, rungs[0], t0 * 2.0 ** -rungs[0])
    return out


def lightcone_check(ct, box, args, rung, dt):
    """The drift with the lightcone (src/drift.cu:36-154, do_lc) of the
    particles of one rung, with the light shell at radius ~0.8 from the
    observer: the first call finds its entries (and sizes the entry buffer),
    the second, with the next conformal-time window, is timed.  Algorithmic
    bytes: 25 B per particle read by the crossing pass, 37 B per particle by
    the drift, 28 B per entry written (the entries stay on the device; the
    timed call returns their number only).  With the CPU baseline on, the first
    call's entries are compared with the oracle's (positions and velocities
    bit for bit; the fraction of equal HEALPix pixels)."""
    nside = 1024
    tau_max = 1.5
    tau1 = tau_max - 0.8
    d0 = box.read() if not args.no_cpu_baseline else None
    e1 = box.drift_lightcone(rung, 1.0, dt, tau1 - dt, tau1, tau_max, nside)
    td = time.perf_counter()
    n2 = box.drift_lightcone(rung, 1.0, dt, tau1, tau1 + dt, tau_max, nside, read=False)
    ms = 1e3 * (time.perf_counter() - td)
    n = box.n
    nbytes = 62.0 * n + 28.0 * n2
    out = {"rung": rung, "nside": nside, "entries": [int(len(e1)), int(n2)], "ms": ms, "GB_per_s": nbytes / (ms * 1e-3) / 1e9,
           "hbm_frac": nbytes / (ms * 1e-3) / 8e12}
    if d0 is not None:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import ct_oracle as oc
        w = oc.lightcone(d0["x"], d0["y"], d0["z"], d0["vel"], d0["rung"], rung, 1.0, dt, tau1 - dt, tau1, tau_max, nside)
        key = lambda e: np.lexsort((e["z"], e["y"], e["x"]))
        g, w = e1[key(e1)], w[key(w)]
        same = len(g) == len(w) and all(np.array_equal(g[f], w[f]) for f in ("x", "y", "z", "vx", "vy", "vz"))
        out["entries_bit_exact_vs_oracle"] = bool(same)
        out["pixel_agreement"] = float(np.mean(g["pix"] == w["pix"])) if same and len(g) else None
    return out


def copy_calibration(mib):
    """One hipMemcpy device to device of mib MiB (the HIP runtime runs it as a
    copy kernel, __amd_rocclr_copyBuffer): the known-size dispatch that scales
    the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    nbytes = ctypes.c_size_t(mib << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), nbytes) or hip.hipMalloc(ctypes.byref(b), nbytes):
        raise RuntimeError("copy calibration: hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, nbytes) or hip.hipMemcpy(b, a, nbytes, 3) or hip.hipDeviceSynchronize():  # 3: device to device
            raise RuntimeError("copy calibration: copy failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)


def cpu_baseline(x, y, z, args, h, gpu_forces=None):
    """The reference's CPU SIMD kick path (kick() + cpu_gravity_*, restated in
    AVX2 FP32 under oracle/) on a bounded sample of sinks of the same box, on
    this host's cores.  The sample grows until it is ~10 s of CPU work.  With
    gpu_forces (forces by original particle index), the CPU kick's forces on
    the sample are compared with the GPU kick's."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ct_oracle as oc
    nthreads = args.cpu_threads or min(os.cpu_count() or 1, 16)
    tb = time.perf_counter()
    tr = oc.OracleTree(x, y, z, args.bucket, args.theta, h, nthreads=nthreads)
    tree_s = time.perf_counter() - tb
    n = len(x)
    vel = np.zeros((n, 3), np.float32)
    rung = np.zeros(n, np.int8)
    p = oc.default_params(args.theta, h, 1.0 / n)
    sample = int(args.cpu_sample)
    start = n // 2 - sample // 2
    while True:
        v = vel.copy()
        r = rung.copy()
        t0 = time.perf_counter()
        f, _ = oc.cpu_simd_kick(tr, p, v, r, nthreads=nthreads, sample=(start, start + sample))
        dt = time.perf_counter() - t0
        if dt * nthreads >= 10.0 or sample >= args.cpu_sample_max or sample >= n:
            break
        sample = min(n, sample * 4, args.cpu_sample_max)
        start = max(0, n // 2 - sample // 2)
    out = {"value": sample / dt, "unit": "particle-kicks/s", "cores": nthreads, "kind": "port",
           "sample": "CPU SIMD kick (the reference's CPU kick path, AVX2 FP32, GPU-path interaction lists) of %d sinks "
                     "[%d, %d) of the same %d-particle box and tree, %.2f s wall on %d threads (tree build %.1f s, "
                     "not timed)" % (sample, start, start + sample, n, dt, nthreads, tree_s)}
    if gpu_forces is not None:
        ids = tr.id[start:start + sample]
        gc = np.stack([f["gx"][start:start + sample], f["gy"][start:start + sample], f["gz"][start:start + sample]], 1).astype(np.float64)
        gg = np.stack([gpu_forces[k][ids] for k in ("gx", "gy", "gz")], 1).astype(np.float64)
        e = np.linalg.norm(gc - gg, axis=1) / np.maximum(np.linalg.norm(gg, axis=1), 1e-300)
        out["force_rel_diff_vs_gpu"] = {"rms": float(np.sqrt(np.mean(e ** 2))), "p99": float(np.percentile(e, 99))}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256 ** 3)
    ap.add_argument("--workload", default="clustered", choices=["clustered", "uniform"])
    ap.add_argument("--theta", type=float, default=0.55)
    ap.add_argument("--bucket", default="auto",
                    help="leaf bucket size, or 'auto': time a kick at each of --bucket-candidates on the resident box "
                         "and keep the fastest")
    ap.add_argument("--bucket-candidates", default="64,96,128,160,192")
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=4096)
    ap.add_argument("--cpu-sample-max", type=int, default=4194304)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--accuracy-sinks", type=int, default=10000)
    ap.add_argument("--theta-accurate", type=float, default=0.4,
                    help="second measurement at an opening angle that meets rms force error <= 1e-3 (0 to skip)")
    ap.add_argument("--precision-sweep", action="store_true",
                    help="instead of the throughput line: FP32 compensated / FP64 / FP32 uncompensated kicks at theta "
                         "0.4, 0.55, 0.7 with their rms and p99 force errors")
    ap.add_argument("--config1", action="store_true",
                    help="the plain-method case: 64^3 uniform, theta 0.7, GPU and CPU SIMD kicks with their force errors")
    ap.add_argument("--ensemble", action="store_true",
                    help="instead of the throughput line: the 64-box ensemble kicked in one batched, graph-captured "
                         "launch sequence against one box per launch, with the autotuned bucket")
    ap.add_argument("--ensemble-boxes", type=int, default=64)
    ap.add_argument("--serve", action="store_true",
                    help="instead of the throughput line: the serving loop (Poisson kick+drift requests, continuous "
                         "batching) at three arrival rates, against one box per launch")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "pp_kernel_traffic.json"))
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="before the kicks, copy this many MiB device to device (a blit kernel that reads and "
                         "writes that many bytes): the calibration dispatch of the HBM traffic counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; runs on %d GPUs are not part of it (independent boxes shard one "
              "per GPU)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import cosmictiger_amd as ct
    if args.bucket != "auto":
        args.bucket = int(args.bucket)
    elif args.ensemble or args.serve:
        args.bucket = 128  # the ensemble mode tunes its own bucket; serving uses the reference default
    if args.config1:
        config1(ct, args)
        return
    if args.ensemble:
        ensemble(ct, args)
        return
    if args.serve:
        serving(ct, args)
        return
    t0 = time.perf_counter()
    x, y, z = make_box(args)
    n = len(x)
    h = ct.softening(n)
    ct.init(0)
    box = ct.Box(x, y, z)
    params = ct.default_kick_params(args.theta, h, 1.0 / n)
    tuning = None
    if args.bucket == "auto":
        tuning = {}
        for b in [int(v) for v in args.bucket_candidates.split(",")]:
            box.tree_build(args.theta, b, h)
            box.kick(params)
            tk = time.perf_counter()
            for _ in range(2):
                box.kick(params)
            tuning[b] = 1e3 * (time.perf_counter() - tk) / 2
            print("bench: bucket %d: %.2f ms per kick" % (b, tuning[b]), file=sys.stderr, flush=True)
        args.bucket = min(tuning, key=tuning.get)
    tb0 = time.perf_counter()
    box.tree_build(args.theta, args.bucket, h)
    tree_s = time.perf_counter() - tb0
    print("bench: %d particles, tree %.2fs, setup %.2fs" % (n, tree_s, time.perf_counter() - t0), file=sys.stderr, flush=True)
    if args.copy_calibration_mib > 0:
        copy_calibration(args.copy_calibration_mib)
    if args.precision_sweep:
        precision_sweep(ct, box, x, y, z, h, n, args)
        box.close()
        return
    for _ in range(args.warmup):
        box.kick(params)
    pp_ms = []
    step_ms = []
    stats = None
    ts = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        box.kick(params)
        step_ms.append(1e3 * (time.perf_counter() - t1))
        stats = box.stats()
        pp_ms.append(stats["pp_ms"])
    elapsed = time.perf_counter() - ts
    ms_per_step = 1e3 * elapsed / args.steps
    value = n * args.steps / elapsed
    flops = W_PP * stats["pp_pairs"]
    pp_s = np.mean(pp_ms) * 1e-3
    achieved = flops / pp_s / 1e12
    traffic = None
    if os.path.exists(args.traffic_file):
        try:
            with open(args.traffic_file) as f:
                tr = json.load(f)
            for e in tr.get("entries", [tr]):
                if e.get("n") == n and e.get("workload") == args.workload and e.get("bucket") == args.bucket:
                    traffic = e.get("bytes_per_launch")
        except Exception:
            traffic = None
    gpu_forces = None
    if not args.no_cpu_baseline:
        # the timed kicks' forces, by original index, for the CPU baseline's check
        d = box.read()
        o = np.argsort(d["id"])
        gpu_forces = {k: d[k][o] for k in ("gx", "gy", "gz")}
        del d
    acc = accuracy(box, x, y, z, h, n, args.accuracy_sinks, args.seed + 1) if args.accuracy_sinks > 0 else None
    at_1e3 = None
    if args.theta_accurate > 0:
        # the same box at the opening angle where the method meets rms <= 1e-3
        box.tree_build(args.theta_accurate, args.bucket, h)
        pa = ct.default_kick_params(args.theta_accurate, h, 1.0 / n)
        box.kick(pa)
        ta = time.perf_counter()
        na = max(1, args.steps // 2)
        for _ in range(na):
            box.kick(pa)
        ea = time.perf_counter() - ta
        at_1e3 = {"theta": args.theta_accurate, "value": n * na / ea, "ms_per_step": 1e3 * ea / na, "steps": na,
                  "accuracy": accuracy(box, x, y, z, h, n, args.accuracy_sinks, args.seed + 1) if args.accuracy_sinks > 0 else None}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(x, y, z, args, h, gpu_forces)
    drift = drift_check(ct, box, args)
    out = {
        "metric": "particle-kicks/s at rms force err <=1e-3 vs FP64 Ewald; p50/p99 step latency",
        "value": value,
        "unit": "particle-kicks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "%s box %d particles (70%% in 2000 NFW halos, 30%% uniform), theta %.2f, bucket %d, "
                               "order 7, Ewald, seed %d" % (args.workload, n, args.theta, args.bucket, args.seed)
                   if args.workload == "clustered" else "uniform box %d particles, theta %.2f, bucket %d, order 7, Ewald, seed %d"
                   % (n, args.theta, args.bucket, args.seed),
                   "n": n, "theta": args.theta, "bucket": args.bucket, "parallelism": "single GPU",
                   "bucket_autotune_ms": tuning},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "kernel": "k_pp (P2P, all leaf-size classes of one kick), FP32 vector ALU (compute bound, no MFMA); "
                               "peak = MI355X published FP32 figure",
                     "flops_per_launch": flops, "launch_ms": float(np.mean(pp_ms))},
        "cpu_baseline": cpu,
        "accuracy": acc,
        "at_rms_1e-3": at_1e3,
        "latency_ms": {"p50": float(np.percentile(step_ms, 50)), "p99": float(np.percentile(step_ms, 99))},
        "phases_ms": {k: stats[k] for k in ("walk_ms", "cells_ms", "l2l_ms", "leaf_ms", "pp_ms", "leafkick_ms", "kick_ms")},
        "events": {k: stats[k] for k in ("pp_pairs", "cc", "ewald_cc", "cp_parts", "pc_pairs", "nodes", "sinks", "checks")},
        "tree_build_s": tree_s,
        "drift": drift,
    }
    box.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
