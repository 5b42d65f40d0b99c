"""Benchmark: ADMM outer iterations of the 4D reconstruction on one MI355X.

One step = one outer ADMM iteration (titer inner steps + prox/dual update + rho
rule) over the whole workload, with every input already resident in HBM.
Default workload: N=1024, Nz=64, Nzp=16, Nrot=8 x 200 = 1600 projections, M=16,
titer=4 (synthetic moving-sphere data).  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--config full|mid|small] [--titer 4] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "voxel-iter/s at N=1024 M=16 (% HBM roofline); batched/served vs plain; p50/p99"
HBM_PEAK = 8.0e12


def kernel_bytes(name, n, nt, m, zc, nzp):
    """Algorithmic HBM bytes of one launch of each kernel over zc slices (each input
    read once, each output written once).  The M real frames of a slice are
    transformed as ceil(M/2) complex pairs, so there are S = zc*ceil(M/2) spectra."""
    N = 2 * n
    V = zc * m * n * n                       # voxels
    V4 = (zc // nzp) * (nzp + 1) * (m + 1) * (n + 1) ** 2  # dual entries
    S = zc * ((m + 1) // 2)                  # pair spectra
    Gs = zc * nt * n                         # polar samples
    return {
        "fwd_rows": 4 * V + S * n * N * 8,
        "fwd_cols": S * (n * N + N * N) * 8,
        "gather": S * N * N * 8 + Gs * 16,
        "ray": Gs * (16 + 4 + 8),
        "scatter": Gs * 8 + S * N * N * 8,
        "adj_cols": S * (N * N + n * N) * 8,
        "adj_rows": S * (n * N + n * n) * 8,
        "update": V * (4 + 4 + 4 + 4),         # f in, div(chi) in, adjoint frames in, f out
        "divchi": 16 * V4 + 4 * V,             # chi = mu - rho psi in (and mu when the rule changed rho); div chi out
        "prox": 8 * V + 48 * V4,
    }.get(name)


def kernel_flops(name, n, nt, m, zc):
    """Algorithmic FLOPs of one launch of the window kernels over zc slices: every
    polar sample has a direct and a mirrored item, each a 9 x 9-tap window sum over
    the 2*ceil(M/2) real columns of the pair spectra (2 FLOPs per multiply-add).
    The FFT and TV kernels are counted by bytes only."""
    if name not in ("gather", "scatter"):
        return None
    return 2 * nt * n * zc * 81 * 2 * ((m + 1) // 2) * 2


FP32_MFMA_PEAK = 157.3e12  # dense FP32 matrix FLOP/s of MI355X


def byte_model(n, nt, m, nz, nzp, titer, pairs=False):
    """Compulsory HBM bytes of one outer iteration (fused-pipeline model).  With
    pairs=False the model counts the reference algorithm's M*Nz spectra; with
    pairs=True the ceil(M/2)*Nz frame-pair spectra this solver transforms (the
    same arithmetic: two real frames per complex transform)."""
    V = n * n * m * nz
    V4 = (n + 1) ** 2 * (m + 1) * (nzp + 1) * nz // nzp
    S = ((m + 1) // 2 if pairs else m) * nz
    G = (2 * n) ** 2 * 8
    H = 2 * n * n * 8
    Gs = n * nt * nz
    inner = (4 * V + S * (2 * H + 2 * G) + 8 * Gs) + (4 * Gs + S * (2 * G + 2 * H) + 8 * V) + 32 * V4
    return titer * inner + 12 * V + 48 * V4


def _hip():
    import ctypes
    return ctypes, ctypes.CDLL("libamdhip64.so")


def copy_probe(nbytes=1 << 30, reps=5):
    """A device-to-device copy of known size on the null stream (for calibrating
    rocprofv3's FETCH_SIZE / WRITE_SIZE counters against a plain copy kernel, and
    as the measured stream-copy bandwidth).  Returns the best achieved bandwidth
    (bytes read + written per second) over `reps` timed copies."""
    ctypes, hip = _hip()
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipEventCreate(ctypes.byref(e0))
    hip.hipEventCreate(ctypes.byref(e1))
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    best = 0.0
    for _ in range(reps):
        hip.hipEventRecord(e0, None)
        hip.hipMemcpyAsync(b, a, ctypes.c_size_t(nbytes), 3, None)  # hipMemcpyDeviceToDevice
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        if ms.value > 0:
            best = max(best, 2 * nbytes / (ms.value * 1e-3))
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    return best


def measured_traffic(config, kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (profiles/traffic_<config>.json, FETCH_SIZE + WRITE_SIZE calibrated on a copy
    of known size), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % config)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else k.get("bytes_per_launch")


CPU_THREADS = 16  # the host share of one GPU on the box


def cpu_baseline(w, titer, threads=None):
    """FP32 NumPy/SciPy port of the same solver (the oracle in float32) on a bounded
    sample of the workload, on `threads` host threads (default min(16, cores)):
    `threads` slices of one chunk, each thread running the reference's per-chunk
    ADMM step (admm.cu:74-122) on its own one-slice chunk (halo slabs from the
    previous iterate, all projections, all basis functions): once with one inner
    step and once with none (prox / dual update only).  The chunk loop is the
    reference's own unit of parallel work (one GPU per chunk, rectv.cu:161-162), so
    the time of an outer iteration scales as titer*inner + prox per slice.
    value = n^2*M*slices / (titer*t_inner + t_prox)."""
    import threading
    from oracle import rectv_oracle as O
    O.set_workers(1)
    threads = threads or min(CPU_THREADS, os.cpu_count() or 1)
    n, nt, m = w["n"], w["ntheta"], w["m"]
    dt = np.float32
    ops = O.Operators(n, nt, m, w["theta"], w["phi"], w["center"], w["lambda1"], dt)
    lam0 = w["lambda0"] * w["lambda1"]
    rng = np.random.default_rng(0)
    f0 = rng.random((threads + 2, m, n, n)).astype(dt) * 0.1

    def one(z, t_iter):
        fext = f0[z:z + 3].copy()
        g = w["data"][z % w["nz"]][None].astype(dt)
        psi = np.zeros((2, m + 1, n + 1, n + 1, 4), dt)
        mu = np.zeros_like(psi)
        O.solver_admm(ops, fext, g, psi, mu, lam0, 0.5, 1, 1, t_iter)

    def timed(t_iter):
        ths = [threading.Thread(target=one, args=(z, t_iter)) for z in range(threads)]
        t0 = time.perf_counter()
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        return time.perf_counter() - t0

    t_step = timed(1)
    t_prox = timed(0)
    t_inner = max(t_step - t_prox, 1e-9)
    t_iter = titer * t_inner + t_prox
    return {"value": n * n * m * threads / t_iter, "unit": "voxel-iter/s", "cores": threads, "kind": "port",
            "sample": "%d slices of one chunk (n=%d, %d projections, M=%d), one thread per slice running the "
                      "per-chunk ADMM step on a one-slice chunk: one inner step %.2f s + prox/dual update %.2f s "
                      "(wall, all threads), FP32 NumPy/SciPy on %d host threads (os.cpu_count() = %s); "
                      "value = n^2*M*slices / (titer*inner + prox) with titer=%d, i.e. extrapolated from one "
                      "slice per thread to the whole volume" % (threads, n, nt, m, t_inner, t_prox, threads,
                                                                os.cpu_count(), titer)}


MIX_FIXTURE = os.path.join(ROOT, "tests", "golden", "mix_samples.npz")


def oracle_sample_errors(jobs, mix):
    """Relative L2 error of the sampled jobs' results against the FP64 oracle results
    committed in tests/golden/mix_samples.npz (a strided subsample of each result,
    made by tests/golden/make_mix_samples.py for the 64-job mixes at niter=64,
    titer=4), or None when the fixture does not cover this run.  Jobs whose stop
    iteration differs from the oracle's are reported, not compared."""
    if not os.path.exists(MIX_FIXTURE) or len(jobs) != 64:
        return None
    z = np.load(MIX_FIXTURE, allow_pickle=False)
    ids = sorted({int(k.split("_")[1]) for k in z.files if k.startswith(mix + "_")})
    out = {}
    for i in ids:
        j, pre = jobs[i], "%s_%d_" % (mix, i)
        if j.niter != 64 or j.titer != 4 or j.result is None:
            return None
        if j.iterations != int(z[pre + "iterations"]):
            out[j.name] = "stopped after %d, oracle after %d" % (j.iterations, int(z[pre + "iterations"]))
            continue
        zs, xs = max(1, -(-j.nz // 4)), max(1, j.n // 16)
        sub = j.result[::zs, :, ::xs, ::xs].astype(np.float64)
        ref = z[pre + "f_sub"].astype(np.float64)
        out[j.name] = float(np.linalg.norm(sub - ref) / np.linalg.norm(ref))
    return out


def batched_bench(count, niter, titer, workers=8, part="both"):
    """Batched-instances mix: `count` uneven jobs solved one after another with the
    plain API (Solver.recon each) and then together (batch.solve_batch); same
    inputs, same arithmetic.  Value = sum of n^2*Nz*M*niter over jobs / wall time,
    host set-up (geometry tables, uploads) included in both.  part = "sequential" or
    "batched" runs only that half (for a kernel trace of one of them)."""
    from rectv_gpu_amd import synthetic, batch
    specs = synthetic.batch_specs(count)
    jobs0 = [synthetic.make_job(sp, niter, titer) for sp in specs]
    jobs_a = [j.fresh() for j in jobs0]
    vox = sum(j.voxels * j.niter for j in jobs_a)
    print("bench: batched mix of %d jobs generated (%.2f G voxel-iterations)" % (count, vox / 1e9),
          file=sys.stderr, flush=True)
    if part != "both":
        t0 = time.perf_counter()
        if part == "sequential":
            batch.solve_sequential(jobs_a)
        else:
            batch.solve_batch(jobs_a, workers=workers)
        t = time.perf_counter() - t0
        print("bench: %s %.2f s" % (part, t), file=sys.stderr, flush=True)
        return {"workload": "%d jobs, niter=%d, titer=%d, %s only" % (count, niter, titer, part),
                "value": round(vox / t, 1), "unit": "voxel-iter/s", "wall_s": round(t, 3), "voxel_iterations": vox}
    t0 = time.perf_counter()
    batch.solve_sequential(jobs_a)
    t_seq = time.perf_counter() - t0
    print("bench: sequential plain solves %.2f s" % t_seq, file=sys.stderr, flush=True)
    ref = [j.result for j in jobs_a]
    del jobs_a
    jobs_b = [j.fresh() for j in jobs0]
    t0 = time.perf_counter()
    batch.solve_batch(jobs_b, workers=workers)
    t_bat = time.perf_counter() - t0
    print("bench: batched %.2f s" % t_bat, file=sys.stderr, flush=True)
    diff = max(float(np.linalg.norm(j.result.astype(np.float64) - r) / max(np.linalg.norm(r), 1e-30))
               for j, r in zip(jobs_b, ref))
    return {"workload": "%d jobs: N in {128,256,384,512}, Nz U{4..32}, Nrot {2,4,8} x {64,100,128} "
                        "projections, M {4,8,16}, Nzp=Nz, niter=%d, titer=%d" % (count, niter, titer),
            "value": round(vox / t_bat, 1), "unit": "voxel-iter/s", "sequential_value": round(vox / t_seq, 1),
            "speedup_vs_sequential": round(t_seq / t_bat, 3), "batched_s": round(t_bat, 3),
            "sequential_s": round(t_seq, 3), "voxel_iterations": vox, "max_rel_l2_vs_plain": diff,
            "sample_rel_l2_vs_fp64_oracle": oracle_sample_errors(jobs_b, "batch")}


def served_bench(count, niter, titer, rate, seed=7, max_active=16, sample_frac=0.1):
    """Serving mix: `count` jobs (the batched mix plus 5% N=1024, Nz=32, M=16 jobs)
    arriving as a Poisson process of `rate` jobs/s, admitted by device footprint,
    iterations interleaved.  Reports throughput over the whole run, p50/p99 job
    latency (arrival to result on the host), and, on a sample of the jobs, the
    relative L2 difference of the served result from a plain solve of the same
    number of iterations and from the FP64 oracle (committed fixture).  Each job is
    checked after every outer iteration (ahead=1), so it stops at the first
    iteration below the tolerance, as the oracle's run does."""
    import threading
    from rectv_gpu_amd import synthetic, batch
    specs = synthetic.serve_specs(count)
    # a job ends at niter or when its relative change per outer iteration drops below 1e-4
    jobs = [synthetic.make_job(sp, niter, titer, tol=1e-4) for sp in specs]
    gaps = np.random.default_rng(seed).exponential(1.0 / rate, count)
    print("bench: serving mix of %d jobs generated" % count, file=sys.stderr, flush=True)
    srv = batch.Server(max_active=max_active, ahead=1, workers=4)

    def producer():
        t0 = time.perf_counter()
        at = np.cumsum(gaps)
        for j, a in zip(jobs, at):
            d = t0 + a - time.perf_counter()
            if d > 0:
                time.sleep(d)
            srv.submit(j)

    th = threading.Thread(target=producer)
    t0 = time.perf_counter()
    th.start()
    try:
        srv.run(until=lambda: not th.is_alive() and srv.pending() == 0)
    finally:
        th.join()
        srv.close()
    t1 = time.perf_counter()
    errors = [j.error for j in jobs if j.error]
    vox = sum(j.voxels * j.iterations for j in jobs if not j.error)
    lat = batch.latency_stats(jobs)
    print("bench: served %d jobs in %.2f s, p50 %.2f s, p99 %.2f s" % (count, t1 - t0, lat["p50_s"], lat["p99_s"]),
          file=sys.stderr, flush=True)
    k = max(1, int(round(sample_frac * count)))
    sample = sorted(range(count), key=lambda i: jobs[i].voxels)[:k]
    diff = 0.0
    for i in sample:
        plain = jobs[i].fresh()
        plain.niter = jobs[i].iterations  # the served job may have stopped early
        r = batch.solve_sequential([plain])[0].result
        diff = max(diff, float(np.linalg.norm(jobs[i].result.astype(np.float64) - r) / max(np.linalg.norm(r), 1e-30)))
    return {"workload": "%d jobs (batched mix + 5%% N=1024 Nz=32 M=16), Poisson arrivals %g jobs/s, niter=%d "
                        "or relative change < 1e-4, titer=%d, max %d active" % (count, rate, niter, titer, max_active),
            "value": round(vox / (t1 - t0), 1), "unit": "voxel-iter/s", "wall_s": round(t1 - t0, 3),
            "latency": lat, "jobs": count, "errors": len(errors), "peak_active": srv.peak_active,
            "sample_max_rel_l2_vs_plain": diff, "sample_jobs": k,
            "iterations": {"min": min(j.iterations for j in jobs), "max": max(j.iterations for j in jobs),
                           "stopped_early": sum(1 for j in jobs if j.iterations < j.niter)},
            "sample_rel_l2_vs_fp64_oracle": oracle_sample_errors(jobs, "serve")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--titer", type=int, default=4)
    ap.add_argument("--config", default="full")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-titer1", action="store_true",
                    help="skip the secondary titer=1 timing (profiling runs: only the headline iterations follow the autotuner)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="accepted for older profiling commands: the 1 GiB device copy (counter calibration, copy bandwidth) always runs")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default gather/scatter launch shapes")
    ap.add_argument("--mode", default="all", choices=["all", "plain", "batched", "served"],
                    help="plain: the headline solve only; batched / served: those forms only; all: each")
    ap.add_argument("--jobs", type=int, default=64, help="jobs in the batched / served mixes")
    ap.add_argument("--job-niter", type=int, default=64, help="outer iterations per batched / served job")
    ap.add_argument("--job-titer", type=int, default=4)
    ap.add_argument("--mix-part", default="both", choices=["both", "sequential", "batched"],
                    help="--mode batched: run both halves of the comparison, or one of them")
    ap.add_argument("--rate", type=float, default=20.0, help="serving arrival rate (jobs/s)")
    ap.add_argument("--light-rate", type=float, default=0.5,
                    help="arrival rate of the second, below-capacity serving run (jobs/s; 0 skips it)")
    ap.add_argument("--light-jobs", type=int, default=12)
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: the solver drives exactly one GPU; --gpus %d is not supported "
              "(sharding the z-chunks across GPUs is not implemented)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    os.environ.setdefault("RECTV_QUIET", "1")
    from rectv_gpu_amd import Solver, synthetic
    if a.mode in ("batched", "served"):
        out = {"metric": METRIC, "n_gpus": 1, "dtype": "f32", "data": "synthetic"}
        if a.mode == "batched":
            out["batched"] = batched_bench(a.jobs, a.job_niter, a.job_titer, part=a.mix_part)
        else:
            out["served"] = served_bench(a.jobs, a.job_niter, a.job_titer, a.rate)
            if a.light_rate > 0:
                out["served_light"] = served_bench(a.light_jobs, a.job_niter, a.job_titer, a.light_rate)
        print(json.dumps(out), flush=True)
        return

    t_setup = time.perf_counter()
    w = synthetic.workload(a.config, seed=a.seed)
    n, nt, m, nz, nzp = w["n"], w["ntheta"], w["m"], w["nz"], w["nzp"]
    lam0, lam1 = w["lambda0"] * w["lambda1"], w["lambda1"]
    print("bench: data generated (%.1f s); config %s n=%d nz=%d nzp=%d ntheta=%d m=%d titer=%d"
          % (time.perf_counter() - t_setup, a.config, n, nz, nzp, nt, m, a.titer), file=sys.stderr, flush=True)
    s = Solver(n, nt, m, nz, nzp, 1)
    s.set_inputs(w["data"], w["theta"], w["phi"], w["center"])
    s.reset()
    tuned = None
    if not a.no_autotune:
        tuned = s.autotune()
        s.reset()
        print("bench: autotuned gather/scatter slices per workgroup %s" % (tuned,), file=sys.stderr, flush=True)
    print("bench: solver ready, %.1f GB on device" % (s.device_bytes() / 1e9), file=sys.stderr, flush=True)
    if a.warmup > 0:
        s.iterate(a.warmup, a.titer, lam0, lam1)
    s.synchronize()
    print("bench: warmup done", file=sys.stderr, flush=True)
    # timed region; per-kernel HIP events are recorded on the solver's stream
    s.set_profiling(True)
    s.synchronize()
    t0 = time.perf_counter()
    s.iterate(a.steps, a.titer, lam0, lam1)
    s.synchronize()
    t1 = time.perf_counter()
    prof = s.profile()
    s.set_profiling(False)
    ms = (t1 - t0) * 1e3 / a.steps
    vox = n * n * nz * m
    value = vox * a.steps / (t1 - t0)
    print("bench: %.1f ms per outer iteration" % ms, file=sys.stderr, flush=True)
    zb = nz  # kernels run over the whole slice batch
    kern = {}
    for name, (cnt, tot) in prof.items():
        b = kernel_bytes(name, n, nt, m, zb, nzp)
        avg = tot / cnt
        kern[name] = {"launches": cnt, "avg_ms": round(avg, 4), "share": 0.0}
        if b is not None:
            kern[name]["alg_bytes"] = b
            kern[name]["gbps"] = round(b / (avg * 1e-3) / 1e9, 1)
        fl = kernel_flops(name, n, nt, m, zb)
        if fl is not None:
            kern[name]["alg_flops"] = fl
            kern[name]["tflops"] = round(fl / (avg * 1e-3) / 1e12, 2)
    total = sum(t for _, t in prof.values())
    for name, (cnt, tot) in prof.items():
        kern[name]["share"] = round(tot / total, 4)
    dom = max(prof.items(), key=lambda kv: kv[1][1])[0]
    db = kernel_bytes(dom, n, nt, m, zb, nzp)
    dfl = kernel_flops(dom, n, nt, m, zb)
    davg = prof[dom][1] / prof[dom][0] * 1e-3
    achieved = db / davg / 1e9
    hbm_view = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": round(achieved / (HBM_PEAK / 1e9), 4)}
    # The binding roofline is the larger of bytes / HBM peak and FLOPs / FP32 matrix peak:
    # the window kernels (about 30 FLOP per algorithmic byte, above the 19.7 FLOP/byte
    # ridge point of MI355X) are bound by the matrix cores, the rest by HBM.
    if dfl is not None and dfl / FP32_MFMA_PEAK > db / HBM_PEAK:
        tf = dfl / davg / 1e12
        roof = {"kernel": dom, "bound": "mfma", "achieved": round(tf, 2), "peak": FP32_MFMA_PEAK / 1e12,
                "unit": "TFLOP/s", "frac": round(tf / (FP32_MFMA_PEAK / 1e12), 4), "alg_flops_per_launch": dfl,
                "hbm_view": hbm_view}
    else:
        roof = dict(hbm_view, kernel=dom)
    roof.update({"traffic": measured_traffic(a.config, dom), "alg_bytes_per_launch": db})
    B = byte_model(n, nt, m, nz, nzp, a.titer)
    Bp = byte_model(n, nt, m, nz, nzp, a.titer, pairs=True)
    # the API default titer = 1 (solver.py:47) on the same solver and data, as a secondary figure
    t1_out = None
    if a.titer != 1 and a.steps > 0 and not a.no_titer1:
        s.reset()
        s.iterate(1, 1, lam0, lam1)
        s.synchronize()
        u0 = time.perf_counter()
        s.iterate(a.steps, 1, lam0, lam1)
        s.synchronize()
        u1 = time.perf_counter()
        t1_out = {"titer": 1, "ms_per_step": round((u1 - u0) * 1e3 / a.steps, 2),
                  "value": round(vox * a.steps / (u1 - u0), 1),
                  "iteration_hbm_frac": round(byte_model(n, nt, m, nz, nzp, 1) / ((u1 - u0) / a.steps) / HBM_PEAK, 4)}
    s.free()
    copy_bw = copy_probe()
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "voxel-iter/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain ADMM solve, N=%d Nz=%d Nzp=%d ntheta=%d (Nrot=%d x %d) M=%d titer=%d"
                   % (n, nz, nzp, nt, w["nrot"], w["ntheta_rot"], m, a.titer),
                   "n": n, "nz": nz, "nzp": nzp, "ntheta": nt, "m": m, "titer": a.titer,
                   "launch_shapes": None if tuned is None else {"gather_slices": tuned[0], "scatter_slices": tuned[1]}},
        # byte model of the reference algorithm (M*Nz spectra) and of the frame-pair
        # pipeline this solver runs (ceil(M/2)*Nz spectra), against the 8 TB/s peak
        # and against the measured device-to-device copy bandwidth
        "iteration_hbm_frac": round(B / (ms * 1e-3) / HBM_PEAK, 4),
        "iteration_model_bytes": B,
        "iteration_hbm_frac_pairs": round(Bp / (ms * 1e-3) / HBM_PEAK, 4),
        "iteration_model_bytes_pairs": Bp,
        "copy_bandwidth_gbps": round(copy_bw / 1e9, 1),
        "iteration_copy_frac_pairs": round(Bp / (ms * 1e-3) / copy_bw, 4) if copy_bw > 0 else None,
        "roofline": roof,
        "kernels": kern,
        "titer1": t1_out,
    }
    if a.mode == "all":
        out["batched"] = batched_bench(a.jobs, a.job_niter, a.job_titer)
        out["served"] = served_bench(a.jobs, a.job_niter, a.job_titer, a.rate)
        if a.light_rate > 0:
            out["served_light"] = served_bench(a.light_jobs, a.job_niter, a.job_titer, a.light_rate)
    if not a.no_cpu:
        print("bench: CPU baseline sample ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(w, a.titer)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
