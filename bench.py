"""Benchmark: log-polar Radon fwd+adj throughput on one MI355X.

A step = one batch of ``--slices`` slices through fwd (lpTransform.fwdp) and
then adj (lpTransform.adjp), inputs resident in HBM.  Workload: N0 = 2048,
Nproj = 1500 angles on [0, pi), centre N0/2, cubic interpolation, no filter,
32 slices per batch (the reference's headline fwd+adj configuration), seeded
random-ellipse phantoms.  Prints one JSON line.

Also reported: the plain path (same kernels, one slice per call), the
per-stage durations from HIP events recorded on the launch stream during the
timed region, the roofline of the dominant stage and of the whole pipeline,
and the FP64 numpy/scipy oracle mirror timed on this host as the CPU baseline.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
METRIC = "fwd+adj slices/s at N=2048,1500 proj: batched vs plain GPU; % of HBM roofline"


def stage_bytes(g, fused_polar=False):
    """Algorithmic HBM bytes per slice of each stage (one launch covers all
    slices of a batch and all three spans).  L = 4*Nrho*Ntheta bytes is one
    real log-polar grid; the half spectra S of a span also occupy L bytes.
    With the forward operator's inverse row pass and polar gather fused, the
    fused kernel is timed as rows_inv (S read, R written) and polar_out is
    empty."""
    N0, N, P0, P = g["N0"], g["N"], g["Nproj0"], g["Nproj"]
    L = 4 * g["Nrho"] * g["Ntheta"]
    fwd = {
        "prep_rows": 4 * N0 * N0 + 4 * N * N,      # read f, write padded x-filtered image
        "col_iir": 8 * N * N,                      # read + write the image
        "gather_lp": 4 * N * N + 3 * L,            # image read once, log-polar grid written per span
        "rows_fwd": 3 * 2 * L,                     # grid read, half spectra written per span
        "cols": 3 * 2 * L,                         # S read + written per span
        "rows_inv": 3 * 2 * L,                     # S read, G written per span
        "polar_out": 3 * L // 2 + 4 * N0 * P0,     # half of each G (|theta| < beta/2) read, R written
    }
    if fused_polar:
        fwd["rows_inv"] = 3 * L + 4 * N0 * P0        # S read per span, R written
        fwd["polar_out"] = 0
    adj = {
        # the cubic path writes only the kept angle rows of the stuffed
        # sinogram (P0 of P = os P0); the column sweep reads them and writes all P
        "sino_rows": 4 * N0 * P0 + 4 * N * P0,
        "col_iir": 4 * N * P0 + 4 * N * P,
        "gather_lp": 4 * N * P + 3 * L,
        "rows_fwd": 3 * 2 * L,
        "cols": 3 * 2 * L,
        "rows_inv": 3 * 2 * L,
        "cart_out": 3 * L + 4 * N0 * N0,
    }
    return fwd, adj


def pipeline_bytes(g):
    """B_alg per slice of fwd+adj (lprec_amd.traffic: the compulsory traffic of
    the fused design, cubic)."""
    from lprec_amd import traffic
    return traffic.b_alg(g)


# Kernel name prefix -> stage of the pipeline (one fwd or adj call launches
# each stage once; the first matching prefix wins).
_STAGE_OF_KERNEL = (
    ("k_image_rows", "prep_rows"), ("k_sino_rows", "sino_rows"), ("k_col_iir", "col_iir"),
    ("k_gather1<0", "gather_lp"), ("k_gather1<1", "gather_lp"), ("k_add_extras", "gather_lp"),
    ("k_gather1<2", "polar_out"), ("k_gather<3", "cart_out"),
    ("k_rows16_fwd", "rows_fwd"), ("k_rows_fwd", "rows_fwd"), ("k_cols16", "cols"), ("k_cols", "cols"),
    ("k_rows16_inv", "rows_inv"), ("k_rows32_inv", "rows_inv"), ("k_rows_inv", "rows_inv"),
)


def _stage_of(kernel_name):
    base = kernel_name.replace("void ", "").split("(")[0].replace("lpk::", "")
    for key, stage in _STAGE_OF_KERNEL:
        if base.startswith(key):
            return stage
    return None


def pmc_stage_kib(csv_path):
    """Sum a rocprofv3 --pmc counter (KiB per dispatch) per pipeline stage and
    average over calls: -> {"fwd.<stage>": KiB per launch, "adj.<stage>": ...}.
    Dispatches are walked in order; a call starts at its first-stage kernel."""
    import csv
    import collections
    rows = sorted(csv.DictReader(open(csv_path)), key=lambda r: int(r["Dispatch_Id"]))
    tot, calls, d = collections.defaultdict(float), collections.Counter(), None
    for r in rows:
        st = _stage_of(r["Kernel_Name"])
        if st is None:
            continue
        if st in ("prep_rows", "sino_rows"):
            d = "fwd" if st == "prep_rows" else "adj"
            calls[d] += 1
        if d is not None:
            tot[d + "." + st] += float(r["Counter_Value"])
    return {k: v / calls[k[:3]] for k, v in tot.items()}


def pmc_copy_factor(csv_path, mib):
    """Bytes per counter KiB, from a known-size device copy (mib MiB read and
    written per launch; the largest-count kernel of the run is the copy)."""
    import csv
    vals = [float(r["Counter_Value"]) for r in csv.DictReader(open(csv_path))
            if "copyBuffer" in r["Kernel_Name"]]
    return mib * (1 << 20) / (sum(vals) / len(vals))


def pmc_summary(root, copy_mib):
    """Calibrated HBM traffic per stage launch from four rocprofv3 --pmc runs
    under root: fetch_bench/ and write_bench/ (FETCH_SIZE / WRITE_SIZE over the
    bench), fetch_copy/ and write_copy/ (the same counters over --copy-ref)."""
    csvp = lambda n: os.path.join(root, n, "run_counter_collection.csv")
    ff = pmc_copy_factor(csvp("fetch_copy"), copy_mib)
    fw = pmc_copy_factor(csvp("write_copy"), copy_mib)
    fetch = pmc_stage_kib(csvp("fetch_bench"))
    write = pmc_stage_kib(csvp("write_bench"))
    return {"fetch_bytes_per_KiB": ff, "write_bytes_per_KiB": fw,
            "stages": {k: [int(fetch[k] * ff), int(write.get(k, 0.0) * fw)] for k in fetch}}


SOURCES = ("kernels.hip", "handle.hip", "fft_lds.hpp", "fft64.hpp", "kernels.hpp", "lp_internal.hpp", "fft_ct.hpp",
           "rtc_kernels.hpp")
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_fwdadj_N2048_P1500_b32.json")


def sources_digest():
    """SHA-256 over the device sources: measured traffic is valid only for the
    kernels it was measured on."""
    import hashlib
    h = hashlib.sha256()
    for name in SOURCES:
        with open(os.path.join(ROOT, "lprec_amd", "csrc", name), "rb") as fh:
            h.update(name.encode() + b"\0" + fh.read())
    return h.hexdigest()


def measured_traffic():
    """Per-stage HBM bytes (read, written) per launch at the default workload
    from the committed rocprofv3 --pmc summary, or None when that summary was
    measured on other kernel sources than the ones present."""
    try:
        with open(TRAFFIC_FILE) as fh:
            rec = json.load(fh)
    except (OSError, ValueError):
        return None
    if rec.get("sources_sha256") != sources_digest():
        return None
    return {k: tuple(v) for k, v in rec["stages"].items()}


def stream_ceiling(mib=4096, reps=10):
    """GB/s of the library's streaming kernels (16-byte non-temporal loads and
    stores) over mib MiB, far beyond the 256 MB memory-side cache: copy (read
    + write bytes), read-only and write-only."""
    import torch
    from lprec_amd import _lib
    n = mib * (1 << 20) // 4
    a = torch.ones(n, device="cuda")
    b = torch.empty(n, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    out = {}
    for kind, name, nbytes in ((0, "copy", 2 * n * 4), (1, "read", n * 4), (2, "write", n * 4)):
        _lib.stream_bench(0, b.data_ptr(), a.data_ptr(), n, kind, 0, stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            _lib.stream_bench(0, b.data_ptr(), a.data_ptr(), n, kind, 0, stream)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name + "_GBps"] = round(nbytes / (ms * 1e-3) / 1e9, 1)
    del a, b
    torch.cuda.empty_cache()
    out["buffer_MiB"] = mib
    return out


def _host_cpu():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(lp, f0, budget_s, R_dev=None, back_dev=None):
    """Time the FP64 oracle mirror (numpy/scipy + the oracle's C sampler) on
    slices of the workload; when the device outputs of slice 0 are given,
    also report their relative L2 difference to the mirror's (fwd on the
    same image; adj on the device's own sinogram)."""
    import numpy as np
    from oracle import mirror
    threads = len(os.sched_getaffinity(0))
    tab = mirror.Tables(lp.glpars, lp.fwdparsi, lp.fwdparamsf, lp.adjparsi, lp.adjparamsf)
    t0 = time.perf_counter()
    n = 0
    R = None
    # slices 0, 1, ... of the batch, one at a time, until the next one would
    # pass the time budget (or the batch is done)
    while True:
        Rn = mirror.fwd(tab, np.asarray(f0[n % len(f0)][None], dtype=np.float64))
        mirror.adj(tab, Rn)
        if R is None:
            R = Rn
        n += 1
        el = time.perf_counter() - t0
        if el + el / n > budget_s or n >= len(f0):
            break
    dt = time.perf_counter() - t0
    out = {"value": n / dt, "unit": "slices/s", "cores": threads, "kind": "port",
           "sample": "slices 0..%d of the batch, fwd+adj each, through the FP64 oracle mirror "
                     "(scipy.fft workers=%d, OpenMP sampler), %.1f s, host CPU: %s" % (n - 1, threads, dt, _host_cpu())}
    if R_dev is not None:
        Rg = R_dev[:1].double().cpu().numpy()
        bg = back_dev[:1].double().cpu().numpy()
        bm = mirror.adj(tab, Rg)
        out["device_vs_mirror_slice0"] = {"fwd_rel_l2": float(np.linalg.norm(Rg - R) / np.linalg.norm(R)),
                                          "adj_rel_l2": float(np.linalg.norm(bg - bm) / np.linalg.norm(bm))}
    return out


def cpu_config1():
    """The oracle mirror on all 4 slices of the small plain configuration
    (N0 = 256, 384 angles, 12-ellipse phantoms): seconds and slices/s."""
    import numpy as np
    from oracle import mirror
    from lprec_amd import lpTransform
    from lprec_amd.phantom import phantom_stack
    lp = lpTransform(256, 384, 4, "none", 128, "cubic")
    lp.precompute(1)
    tab = mirror.Tables(lp.glpars, lp.fwdparsi, lp.fwdparamsf, lp.adjparsi, lp.adjparamsf)
    f = phantom_stack(4, 256, 12, seed=256).astype(np.float64)
    t0 = time.perf_counter()
    mirror.adj(tab, mirror.fwd(tab, f))
    dt = time.perf_counter() - t0
    return {"workload": "fwd+adj, N0=256, 384 angles, 4 slices", "seconds": round(dt, 3),
            "value": round(4 / dt, 2), "unit": "slices/s", "cores": len(os.sched_getaffinity(0)), "kind": "port"}


def _timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def _schedule(args):
    """--tune key=value,... -> schedule dict (None when empty)."""
    sched = {}
    for kv in filter(None, args.tune.split(",")):
        k, v = kv.split("=")
        sched[k] = int(v)
    return sched or None


def run_fbp(args):
    """FBP (adjp with a filter) of 32 noisy slices at N0 = 2048, 1500 angles and
    a centre offset drawn from U(-3, 3) px (padded grid N = 2064: FFT lengths
    2^4 3 43 and 2^5 3 43), ramp and Parzen filters."""
    import numpy as np
    import torch
    from lprec_amd import lpTransform
    from lprec_amd.phantom import scan
    N0, P0, ns = 2048, 1500, 32
    cor = N0 / 2 + float(np.random.default_rng(2064).uniform(-3, 3))
    f_true, R_host = scan(ns, N0, P0, 20, seed=2064, photons=1e4, cor=cor)
    R = torch.from_numpy(R_host).cuda()
    rec = torch.zeros((ns, N0, N0), device="cuda")
    out = {}
    for filt in ("ramp", "parzen"):
        lp = lpTransform(N0, P0, ns, filt, cor, "cubic")
        lp.precompute(0)
        lp.initcmem(0, 0, schedule=_schedule(args))
        h = lp.clphandle[0]
        g = h.info()
        for _ in range(args.warmup):
            lp.adjp(rec, R, 0)
        torch.cuda.synchronize()
        h.profile(True)
        dt = _timed(lambda: lp.adjp(rec, R, 0), args.steps, 0)
        stage_ms, ncalls = h.profile_read(1)
        h.profile(False)
        rate = ns * args.steps / dt
        from lprec_amd import traffic
        b_alg = traffic.b_adj(g, filtered=True)
        # reconstruction error against the phantom inside the disk (scale-free)
        x = rec.cpu().numpy().astype(np.float64)
        a = float(np.sum(x * f_true) / np.sum(x * x))
        err = float(np.linalg.norm(a * x - f_true) / np.linalg.norm(f_true))
        out[filt] = {"value": round(rate, 2), "ms_per_call": round(dt / args.steps * 1e3, 3),
                     "B_alg_bytes_per_slice": int(b_alg), "roofline_frac": round(b_alg * rate / HBM_PEAK, 4),
                     "rel_error_vs_phantom": round(err, 4),
                     "stage_ms_per_call": {k: round(v / max(ncalls, 1), 3) for k, v in stage_ms.items()}}
        lp.close()
    line = {"metric": "FBP slices/s at N0=2048, 1500 proj, centre offset (N=2064), 32 slices",
            "value": out["ramp"]["value"], "unit": "slices/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic: 20-ellipse phantoms, analytic sinograms, Poisson 1e4 photons",
            "config": {"workload": "FBP ramp/parzen, N0=2048, Nproj=1500, cor=%.3f (N=%d), 32 slices"
                                   % (cor, g["N"]), "N0": N0, "Nproj": P0, "cor": cor},
            "filters": out}
    print(json.dumps(line), flush=True)


def run_iterative(args):
    """EM reconstruction of 128 slices at N0 = 2048, 2048 angles, 50
    iterations, each iteration (fwd, ratio, adj, update) replayed from one
    captured HIP graph (lpmethods.GraphedEM).  Sinograms: 4 distinct
    30-ellipse phantoms (exact FP64 line integrals, Poisson noise at 5e3
    photons, -log), each repeated over 32 slices."""
    import numpy as np
    import torch
    from lprec_amd import lpTransform, lpmethods
    from lprec_amd.phantom import scan
    N0, P0, ns, iters, distinct = 2048, 2048, 128, 50, 4
    f8, g8 = scan(distinct, N0, P0, 30, seed=5000, photons=5e3)
    reps = ns // distinct
    tomo = torch.from_numpy(np.maximum(np.repeat(g8, reps, axis=0), 0)).cuda()
    lp = lpTransform(N0, P0, ns, "none", N0 / 2, "cubic")
    lp.precompute(1)
    lp.initcmem(1, 0, schedule=_schedule(args))
    g = lp.clphandle[0].info()
    init = torch.full((ns, N0, N0), 1e-3, device="cuda")
    em = lpmethods.GraphedEM(lp, init, tomo, 1e-3, 0)
    # eager reference timing of a few iterations, then the graph
    eager_steps = 3
    em_e = lpmethods.GraphedEM(lp, init, tomo, 1e-3, 0)
    t_eager = _timed(lambda: em_e.step(), eager_steps, 1) / eager_steps
    del em_e
    em.capture()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    em.step(iters)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    rate = ns * iters / dt
    b_alg = pipeline_bytes(g)
    # error against the phantom inside the disk (scale-free), per distinct slice
    x = em.recon[::reps].cpu().numpy().astype(np.float64)
    errs = []
    for i in range(distinct):
        a = float(np.sum(x[i] * f8[i]) / np.sum(x[i] * x[i]))
        errs.append(float(np.linalg.norm(a * x[i] - f8[i]) / np.linalg.norm(f8[i])))
    line = {"metric": "EM slice-iterations/s, 128 slices, N0=2048, 2048 proj, graph-captured fwd/adj step",
            "value": round(rate, 2), "unit": "slice-iterations/s", "n_gpus": 1, "steps": iters, "warmup": 0,
            "ms_per_iteration": round(dt / iters * 1e3, 3), "eager_ms_per_iteration": round(t_eager * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic: 4 distinct 30-ellipse phantoms x 32 slices, exact sinograms, Poisson 5e3 photons",
            "config": {"workload": "EM (lpmethods.em update, eps 1e-3), 128 slices, N0=2048, Nproj=2048, "
                                   "50 iterations, one HIP graph per iteration", "N0": N0, "Nproj": P0,
                       "slices": ns, "iterations": iters},
            "pipeline_roofline": {"B_alg_bytes_per_slice_iteration": int(b_alg),
                                  "frac": round(b_alg * rate / HBM_PEAK, 4)},
            "rel_error_vs_phantom": round(float(np.mean(errs)), 4),
            "reference_iterate": "the reference's CUDA/cupy path cannot run here; the graphed EM iterate is "
                                 "checked against the FP64 oracle EM (oracle/solvers.py) at this geometry "
                                 "in tests/test_gpu_batch.py"}
    print(json.dumps(line), flush=True)


def run_autograd(args):
    """Differentiable operator workload: 100 Adam steps of ||radon(f) - g||^2
    + 1e-3 TV(f) at N0 = 1024, 1024 angles, 16 slices (autograd backward =
    the exact transpose kernels), against noisy sinograms (1e4 photons) of
    seeded 8-ellipse phantoms."""
    import numpy as np
    import torch
    from lprec_amd import lpTransform
    from lprec_amd.autograd import adam_tv
    from lprec_amd.phantom import scan
    N0, P0, ns, iters = 1024, 1024, 16, 100
    f_true, g = scan(ns, N0, P0, 8, seed=1024, photons=1e4)
    tomo = torch.from_numpy(g).cuda()
    lp = lpTransform(N0, P0, ns, "none", N0 / 2, "cubic")
    lp.precompute(1)
    lp.initcmem(1, 0, schedule=_schedule(args))
    adam_tv(lp, tomo, num_iter=2, reg_par=1e-3, lr=0.1)          # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rec, losses = adam_tv(lp, tomo, num_iter=iters, reg_par=1e-3, lr=0.1)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ft = torch.from_numpy(f_true).cuda()
    err = float(torch.linalg.norm(rec - ft) / torch.linalg.norm(ft))
    line = {"metric": "Adam-TV steps/s through the autograd operators, N0=1024, 1024 proj, 16 slices",
            "value": round(iters / dt, 2), "unit": "steps/s", "n_gpus": 1, "steps": iters, "warmup": 2,
            "ms_per_step": round(dt / iters * 1e3, 3), "slices_per_s": round(ns * iters / dt, 2),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic: seeded 8-ellipse phantoms, exact sinograms, Poisson 1e4 photons",
            "config": {"workload": "Adam (lr 0.1) on ||radon(f)-g||^2 + 1e-3 TV(f), steepest-descent start, "
                                   "backward = exact transposes", "N0": N0, "Nproj": P0, "slices": ns},
            "loss_first_last": [losses[0], losses[-1]], "rel_error_vs_phantom": round(err, 4)}
    print(json.dumps(line), flush=True)


def run_multiscan(args):
    """96 scans, N0 from {512, 1024, 2048}, Nproj from {360, 720, 1500}, 1-16
    slices each: fwd+adj of every scan per step, one graph launch per step."""
    import numpy as np
    import torch
    from lprec_amd.batch import MultiScan, random_scans
    from lprec_amd.traffic import bytes_per_slice
    from lprec_amd.phantom import phantom_stack
    scans = random_scans(96, seed=96)
    t0 = time.perf_counter()
    ms = MultiScan(scans, cache_dir=os.environ.get("LPREC_AMD_CACHE"))
    t_init = time.perf_counter() - t0
    for i, s in enumerate(scans):
        ms.scan_views(i)[0].copy_(torch.from_numpy(phantom_stack(s.nslices, s.N0, 12, seed=i)))
    tuned = None
    if args.autotune:
        tuned = {k: v[0] for k, v in ms.autotune(reps=2, sweeps=1).items()}
    ms.capture()
    nsl = ms.nslices
    b_alg = sum(bytes_per_slice(s.N0, s.Nproj, s.cor) * s.nslices for s in scans)
    t_graph = _timed(ms.replay, args.steps, args.warmup)
    t_conc = _timed(lambda: ms.step_eager(True), args.steps, args.warmup)
    t_seq = _timed(lambda: ms.step_eager(False), args.steps, args.warmup)
    rate = nsl * args.steps / t_graph
    line = {"metric": "uneven multi-scan fwd+adj slices/s (96 scans, one graph launch per step)",
            "value": round(rate, 2), "unit": "slices/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(t_graph / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic: seeded random-ellipse phantoms",
            "config": {"workload": "96 scans, N0 in {512,1024,2048}, Nproj in {360,720,1500}, 1-16 slices",
                       "scans": len(scans), "slices": nsl, "geometries": len(ms.groups)},
            "scans_per_s": round(len(scans) * args.steps / t_graph, 2),
            "pipeline_roofline_frac": round(b_alg * args.steps / t_graph / HBM_PEAK, 4),
            "eager_concurrent_slices_per_s": round(nsl * args.steps / t_conc, 2),
            "eager_sequential_slices_per_s": round(nsl * args.steps / t_seq, 2),
            "init_s": round(t_init, 1), "schedules": tuned}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--slices", type=int, default=32)
    ap.add_argument("--N", type=int, default=2048)
    ap.add_argument("--nproj", type=int, default=1500)
    ap.add_argument("--plain-slices", type=int, default=8)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ceiling", action="store_true", help="skip the streaming-kernel HBM ceiling")
    ap.add_argument("--tune", default="", help="schedule knobs key=value,... (the keys of lprec_amd.tune.KNOBS)")
    ap.add_argument("--copy-ref", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--pmc-summary", default="", metavar="DIR",
                    help="print calibrated per-stage traffic from rocprofv3 --pmc runs under DIR and exit")
    ap.add_argument("--pmc-copy-mib", type=int, default=4096)
    ap.add_argument("--write-traffic", action="store_true",
                    help="with --pmc-summary: store the table, stamped with the device-source digest, as the "
                         "traffic the default run reports")
    ap.add_argument("--workload", default="fwdadj", choices=("fwdadj", "fbp", "multiscan", "em", "autograd"),
                    help="fwdadj: the headline line; fbp: FBP at the centre-offset geometry (ramp and "
                         "parzen); multiscan: 96 uneven scans in one graph launch; em: 50 graph-replayed "
                         "EM iterations on 128 slices at N0=2048, 2048 angles; autograd: 100 Adam-TV steps "
                         "at N0=1024 through the differentiable operators")
    ap.add_argument("--autotune", action="store_true", help="multiscan: autotune every geometry first")
    args = ap.parse_args()
    if args.pmc_summary:
        summ = pmc_summary(args.pmc_summary, args.pmc_copy_mib)
        if args.write_traffic:
            summ["sources_sha256"] = sources_digest()
            summ["workload"] = "fwd+adj, N0=2048, 1500 angles, 32 slices, default schedule"
            with open(TRAFFIC_FILE, "w") as fh:
                json.dump(summ, fh, indent=1)
        print(json.dumps(summ, indent=1))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path here "
              "(slices would shard across GPUs with no exchange). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    from lprec_amd import lpTransform
    from lprec_amd.phantom import phantom_stack

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.workload == "fbp":
        return run_fbp(args)
    if args.workload == "multiscan":
        return run_multiscan(args)
    if args.workload == "em":
        return run_iterative(args)
    if args.workload == "autograd":
        return run_autograd(args)
    if args.copy_ref:
        # known-size copy: reads and writes copy_ref MiB per launch
        n = args.copy_ref * (1 << 20) // 4
        a = torch.ones(n, device="cuda")
        b = torch.empty(n, device="cuda")
        b.copy_(a)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(max(args.steps, 1)):
            b.copy_(a)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1) / max(args.steps, 1)
        print(json.dumps({"copy_ref_MiB": args.copy_ref, "launches": max(args.steps, 1) + 1,
                          "ms_per_copy": round(ms, 4),
                          "read_plus_write_GBps": round(2 * args.copy_ref * (1 << 20) / (ms * 1e-3) / 1e9, 1)}))
        return
    N0, P0, ns = args.N, args.nproj, args.slices
    cor = N0 / 2
    from lprec_amd.tune import KNOBS as knobs
    sched = {}
    for kv in filter(None, args.tune.split(",")):
        k, v = kv.split("=")
        sched[k] = int(v)
    lp = lpTransform(N0, P0, ns, "none", cor, "cubic")
    lp.precompute(1)
    lp.initcmem(1, 0, schedule=sched)
    h = lp.clphandle[0]
    g = h.info()
    f_host = phantom_stack(ns, N0, 12, seed=2048)
    f = torch.from_numpy(f_host).cuda()
    R = torch.zeros((ns, P0, N0), device="cuda")
    back = torch.zeros((ns, N0, N0), device="cuda")

    def step():
        lp.fwdp(R, f, 0)
        lp.adjp(back, R, 0)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    h.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    fwd_ms, nf = h.profile_read(0)
    adj_ms, na = h.profile_read(1)
    h.profile(False)
    value = ns * args.steps / dt

    # plain path: same kernels, one slice per call, tables re-read every call
    nplain = min(args.plain_slices, ns)
    plain = None
    if nplain > 0:
        lp1 = lpTransform(N0, P0, 1, "none", cor, "cubic")
        lp1.precompute(1)
        lp1.initcmem(1, 0)
        fs = [f[i:i + 1].contiguous() for i in range(nplain)]
        R1 = torch.zeros((1, P0, N0), device="cuda")
        b1 = torch.zeros((1, N0, N0), device="cuda")
        for x in fs[:2]:
            lp1.fwdp(R1, x, 0)
            lp1.adjp(b1, R1, 0)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for x in fs:
            lp1.fwdp(R1, x, 0)
            lp1.adjp(b1, R1, 0)
        torch.cuda.synchronize()
        plain = nplain / (time.perf_counter() - t1)

    # roofline of the dominant stage (largest summed duration)
    sb_fwd, sb_adj = stage_bytes(g, fused_polar=fwd_ms.get("polar_out", 1.0) == 0.0)
    stages = {}
    for name, ms in fwd_ms.items():
        stages["fwd." + name] = (ms / max(nf, 1), sb_fwd[name] * ns)
    for name, ms in adj_ms.items():
        stages["adj." + name] = (ms / max(na, 1), sb_adj[name] * ns)
    stages = {k: v for k, v in stages.items() if v[1] > 0}
    dom = max(stages, key=lambda k: stages[k][0])
    dom_ms, dom_bytes = stages[dom]
    achieved = dom_bytes / (dom_ms * 1e-3)
    b_alg = pipeline_bytes(g)
    default_run = (N0, P0, ns, args.tune) == (2048, 1500, 32, "")
    table = measured_traffic() if default_run else None
    traffic = sum(table[dom]) if table and dom in table else None
    ceiling = stream_ceiling() if not args.no_ceiling else None
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded random-ellipse phantoms (12 ellipses per slice)",
        "config": {"workload": "lprec fwd+adj, N0=2048, Nproj=1500 on [0,pi), centre N0/2, cubic, no filter, "
                               "32 slices per batch (internal N=%d, os=%d, Nproj=%d, log-polar %dx%d)"
                               % (g["N"], g["os"], g["Nproj"], g["Nrho"], g["Ntheta"]),
                   "N0": N0, "Nproj": P0, "slices_per_batch": ns, "interp": "cubic", "filter": "none"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic,
                     "alg_bytes_per_launch": int(dom_bytes), "avg_launch_ms": round(dom_ms, 4),
                     "traffic_read_write": list(table[dom]) if traffic else None,
                     "measured_ceiling": ceiling,
                     "frac_of_copy_ceiling": (round(achieved / (ceiling["copy_GBps"] * 1e9), 4)
                                              if ceiling else None)},
        "pipeline_roofline": {"B_alg_bytes_per_slice": int(b_alg),
                              "achieved_GBps": round(b_alg * value / 1e9, 1),
                              "frac": round(b_alg * value / HBM_PEAK, 4),
                              "frac_of_copy_ceiling": (round(b_alg * value / (ceiling["copy_GBps"] * 1e9), 4)
                                                       if ceiling else None)},
        "plain_path": ({"value": round(plain, 2), "unit": "slices/s", "batched_over_plain": round(value / plain, 3)}
                       if plain else None),
        "stage_ms_per_step": {k: round(v[0], 4) for k, v in stages.items()},
        "schedule": {k: h.get_tuning(i) for k, i in knobs.items()},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(lp, f_host, args.cpu_budget, R, back)
        out["cpu_config1"] = cpu_config1()
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
