"""Schedule autotuner for one operator handle.

The kernels expose seven schedule knobs through ``lp_set_tuning`` that can
change after the tables are built (log-polar rows per row-FFT block, slices
per FFT block, prefilter column chunk, slices per gather block, FFT kernel
family, the forward operator's fused inverse row pass + polar gather, slices
per block of the polar output gather).  ``autotune`` times fwd+adj (and,
optionally, the transposes) on the handle's own stream with device events
and walks the knobs one at a time (coordinate descent, two sweeps), keeping
the fastest setting of each; settings the handle rejects (LDS or register
budget) are skipped.  Results can be stored per geometry and slice count in a
JSON file and re-applied without timing.
"""
import json
import os

import torch

from ._lib import LprecError

KNOBS = {"rows": 0, "zgroup": 1, "colchunk": 2, "gzgroup": 3, "fft16": 4, "boxcap": 5, "fuse": 6, "polart": 7,
         "staged": 8, "boxstride": 9, "rtc": 10, "rtccols": 11, "rtcvar": 12, "pzgroup": 13, "direct": 14}
# knobs that can change after the tables are built (boxcap and boxstride are fixed at init)
RUNTIME_KNOBS = ("rows", "zgroup", "colchunk", "gzgroup", "fft16", "fuse", "pzgroup")
SPACE = {
    "rows": (4, 8, 16),
    "zgroup": (1, 2, 4, 8),
    "gzgroup": (1, 2, 4, 8),
    "pzgroup": (4, 8, 16),
    "colchunk": (32, 64, 128),
    "fft16": (2, 1, 0),
    "fuse": (1, 0),
}


def get_schedule(h):
    return {k: h.get_tuning(KNOBS[k]) for k in RUNTIME_KNOBS}


def set_schedule(h, sched):
    """Apply a schedule; raises LprecError for values the handle rejects."""
    for k, v in sched.items():
        h.set_tuning(KNOBS[k], int(v))


def _time(lp, gpu, f, R, back, reps, transposes):
    stream = torch.cuda.current_stream(gpu)
    lp.fwdp(R, f, gpu)                      # warm-up with this schedule
    lp.adjp(back, R, gpu)
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(reps):
        lp.fwdp(R, f, gpu)
        lp.adjp(back, R, gpu)
        if transposes:
            lp.fwdp_transpose(back, R, gpu)
            lp.adjp_transpose(R, back, gpu)
    stop.record(stream)
    stop.synchronize()
    return start.elapsed_time(stop) / reps


def autotune(lp, gpu=0, nslices=None, reps=3, sweeps=2, transposes=False, space=None, log=None):
    """Tune the handle ``lp.clphandle[gpu]`` in place for batches of
    ``nslices`` (default: Nslices).  -> (best schedule dict, best ms per
    fwd+adj, list of (schedule, ms) tried)."""
    space = dict(SPACE if space is None else space)
    h = lp.clphandle[gpu]
    ns = lp.Nslices if nslices is None else int(nslices)
    dev = torch.device("cuda", gpu)
    g = torch.Generator(device=dev).manual_seed(0)
    f = torch.rand((ns, lp.N, lp.N), device=dev, generator=g)
    R = torch.zeros((ns, lp.Nproj, lp.N), device=dev)
    back = torch.zeros((ns, lp.N, lp.N), device=dev)
    best = get_schedule(h)
    best_ms = _time(lp, gpu, f, R, back, reps, transposes)
    tried = [(dict(best), best_ms)]
    for _ in range(sweeps):
        changed = False
        for knob, values in space.items():
            for v in values:
                if v == best[knob]:
                    continue
                cand = dict(best)
                cand[knob] = v
                try:
                    set_schedule(h, cand)
                except LprecError:
                    set_schedule(h, best)
                    continue
                ms = _time(lp, gpu, f, R, back, reps, transposes)
                tried.append((dict(cand), ms))
                if log:
                    log("%s -> %.3f ms" % (cand, ms))
                if ms < best_ms:
                    best, best_ms, changed = cand, ms, True
                set_schedule(h, best)
        if not changed:
            break
    set_schedule(h, best)
    return best, best_ms, tried


def schedule_key(lp, nslices):
    return "N%d_P%d_c%r_%s_%s_ns%d" % (lp.N, lp.Nproj, float(lp.cor), "cubic" if lp.cubic else "linear",
                                      lp.filter_type, int(nslices))


def load_schedule(path, lp, nslices):
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get(schedule_key(lp, nslices))


def store_schedule(path, lp, nslices, sched):
    data = {}
    if os.path.exists(path):
        with open(path) as fh:
            data = json.load(fh)
    data[schedule_key(lp, nslices)] = sched
    tmp = path + ".tmp%d" % os.getpid()
    with open(tmp, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)
