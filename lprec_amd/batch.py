"""Many scans of uneven size in one graph launch.

A ``MultiScan`` takes a list of scans (N0, Nproj, slices, cor), groups them by
geometry, and builds one operator handle per geometry holding all of that
geometry's slices, so one fwd and one adj call serve every scan of the group
(the reference runs one lpTransform per geometry and loops over slice chunks,
tests/test_gpus_many.py:11-84).  Each group's fwd -> adj chain is issued on a
stream of its own; ``capture()`` records all chains, forked from and joined
back to one capture stream, into a single HIP graph (torch.cuda.CUDAGraph),
and ``replay()`` launches the whole batch at once.  Inputs and outputs live in
per-group device buffers; ``scan_views(i)`` gives scan i's slices of them.
"""
import torch

from .lpTransform import lpTransform
from . import tune as tuner
from .traffic import bytes_per_slice  # B_alg of one fwd+adj slice (shared with bench.py)


class Scan:
    def __init__(self, N0, Nproj, nslices, cor=None):
        self.N0, self.Nproj, self.nslices = int(N0), int(Nproj), int(nslices)
        self.cor = self.N0 / 2 if cor is None else cor

    def key(self):
        return (self.N0, self.Nproj, float(self.cor))


class _Group:
    def __init__(self, key, scans, filter_type, interp, gpu, cache_dir):
        N0, P0, cor = key
        self.scans = scans                      # {index into MultiScan.scans: Scan}
        self.offsets = []                       # (scan index, first slice, end slice)
        total = 0
        self.N0, self.P0 = N0, P0
        self.lp = lpTransform(N0, P0, max(1, sum(s.nslices for s in scans.values())), filter_type, cor, interp)
        self.lp.precompute(1, cache_dir=cache_dir)
        self.lp.initcmem(1, gpu)
        for i, s in scans.items():
            self.offsets.append((i, total, total + s.nslices))
            total += s.nslices
        self.total = total
        dev = torch.device("cuda", gpu)
        self.f = torch.zeros((total, N0, N0), device=dev)
        self.R = torch.zeros((total, P0, N0), device=dev)
        self.back = torch.zeros((total, N0, N0), device=dev)
        self.stream = torch.cuda.Stream(device=dev)   # default priority

    def run(self, gpu):
        self.lp.fwdp(self.R, self.f, gpu)
        self.lp.adjp(self.back, self.R, gpu)


class MultiScan:
    """fwd then adj of every scan in a list, grouped by geometry."""

    def __init__(self, scans, gpu=0, filter_type="none", interp="cubic", cache_dir=None):
        self.gpu = int(gpu)
        self.scans = list(scans)
        by_key = {}
        for i, s in enumerate(self.scans):
            by_key.setdefault(s.key(), {})[i] = s
        # largest groups first: their chains start first in eager mode
        keys = sorted(by_key, key=lambda k: -sum(s.nslices for s in by_key[k].values()) * k[0] * k[0])
        self.groups = [_Group(k, by_key[k], filter_type, interp, self.gpu, cache_dir) for k in keys]
        self.where = {}
        for gi, g in enumerate(self.groups):
            for i, a, b in g.offsets:
                self.where[i] = (gi, a, b)
        self.graph = None

    @property
    def nslices(self):
        return sum(s.nslices for s in self.scans)

    def scan_views(self, i):
        """-> (f, R, back) views of scan i in its group's buffers."""
        gi, a, b = self.where[i]
        g = self.groups[gi]
        return g.f[a:b], g.R[a:b], g.back[a:b]

    def autotune(self, reps=2, sweeps=1, schedule_file=None, log=None):
        """Tune every group's handle for its batch (or re-apply stored schedules)."""
        out = {}
        for g in self.groups:
            sched = tuner.load_schedule(schedule_file, g.lp, g.total) if schedule_file else None
            if sched is not None:
                tuner.set_schedule(g.lp.clphandle[self.gpu], sched)
                out[tuner.schedule_key(g.lp, g.total)] = (sched, None)
                continue
            best, ms, _ = tuner.autotune(g.lp, self.gpu, g.total, reps=reps, sweeps=sweeps, log=log)
            out[tuner.schedule_key(g.lp, g.total)] = (best, ms)
            if schedule_file:
                tuner.store_schedule(schedule_file, g.lp, g.total, best)
        return out

    def step_eager(self, concurrent=True):
        """One fwd+adj of every scan, issued now (group chains on their own
        streams when ``concurrent``, else one after another on the current
        stream)."""
        cur = torch.cuda.current_stream(self.gpu)
        if not concurrent:
            for g in self.groups:
                g.run(self.gpu)
            return
        for g in self.groups:
            g.stream.wait_stream(cur)
            with torch.cuda.stream(g.stream):
                g.run(self.gpu)
        for g in self.groups:
            cur.wait_stream(g.stream)

    def capture(self):
        """Record step_eager() into one graph (after one eager warm-up step)."""
        self.step_eager()
        torch.cuda.synchronize(self.gpu)
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream(device=torch.device("cuda", self.gpu))
        side.wait_stream(torch.cuda.current_stream(self.gpu))
        with torch.cuda.stream(side):
            with torch.cuda.graph(graph, stream=side):
                self.step_eager()
        torch.cuda.current_stream(self.gpu).wait_stream(side)
        self.graph = graph
        return graph

    def replay(self):
        if self.graph is None:
            raise RuntimeError("capture() first")
        self.graph.replay()


def random_scans(n, seed, sizes=(512, 1024, 2048), angles=(360, 720, 1500), slices=(1, 16)):
    """n scans with N0, Nproj drawn uniformly from the given sets and 1..16
    slices each (the uneven multi-scan workload)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return [Scan(int(rng.choice(sizes)), int(rng.choice(angles)), int(rng.integers(slices[0], slices[1] + 1)))
            for _ in range(n)]


__all__ = ["Scan", "MultiScan", "random_scans", "bytes_per_slice"]
