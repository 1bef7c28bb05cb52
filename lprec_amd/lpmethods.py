"""Reconstruction methods built on ``lpTransform.fwdp``/``adjp`` (torch).

Counterpart of src/lprec/lpmethods.py:7-287: the same functions, arguments
and iterations (FBP, gradient descent with a Barzilai-Borwein step, conjugate
gradient on the normal equations, EM, and the Chambolle-Pock TV / TV-entropy /
TV-L1 schemes), with torch tensors on the operator's GPU in place of cupy.
Every fwd/adj pair runs on torch's current stream, so the element-wise
updates and the operator calls are ordered without host synchronisation.

``init_recon`` and ``tomo0`` may be numpy arrays (the result is returned as
numpy, as the reference's ``recon.get()`` does) or torch tensors (the result
stays a tensor on the operator's device).  ``lp`` is anything with the
``fwdp(R, f, gpu)`` / ``adjp(f, R, gpu)`` methods; its optional ``device``
attribute overrides ``cuda:<gpu>`` (used to run the iterations on the CPU
with a dense test operator).
"""
import numpy as np
import torch


def _device(lp, gpu):
    dev = getattr(lp, "device", None)
    return torch.device(dev) if dev is not None else torch.device("cuda", int(gpu))


def _setup(lp, init_recon, tomo0, gpu):
    dev = _device(lp, gpu)
    as_numpy = not isinstance(init_recon, torch.Tensor)
    recon = torch.as_tensor(np.asarray(init_recon) if as_numpy else init_recon, dtype=torch.float32,
                            device=dev).clone()
    tomo = torch.as_tensor(np.asarray(tomo0) if not isinstance(tomo0, torch.Tensor) else tomo0,
                           dtype=torch.float32, device=dev).contiguous()
    return recon.contiguous(), tomo, as_numpy


def _out(recon, as_numpy):
    return recon.cpu().numpy() if as_numpy else recon


def fbp(lp, init_recon, tomo0, num_iter, reg_par, gpu):
    """Filtered back-projection: one adj call (lpmethods.py:7-19).  The filter
    is the one the transform was built with."""
    recon, tomo, as_numpy = _setup(lp, init_recon, tomo0, gpu)
    lp.adjp(recon, tomo.clone(), gpu)
    return _out(recon, as_numpy)


def grad(lp, init_recon, tomo0, num_iter, reg_par, gpu):
    """Gradient descent on 1/2||R recon - tomo||^2 (lpmethods.py:22-59).
    reg_par < 0: per-slice Barzilai-Borwein step (1e-3 on the first
    iteration); otherwise the fixed step reg_par."""
    recon, tomo, as_numpy = _setup(lp, init_recon, tomo0, gpu)
    ns = tomo.shape[0]
    recon0 = recon
    grd = torch.zeros_like(recon)
    grd0 = torch.zeros_like(recon)
    g = torch.zeros_like(tomo)
    for i in range(num_iter):
        lp.fwdp(g, recon, gpu)
        lp.adjp(grd, (2 * (g - tomo)).contiguous(), gpu)
        if reg_par < 0:
            if i == 0:
                lam = torch.full((ns,), 1e-3, dtype=torch.float32, device=recon.device)
            else:
                dg = grd - grd0
                lam = ((recon - recon0) * dg).sum(dim=(1, 2)) / (dg * dg).sum(dim=(1, 2))
        else:
            lam = torch.full((ns,), float(reg_par), dtype=torch.float32, device=recon.device)
        recon0 = recon.clone()
        grd0 = grd.clone()
        recon = (recon - lam.view(ns, 1, 1) * grd).contiguous()
    return _out(recon, as_numpy)


def cg(lp, init_recon, tomo0, num_iter, reg_par, gpu):
    """Conjugate gradient on R*R recon = R* tomo (lpmethods.py:62-97); the
    inner products run over all slices together, as in the reference."""
    recon, tomo, as_numpy = _setup(lp, init_recon, tomo0, gpu)
    b = torch.zeros_like(recon)
    f = torch.zeros_like(recon)
    g = torch.zeros_like(tomo)
    lp.adjp(b, tomo.clone(), gpu)
    lp.fwdp(g, recon, gpu)
    lp.adjp(f, g, gpu)
    r = b - f
    p = r.clone()
    rsold = torch.sum(r * r)
    for _ in range(num_iter):
        lp.fwdp(g, p, gpu)
        lp.adjp(f, g, gpu)
        alpha = rsold / torch.sum(p * f)
        recon = recon + alpha * p
        r = r - alpha * f
        rsnew = torch.sum(r * r)
        p = (r + (rsnew / rsold) * p).contiguous()
        rsold = rsnew
    return _out(recon, as_numpy)


def em(lp, init_recon, tomo0, num_iter, reg_par, gpu):
    """Expectation maximisation, recon *= R*(tomo / (R recon + eps)) / R*(1)
    with eps = reg_par (lpmethods.py:100-128)."""
    recon, tomo, as_numpy = _setup(lp, init_recon, tomo0, gpu)
    xi = torch.zeros_like(recon)
    upd = torch.zeros_like(recon)
    g = torch.zeros_like(tomo)
    eps = np.float32(reg_par)
    lp.adjp(xi, torch.ones_like(tomo), gpu)
    xi = xi + 1e-5
    for _ in range(num_iter):
        lp.fwdp(g, recon, gpu)
        lp.adjp(upd, (tomo / (g + float(eps))).contiguous(), gpu)
        recon = (recon * (upd / xi)).contiguous()
    return _out(recon, as_numpy)


class GraphedEM:
    """EM iterations (the update of em()) with one whole iteration -- fwd,
    ratio, adj, multiplicative update -- recorded once into a HIP graph
    (torch.cuda.CUDAGraph) and replayed per iteration.  Every buffer is
    preallocated and every operation writes in place, so the recorded
    iteration allocates nothing; the arithmetic and its order are those of
    em().  ``step(n)`` runs n iterations (eagerly until ``capture()``)."""

    def __init__(self, lp, init_recon, tomo0, reg_par, gpu=0):
        self.recon, self.tomo, self.as_numpy = _setup(lp, init_recon, tomo0, gpu)
        self.lp, self.gpu = lp, gpu
        self.g = torch.zeros_like(self.tomo)
        self.ratio = torch.zeros_like(self.tomo)
        self.upd = torch.zeros_like(self.recon)
        self.xi = torch.zeros_like(self.recon)
        self.eps = float(np.float32(reg_par))
        lp.adjp(self.xi, torch.ones_like(self.tomo), gpu)
        self.xi += 1e-5
        self.graph = None

    def _iteration(self):
        self.lp.fwdp(self.g, self.recon, self.gpu)
        torch.add(self.g, self.eps, out=self.ratio)
        torch.div(self.tomo, self.ratio, out=self.ratio)
        self.lp.adjp(self.upd, self.ratio, self.gpu)
        self.upd.div_(self.xi)
        self.recon.mul_(self.upd)

    def capture(self):
        """Record one iteration (recording runs nothing, so the iterate does
        not advance)."""
        dev = self.recon.device
        torch.cuda.synchronize(dev)
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream(device=dev)   # default priority
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):
            with torch.cuda.graph(graph, stream=side):
                self._iteration()
        torch.cuda.current_stream(dev).wait_stream(side)
        self.graph = graph
        return graph

    def step(self, n=1):
        for _ in range(n):
            if self.graph is None:
                self._iteration()
            else:
                self.graph.replay()

    def result(self):
        return _out(self.recon.clone(), self.as_numpy)


def _tv_dual(recon, prox0x, prox0y, c, lam):
    # proximal step of the TV dual (lpmethods.py:158-163)
    prox0x[:, :, :-1] += c * (recon[:, :, 1:] - recon[:, :, :-1])
    prox0y[:, :-1, :] += c * (recon[:, 1:, :] - recon[:, :-1, :])
    nprox = torch.clamp(torch.sqrt(prox0x * prox0x + prox0y * prox0y) / lam, min=1.0)
    return prox0x / nprox, prox0y / nprox


def _div(prox0x, prox0y):
    # discrete divergence (lpmethods.py:170-173)
    div0 = torch.empty_like(prox0x)
    div0[:, :, 1:] = prox0x[:, :, 1:] - prox0x[:, :, :-1]
    div0[:, :, 0] = prox0x[:, :, 0]
    div0[:, 1:, :] += prox0y[:, 1:, :] - prox0y[:, :-1, :]
    div0[:, 0, :] += prox0y[:, 0, :]
    return div0


def _chambolle_pock(lp, init_recon, tomo0, num_iter, reg_par, gpu, prox_data):
    recon, tomo, as_numpy = _setup(lp, init_recon, tomo0, gpu)
    g = torch.zeros_like(tomo)
    p = torch.zeros_like(recon)
    recon0 = recon
    prox0x = torch.zeros_like(recon)
    prox0y = torch.zeros_like(recon)
    prox1 = torch.zeros_like(tomo)
    lam = reg_par
    c = 0.35  # the reference's fixed step, 1/power-method estimate
    for _ in range(num_iter):
        prox0x, prox0y = _tv_dual(recon, prox0x, prox0y, c, lam)
        lp.fwdp(g, recon, gpu)
        prox1 = prox_data(prox1, g, tomo, c).contiguous()
        recon = recon0
        div0 = _div(prox0x, prox0y)
        lp.adjp(p, prox1, gpu)
        recon0 = recon0 - c * p + c * div0
        recon = (2 * recon0 - recon).contiguous()
    return _out(recon, as_numpy)


def tv(lp, init_recon, tomo0, num_iter, reg_par, gpu):
    """1/2||R recon - tomo||^2 + reg_par TV(recon), Chambolle-Pock
    (lpmethods.py:131-180)."""
    return _chambolle_pock(lp, init_recon, tomo0, num_iter, reg_par, gpu,
                           lambda prox1, g, tomo, c: (prox1 + c * g - c * tomo) / (1 + c))


def tve(lp, init_recon, tomo0, num_iter, reg_par, gpu):
    """sum(R recon - tomo log R recon) + reg_par TV(recon) (lpmethods.py:182-234)."""
    def prox(prox1, g, tomo, c):
        a = prox1 + c * g - 1
        tmp = 0.5 * (a + torch.sqrt(a * a + 4 * c * tomo))
        tmp = torch.clamp(tmp, min=0.0)
        return prox1 + c * g - tmp
    return _chambolle_pock(lp, init_recon, tomo0, num_iter, reg_par, gpu, prox)


def tvl1(lp, init_recon, tomo0, num_iter, reg_par, gpu):
    """||R recon - tomo||_1 + reg_par TV(recon) (lpmethods.py:237-288)."""
    def prox(prox1, g, tomo, c):
        tmp = prox1 + c * g - c * tomo
        tmp2 = torch.clamp(torch.abs(tmp) - 1, min=0.0)
        return tmp - tmp2 * torch.sign(tmp)
    return _chambolle_pock(lp, init_recon, tomo0, num_iter, reg_par, gpu, prox)


METHODS = {"fbp": fbp, "grad": grad, "cg": cg, "em": em, "tv": tv, "tve": tve, "tvl1": tvl1}
