"""torch.autograd wrappers of the log-polar operators.

``radon(f, lp)`` applies ``lp.fwdp`` and ``backprojection(R, lp)`` applies
``lp.adjp`` (the reference's back-projection, FBP filter included); their
gradients run the hand-written HIP kernels directly, so a reconstruction loss
over them has no autograd tape inside the FFTs.  The backward of each
operator is its exact transpose (``fwdp_transpose`` / ``adjp_transpose``), and
the backward of a transpose is the operator again, so higher derivatives and
``torch.autograd.gradcheck``-style tests hold to FP32 rounding.  With
``backward="reference"`` the backward of ``radon`` is ``adjp`` and that of
``backprojection`` is ``fwdp`` -- the pairing the reference's iterative
methods use (lpmethods.py:43-44), which is adjoint only to ~1e-3
(initsadj.py:136 vs initsfwd.py:86).

Batches with more slices than the transform's ``Nslices`` are processed in
chunks.  Tensors must live on ``cuda:<gpu>`` and be float32.
"""
import torch

_OPS = {
    # name: (lpTransform method, output is a sinogram?)
    "fwd": ("fwdp", True),
    "fwd_T": ("fwdp_transpose", False),
    "adj": ("adjp", False),
    "adj_T": ("adjp_transpose", True),
}


def _apply(lp, name, x, gpu):
    method, to_sino = _OPS[name]
    x = x.contiguous()
    ns = x.shape[0]
    shape = (ns, lp.Nproj, lp.N) if to_sino else (ns, lp.N, lp.N)
    out = torch.empty(shape, dtype=torch.float32, device=x.device)
    step = lp.Nslices
    fn = getattr(lp, method)
    for s0 in range(0, ns, step):
        s1 = min(ns, s0 + step)
        fn(out[s0:s1], x[s0:s1], gpu)
    return out


class _LinearOp(torch.autograd.Function):
    """y = op(x) for a linear operator; the backward applies `partner`, whose
    own backward is `op` again (so double backward works)."""

    @staticmethod
    def forward(ctx, x, lp, name, gpu, partner):
        ctx.lp, ctx.gpu, ctx.name, ctx.partner = lp, gpu, name, partner
        return _apply(lp, name, x, gpu)

    @staticmethod
    def backward(ctx, g):
        gx = _LinearOp.apply(g, ctx.lp, ctx.partner, ctx.gpu, ctx.name)
        return gx, None, None, None, None


def _check(x, lp, name, gpu):
    if not isinstance(x, torch.Tensor) or not x.is_cuda or x.dtype != torch.float32 or x.dim() != 3:
        raise ValueError("%s needs a 3-D float32 CUDA tensor" % name)
    if x.device.index != gpu:
        raise ValueError("tensor is on cuda:%d, the operator on cuda:%d" % (x.device.index, gpu))


def radon(f, lp, gpu=0, backward="transpose"):
    """R = fwdp(f), differentiable.  f [ns, N, N] -> R [ns, Nproj, N]."""
    _check(f, lp, "radon", gpu)
    if backward == "transpose":
        return _LinearOp.apply(f, lp, "fwd", gpu, "fwd_T")
    if backward == "reference":
        return _LinearOp.apply(f, lp, "fwd", gpu, "adj")
    raise ValueError("backward must be 'transpose' or 'reference'")


def backprojection(R, lp, gpu=0, backward="transpose"):
    """f = adjp(R) (FBP filter included), differentiable.  R [ns, Nproj, N] -> f [ns, N, N]."""
    _check(R, lp, "backprojection", gpu)
    if backward == "transpose":
        return _LinearOp.apply(R, lp, "adj", gpu, "adj_T")
    if backward == "reference":
        return _LinearOp.apply(R, lp, "adj", gpu, "fwd")
    raise ValueError("backward must be 'transpose' or 'reference'")


def tv_norm(f, eps=1e-8):
    """Isotropic total variation sum sqrt(dx^2 + dy^2 + eps) per slice, summed
    (forward differences, zero flux at the far edges)."""
    dx = torch.zeros_like(f)
    dy = torch.zeros_like(f)
    dx[:, :, :-1] = f[:, :, 1:] - f[:, :, :-1]
    dy[:, :-1, :] = f[:, 1:, :] - f[:, :-1, :]
    return torch.sqrt(dx * dx + dy * dy + eps).sum()


def steepest_descent_start(lp, tomo, gpu=0):
    """alpha F^T tomo with alpha = <F F^T tomo, tomo> / ||F F^T tomo||^2 per
    batch: the first steepest-descent step of ||F f - tomo||^2 from f = 0."""
    with torch.no_grad():
        b = _apply(lp, "fwd_T", tomo, gpu)
        Fb = _apply(lp, "fwd", b, gpu)
        alpha = (Fb.double() * tomo.double()).sum() / (Fb.double() * Fb.double()).sum().clamp(min=1e-300)
        return (b * float(alpha)).contiguous()


def adam_tv(lp, tomo, num_iter=100, reg_par=1e-3, lr=0.1, init="sd", gpu=0, callback=None):
    """Minimise ||radon(f) - tomo||^2 + reg_par TV(f) with Adam (step ``lr``
    in image units, for images of order one).  ``init``: "sd" (the first
    steepest-descent step, default), "zero", or a tensor.

    ``tomo`` [ns, Nproj, N] (CUDA float32).  Returns (f, list of losses)."""
    dev = torch.device("cuda", gpu)
    tomo = tomo.to(dev, torch.float32).contiguous()
    if isinstance(init, torch.Tensor):
        f = init.to(dev, torch.float32).clone()
    elif init == "zero":
        f = torch.zeros((tomo.shape[0], lp.N, lp.N), device=dev)
    elif init == "sd":
        f = steepest_descent_start(lp, tomo, gpu)
    else:
        raise ValueError("init must be 'sd', 'zero' or a tensor")
    f.requires_grad_(True)
    opt = torch.optim.Adam([f], lr=lr)
    losses = []
    for it in range(num_iter):
        opt.zero_grad(set_to_none=True)
        r = radon(f, lp, gpu) - tomo
        loss = (r * r).sum() + reg_par * tv_norm(f)
        loss.backward()
        opt.step()
        # kept on the device: no host synchronisation per step
        losses.append(loss.detach())
        if callback is not None:
            callback(it, f.detach(), float(losses[-1]))
    return f.detach(), [float(v) for v in losses]
