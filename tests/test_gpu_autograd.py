"""Differentiable operators (lprec_amd.autograd) on the GPU.

Workload of the differentiable-operator case: N = 1024, 1024 angles, 16
slices, seeded ellipse phantoms, noisy sinograms (1e4 photons).

* central finite differences of L(f) = 1/2 ||radon(f) - g||^2 on 64 random
  pixel directions with h = 1e-3, taken at f = 0: there the two evaluations
  are +-h F e and the quotient is free of the FP32 rounding of F f, which at
  a non-zero f (|L| ~ 1e4 per slice) swamps a 2 h <F e, r> ~ 1e-3 difference;
  and at f = phantom/2 with h = 1 (L is quadratic, so the quotient is exact
  up to that rounding, ~1e-3);
* the same for the back-projection with the Parzen FBP filter;
* double backward: the gradient of <F^T u, v> with respect to u is F v;
* 100 Adam steps of ||radon(f) - g||^2 + 1e-3 TV(f).
"""
import numpy as np
import pytest

from lprec_amd import lpTransform
from lprec_amd.autograd import adam_tv, backprojection, radon
from lprec_amd.phantom import scan

pytestmark = pytest.mark.gpu

N0, P0, NS = 1024, 1024, 16


@pytest.fixture(scope="module")
def setup(torch_cuda):
    torch = torch_cuda
    lp = lpTransform(N0, P0, NS, "none", N0 / 2, "cubic")
    lp.precompute(1)
    lp.initcmem(1, 0)
    f, g = scan(NS, N0, P0, 8, seed=1024, photons=1e4)
    return lp, torch.from_numpy(f).cuda(), torch.from_numpy(g).cuda()


def _fd_check(torch, op, lp, x0, target, h, backward):
    """-> (max |fd - grad| / max |grad|) over 64 random entries of x0."""
    x = x0.clone().requires_grad_(True)
    y = op(x, lp, 0, backward=backward)
    loss = 0.5 * ((y.double() - target.double()) ** 2).sum()
    (grad,) = torch.autograd.grad(loss, x)
    rng = np.random.default_rng(64)
    idx = [tuple(int(rng.integers(0, n)) for n in x0.shape) for _ in range(64)]
    fd, gv = [], []
    with torch.no_grad():
        for (s, i, j) in idx:
            xs = x0[s:s + 1].clone()
            vals = []
            for sign in (1.0, -1.0):
                xp = xs.clone()
                xp[0, i, j] += sign * h
                yp = op(xp, lp, 0)
                vals.append(0.5 * float(((yp.double() - target[s:s + 1].double()) ** 2).sum()))
            fd.append((vals[0] - vals[1]) / (2 * h))
            gv.append(float(grad[s, i, j]))
    fd, gv = np.array(fd), np.array(gv)
    return float(np.max(np.abs(fd - gv)) / np.max(np.abs(gv)))


def test_radon_gradient_finite_differences(setup, torch_cuda):
    lp, f, g = setup
    zero = torch_cuda.zeros_like(f)
    e_small = _fd_check(torch_cuda, radon, lp, zero, g, 1e-3, "transpose")
    e_ref = _fd_check(torch_cuda, radon, lp, zero, g, 1e-3, "reference")
    e_mid = _fd_check(torch_cuda, radon, lp, 0.5 * f, g, 1.0, "transpose")
    print("radon FD: h=1e-3 at 0: %.2e (reference adj as the gradient: %.2e); h=1 at f/2: %.2e"
          % (e_small, e_ref, e_mid), flush=True)
    assert e_small < 1e-4
    assert e_mid < 5e-3


def test_backprojection_gradient_finite_differences(torch_cuda):
    torch = torch_cuda
    n0, p0, ns = 512, 600, 2
    lp = lpTransform(n0, p0, ns, "parzen", n0 / 2 + 1.5, "cubic")
    lp.precompute(1)
    lp.initcmem(1, 0)
    rng = np.random.default_rng(5)
    R0 = torch.from_numpy(rng.standard_normal((ns, p0, n0)).astype(np.float32)).cuda()
    tgt = torch.from_numpy(rng.standard_normal((ns, n0, n0)).astype(np.float32)).cuda()
    e_small = _fd_check(torch, backprojection, lp, torch.zeros_like(R0), tgt, 1e-3, "transpose")
    e_mid = _fd_check(torch, backprojection, lp, R0, tgt, 1.0, "transpose")
    print("backprojection FD: h=1e-3 at 0: %.2e; h=1 at R0: %.2e" % (e_small, e_mid), flush=True)
    assert e_small < 1e-4
    assert e_mid < 5e-3


def test_double_backward(setup, torch_cuda):
    torch = torch_cuda
    lp, f, g = setup
    u = torch.randn_like(g).requires_grad_(True)
    v = torch.randn_like(f)
    x = f.clone().requires_grad_(True)
    (gx,) = torch.autograd.grad((radon(x, lp) * u).sum(), x, create_graph=True)   # F^T u
    (gu,) = torch.autograd.grad((gx * v).sum(), u)                                  # F v
    Fv = torch.zeros_like(g)
    lp.fwdp(Fv, v, 0)
    assert torch.allclose(gu, Fv, rtol=0, atol=1e-6 * float(Fv.abs().max()))


def test_adam_tv_reconstruction(setup, torch_cuda):
    torch = torch_cuda
    lp, f, g = setup
    rec, losses = adam_tv(lp, g, num_iter=100, reg_par=1e-3, lr=0.1)
    err = float(torch.linalg.norm(rec - f) / torch.linalg.norm(f))
    print("Adam-TV: loss %.4e -> %.4e, relative error to the phantom %.3f" % (losses[0], losses[-1], err), flush=True)
    # (16 CG iterations on the same data reach 0.147)
    assert losses[-1] < 0.02 * losses[0]
    assert err < 0.2
