"""Host logic of the reconstruction methods, the slice scheduler and the
precompute cache (CPU only).

The iteration logic of lprec_amd.lpmethods (lpmethods.py:7-287 of the
reference) is exercised on the CPU with small dense operators whose answers
are known in closed form or from numpy.linalg; the GPU runs of the same
functions against the reference's recorded norms are in test_gpu_methods.py.
"""
import numpy as np
import pytest
import torch

from lprec_amd import lpmethods, precache
from lprec_amd.recon import check_uniform_theta, lprec


class DenseOp:
    """fwdp/adjp of a dense per-slice matrix A [m, n*n] on CPU tensors."""

    device = "cpu"

    def __init__(self, A, n):
        self.A = torch.as_tensor(A, dtype=torch.float32)
        self.n = n
        self.calls = 0

    def fwdp(self, R, f, gpu):
        self.calls += 1
        ns = f.shape[0]
        R.copy_((f.reshape(ns, -1) @ self.A.T).reshape(R.shape))

    def adjp(self, f, R, gpu):
        ns = R.shape[0]
        f.copy_((R.reshape(ns, -1) @ self.A).reshape(f.shape))


def _problem(seed=0, n=4, m=24, ns=2):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n * n)) / np.sqrt(m)
    x = rng.standard_normal((ns, n, n))
    b = (x.reshape(ns, -1) @ A.T).reshape(ns, m, 1)
    return A, x, b.astype(np.float32)


def test_cg_reaches_least_squares():
    A, x, b = _problem()
    op = DenseOp(A, 4)
    out = lpmethods.cg(op, np.zeros((2, 4, 4), np.float32), b, 40, 0, 0)
    assert isinstance(out, np.ndarray)
    assert np.allclose(out, x, atol=2e-3)


def test_grad_fixed_step_closed_form():
    # identity operator: x_k = b (1 - (1 - 2 lam)^k)
    ident = DenseOp(np.eye(9), 3)
    b = np.random.default_rng(1).standard_normal((1, 9, 1)).astype(np.float32)
    lam, k = 0.1, 7
    out = lpmethods.grad(ident, np.zeros((1, 3, 3), np.float32), b, k, lam, 0)
    assert np.allclose(out.reshape(-1), b.reshape(-1) * (1 - (1 - 2 * lam) ** k), atol=1e-6)


def test_grad_barzilai_borwein_converges():
    # (once the iterate is exact, the BB quotient is 0/0 in the reference too,
    # so the run stops short of that)
    A, x, b = _problem(2)
    out = lpmethods.grad(DenseOp(A, 4), np.zeros((2, 4, 4), np.float32), b, 60, -1, 0)
    assert np.allclose(out, x, atol=1e-3)


def test_em_identity_fixed_point():
    # identity, eps = 0: one step gives b / (1 + 1e-5) and stays there
    ident = DenseOp(np.eye(4), 2)
    b = np.array([[[1.0], [2.0], [0.5], [3.0]]], np.float32)
    out = lpmethods.em(ident, np.ones((1, 2, 2), np.float32), b, 5, 0.0, 0)
    assert np.allclose(out.reshape(-1), b.reshape(-1) / (1 + 1e-5), rtol=1e-6)


def _dense_fns(A, n):
    ns_shape = lambda v: v.reshape(v.shape[0], -1)  # noqa: E731
    return (lambda f: (ns_shape(f) @ A.T).reshape(f.shape[0], -1, 1),
            lambda R: (ns_shape(R) @ A).reshape(R.shape[0], n, n))


def test_oracle_grad_pins():
    """oracle/solvers.py grad (FP64): identity operator with a fixed step has
    the closed form b (1 - (1 - 2 lam)^k); the Barzilai-Borwein run reaches
    the least-squares solution; its second step equals the BB quotient of
    the first two iterates computed from the definition."""
    from oracle import solvers
    fwd, adj = _dense_fns(np.eye(9), 3)
    b = np.random.default_rng(1).standard_normal((1, 9, 1))
    out, _ = solvers.grad(fwd, adj, np.zeros((1, 3, 3)), b, 7, 0.1)
    assert np.allclose(out.reshape(-1), b.reshape(-1) * (1 - 0.8 ** 7), atol=1e-12)
    A, x, b = _problem(2)
    fwd, adj = _dense_fns(A, 4)
    b = b.astype(np.float64)
    out, kept = solvers.grad(fwd, adj, np.zeros((2, 4, 4)), b, 120, -1, keep=(1, 2))
    xls = np.stack([np.linalg.lstsq(A, b[z, :, 0], rcond=None)[0] for z in range(2)]).reshape(2, 4, 4)
    assert np.allclose(out, xls, atol=1e-10)
    # step 2 from the definition: s = x1 - x0, y = g1 - g0, lam = <s, y> / <y, y>
    g = lambda v: adj(2 * (fwd(v) - b))  # noqa: E731
    x0 = np.zeros((2, 4, 4))
    x1 = x0 - 1e-3 * g(x0)
    s_, y_ = x1 - x0, g(x1) - g(x0)
    lam = np.sum(s_ * y_, axis=(1, 2)) / np.sum(y_ * y_, axis=(1, 2))
    assert np.allclose(kept[1], x1, atol=1e-14)
    assert np.allclose(kept[2], x1 - lam[:, None, None] * g(x1), atol=1e-12)


def test_oracle_em_pins():
    """oracle/solvers.py em (FP64): on a consistent positive system the
    iterates stay positive, the Kullback-Leibler objective decreases
    monotonically, and they approach the exact solution."""
    from oracle import solvers
    rng = np.random.default_rng(4)
    n, m = 3, 20
    A = rng.uniform(0.1, 1.0, (m, n * n))
    x = rng.uniform(0.5, 2.0, (1, n, n))
    fwd, adj = _dense_fns(A, n)
    b = fwd(x)
    kl = []
    xk = np.ones((1, n, n))
    for _ in range(4):
        xk, _ = solvers.em(fwd, adj, xk, b, 500, 0.0)
        Ax = fwd(xk)
        kl.append(float(np.sum(b * np.log(b / Ax) - b + Ax)))
        assert np.all(xk > 0)
    # (the 1e-5 added to R*(1) moves the fixed point by ~1e-5)
    assert all(k1 < k0 for k0, k1 in zip(kl, kl[1:]))
    assert kl[-1] < 1e-5 * kl[0] and np.max(np.abs(xk - x)) < 1e-4


@pytest.mark.parametrize("name", ["tv", "tvl1", "tve"])
def test_chambolle_pock_without_regularisation_recovers_data(name):
    # identity operator, vanishing TV weight: the minimiser is b (b > 0 for tve)
    ident = DenseOp(np.eye(16), 4)
    b = np.random.default_rng(3).uniform(0.5, 2.0, (2, 16, 1)).astype(np.float32)
    out = lpmethods.METHODS[name](ident, np.zeros((2, 4, 4), np.float32), b, 400, 1e-8, 0)
    assert np.allclose(out.reshape(2, -1), b.reshape(2, -1), atol=1e-3), np.abs(out.reshape(2, -1) - b.reshape(2, -1)).max()


def test_tv_smooths():
    # a strong TV weight lowers the total variation of the reconstruction
    ident = DenseOp(np.eye(64), 8)
    rng = np.random.default_rng(4)
    b = rng.standard_normal((1, 64, 1)).astype(np.float32)
    weak = lpmethods.tv(ident, np.zeros((1, 8, 8), np.float32), b, 300, 1e-6, 0)
    strong = lpmethods.tv(ident, np.zeros((1, 8, 8), np.float32), b, 300, 0.5, 0)

    def tvnorm(u):
        return np.abs(np.diff(u, axis=1)).sum() + np.abs(np.diff(u, axis=2)).sum()
    assert tvnorm(strong) < 0.5 * tvnorm(weak)


def test_methods_accept_tensors_and_return_tensors():
    A, x, b = _problem(5)
    out = lpmethods.cg(DenseOp(A, 4), torch.zeros(2, 4, 4), torch.from_numpy(b), 40, 0, 0)
    assert isinstance(out, torch.Tensor)
    assert torch.allclose(out, torch.from_numpy(x).float(), atol=2e-3)


def test_fbp_is_one_adjoint():
    A, x, b = _problem(6)
    op = DenseOp(A, 4)
    out = lpmethods.fbp(op, np.zeros((2, 4, 4), np.float32), b, 0, 0, 0)
    assert np.allclose(out.reshape(2, -1), b.reshape(2, -1) @ A, atol=1e-5)


def test_lprec_scheduler_splits_chunks_over_devices():
    A, _, _ = _problem(7, ns=1)
    rng = np.random.default_rng(8)
    tomo = rng.standard_normal((7, 24, 1)).astype(np.float32)
    ops = []

    def make(chunk):
        op = DenseOp(A, 4)
        ops.append(chunk)
        return op
    # 'N' of the fake sinogram is 1 column; the image side is 4x4 per slice
    one = lprec(tomo, 0.5, np.zeros((7, 4, 4), np.float32), lpmethod="cg", num_iter=10, gpu_list=[0],
                chunk=3, make_transform=make)
    two = lprec(tomo, 0.5, np.zeros((7, 4, 4), np.float32), lpmethod="cg", num_iter=10, gpu_list=[0, 1],
                chunk=3, make_transform=make)
    assert ops == [3, 3]
    assert np.allclose(one, two, atol=1e-6)
    with pytest.raises(ValueError):
        lprec(tomo, 0.5, lpmethod="sirt", make_transform=make)


def test_theta_must_be_uniform():
    check_uniform_theta(np.arange(10) * np.pi / 10, 10)
    with pytest.raises(ValueError):
        check_uniform_theta(np.linspace(0, np.pi, 10), 10)
    with pytest.raises(ValueError):
        check_uniform_theta(np.arange(9) * np.pi / 9, 10)


def test_precompute_cache_round_trip(tmp_path):
    from lprec_amd import _lib
    from lprec_amd.lpTransform import lpTransform
    d = str(tmp_path)
    assert precache.load(d, 64, 48, 3, 32.5, True, "ramp", True) is None
    lp = lpTransform(64, 48, 3, "ramp", 32.5, "cubic")
    lp.precompute(0, cache_dir=d)                 # stores adj-only tables
    assert precache.load(d, 64, 48, 5, 32.5, True, "ramp", True) is None
    lp.precompute(1, cache_dir=d)                 # fwd request: recomputes, stores all
    fresh = _lib.precompute(64, 48, 5, 32.5, True, "ramp", True)
    got = precache.load(d, 64, 48, 5, 32.5, True, "ramp", True)
    for k in fresh:
        assert np.array_equal(got[k], fresh[k]), k
    assert got["glpars"][6] == 5
    lp2 = lpTransform(64, 48, 2, "ramp", 32.5, "cubic")
    lp2.precompute(1, cache_dir=d)
    assert np.array_equal(lp2.fwdparamsf, fresh["fwdparamsf"]) and lp2.glpars[6] == 2


def test_autograd_rejects_host_tensors():
    from lprec_amd.autograd import backprojection, radon
    with pytest.raises(ValueError):
        radon(torch.zeros(1, 4, 4), object())
    with pytest.raises(ValueError):
        backprojection(torch.zeros(1, 4, 4, dtype=torch.float64), object())
