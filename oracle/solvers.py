"""FP64 numpy restatements of the reference's iterative methods (TEST
INFRASTRUCTURE), on any operator pair given as functions -- the mirror's
``fwd``/``adj`` for parity of the device iterates, or dense matrices for the
pins in tests/test_methods.py.

* ``grad``  gradient descent on 1/2 ||R x - g||^2 with the per-slice
  Barzilai-Borwein step when reg_par < 0 (1e-3 on the first iteration),
  src/lprec/lpmethods.py:22-59.
* ``em``    x <- x * R*(g / (R x + eps)) / (R*(1) + 1e-5), eps = reg_par,
  src/lprec/lpmethods.py:100-128.
"""
import numpy as np


def grad(fwd, adj, init, tomo, num_iter, reg_par, keep=()):
    """-> final iterate, {i: iterate after i iterations for i in keep}."""
    x = np.array(init, dtype=np.float64)
    tomo = np.asarray(tomo, dtype=np.float64)
    ns = x.shape[0]
    x0 = x.copy()
    gr0 = np.zeros_like(x)
    kept = {}
    for i in range(num_iter):
        gr = adj(2 * (fwd(x) - tomo))
        if reg_par < 0:
            if i == 0:
                lam = np.full(ns, 1e-3)
            else:
                dg = gr - gr0
                lam = np.sum((x - x0) * dg, axis=(1, 2)) / np.sum(dg * dg, axis=(1, 2))
        else:
            lam = np.full(ns, float(reg_par))
        x0 = x.copy()
        gr0 = gr
        x = x - lam[:, None, None] * gr
        if i + 1 in keep:
            kept[i + 1] = x.copy()
    return x, kept


def em(fwd, adj, init, tomo, num_iter, reg_par, keep=()):
    """-> final iterate, {i: iterate after i iterations for i in keep}."""
    x = np.array(init, dtype=np.float64)
    tomo = np.asarray(tomo, dtype=np.float64)
    xi = adj(np.ones_like(tomo)) + 1e-5
    kept = {}
    for i in range(num_iter):
        x = x * (adj(tomo / (fwd(x) + reg_par)) / xi)
        if i + 1 in keep:
            kept[i + 1] = x.copy()
    return x, kept
