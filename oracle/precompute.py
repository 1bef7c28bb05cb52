"""FP64 numpy restatement of the reference precompute (TEST INFRASTRUCTURE).

Follows, function by function:

* ``create_gl``      src/lprec/initsgl.py:25-57  (+ getparameters :60-70,
                     osg :73-77, splineB3 :80-99)
* ``create_fwd``     src/lprec/initsfwd.py:15-66 (+ fzeta_loop_weights
                     :69-93, savePfwdpars :96-126)
* ``create_adj``     src/lprec/initsadj.py:19-116 (+ fzeta_loop_weights_adj
                     :119-143, take_filter :146-173, wint :176-203,
                     savePadjpars :205-244)

Precision: everything is FP64 except the two grids the reference explicitly
rounds to float32 (``thsp``/``rhosp``, initsgl.py:42-43), which define the
log-polar lattice.  The reference itself evaluates much of this in float32 and
complex64 (cupy dtype rules), so its tables differ from these by float32
rounding; membership tests (``<= 1``) can flip for points within ~1e-7 of a
boundary.  Outputs are the reference's flat wire arrays (params.cu:40-145).

Pins:
* ``splineB3``/``B3com``  -> closed form (4 + 2 cos(2 pi k / M)) / 6.
* ``fzeta``               -> direct FP64 quadrature of the same sum for a few
                             rows (tests/test_oracle_precompute.py).
* ``wint``                -> integrates polynomials of degree < 12 exactly.
* geometry                -> invariants: span partition of the sinogram,
                             rotation invariance of ``cids``, counts.
* end to end              -> the reference fixtures f/R at the reference
                             tolerances through ``oracle.mirror``.
"""
import math

import numpy as np

NSPAN = 3
FILTERS = ("ramp", "shepp-logan", "cosine", "cosine2", "hamming", "hann", "parzen")


class Geometry:
    """Global parameters (the reference's ``Pgl``, initsgl.py:3-23)."""


def padded_size(N0, cor):
    # initsgl.py:29 (float |N0/2 - cor|)
    return int(math.ceil((N0 + abs(N0 / 2 - cor) * 2.0) / 16.0) * 16)


def os_angles(N, Nproj):
    # initsgl.py:32 -- Python round() is round-half-to-even
    return int(max(round(3.0 * N / 2.0 / Nproj), 1))


def osg(aR, theta):
    # initsgl.py:73-77
    t = np.linspace(-np.pi / 2, np.pi / 2, 1000)
    w = aR * np.cos(t) + (1 - aR) + 1j * aR * np.sin(t)
    return np.max(np.log(np.abs(w)) + np.log(np.cos(theta - np.arctan2(w.imag, w.real))))


def splineB3(x2, r):
    # initsgl.py:80-99, evaluated in FP64 on the (float32) grid values
    x2 = np.asarray(x2, dtype=np.float64)
    sizex = len(x2)
    x2 = x2 - (x2[-1] + x2[0]) / 2
    stepx = x2[1] - x2[0]
    ri = int(math.ceil(2 * r))
    r = r * stepx
    c = int(math.ceil((sizex + 1) / 2.0))
    x2c = x2[c - 1]
    x = x2[c - ri - 1:c + ri]
    d = np.abs(x - x2c) / r
    B3 = np.zeros_like(x)
    for ix in range(-ri, ri + 1):
        i = ix + ri
        if d[i] < 1:
            B3[i] = (3 * d[i] ** 3 - 6 * d[i] ** 2 + 4) / 6
        elif d[i] < 2:
            B3[i] = (-d[i] ** 3 + 6 * d[i] ** 2 - 12 * d[i] + 8) / 6
    B3f = np.zeros_like(x2)
    B3f[c - ri - 1:c + ri] = B3
    return B3f


def create_gl(N0, Nproj, Nslices, cor, interp_type):
    """initsgl.py:25-57 -> (Geometry, glpars float32)."""
    P = Geometry()
    beta = np.pi / NSPAN
    N = padded_size(N0, cor)
    osangles = os_angles(N, Nproj)
    Nproj = osangles * Nproj
    proj = np.arange(0, Nproj) * np.pi / Nproj - beta / 2
    s = np.linspace(-1, 1, N)
    # getparameters, initsgl.py:60-70
    aR = np.sin(beta / 2) / (1 + np.sin(beta / 2))
    am = (np.cos(beta / 2) - np.sin(beta / 2)) / (1 + np.sin(beta / 2))
    g = osg(aR, beta / 2)
    Ntheta = N
    Nrho = 2 * N
    dtheta = (2 * beta) / Ntheta
    drho = (g - np.log(am)) / Nrho
    # initsgl.py:42-43: the log-polar lattice is rounded to float32
    thsp = (np.arange(-Ntheta / 2, Ntheta / 2) * np.float64(np.float32(dtheta))).astype(np.float32)
    rhosp = (np.arange(-Nrho, 0) * drho).astype(np.float32)
    erho = np.tile(np.exp(rhosp.astype(np.float64))[:, None], [1, Ntheta])
    B3th = np.fft.fft(np.fft.ifftshift(splineB3(thsp, 1)))
    B3rho = np.fft.fft(np.fft.ifftshift(splineB3(rhosp, 1)))
    B3com = np.outer(B3rho, B3th)
    P.Nspan, P.N, P.N0, P.Nproj, P.Nslices = NSPAN, N, N0, Nproj, Nslices
    P.Ntheta, P.Nrho, P.proj, P.s, P.thsp, P.rhosp = Ntheta, Nrho, proj, s, thsp, rhosp
    P.aR, P.beta, P.B3com, P.am, P.g, P.cor = aR, beta, B3com, am, g, cor
    P.osangles, P.interp_type = osangles, interp_type
    head = np.array([N, N0, Ntheta, Nrho, NSPAN, Nproj, Nslices, cor, osangles,
                     interp_type == "cubic"], dtype=np.float32)
    glpars = np.concatenate((head, erho.flatten().astype(np.float32)))
    return P, glpars


def _h_weights(Nl):
    # end-corrected trapezoid weights, initsfwd.py:74-83 (== initsadj.py:125-134)
    h = np.ones(Nl)
    correcting = 1 + np.array([-216254335, 679543284, -1412947389, 2415881496, -3103579086,
                               2939942400, -2023224114, 984515304, -321455811, 63253516,
                               -5675265]) / 958003200.0
    correcting[0] = 2 * (correcting[0] - 0.5)
    h[0] = h[0] * correcting[0]
    for j in range(1, len(correcting)):
        h[j] = h[j] * correcting[j]
        h[-j] = h[-j] * correcting[j]
    return h


def fzeta(Ntheta, Nrho, betas, rhos, a, osthlarge, adjoint, rows=None):
    """fzeta_loop_weights (initsfwd.py:69-93) / _adj (initsadj.py:119-143).

    Row j (kappa = j - Nrho/2) is the centred DFT over the 4x oversampled
    theta grid of h(th) * cos(th)**e with e = -2 pi i kappa/rhos - 1 - a (fwd)
    or e = 2 pi i kappa/rhos - a (adj); central Ntheta columns kept, scaled by
    the theta step, row 0 and column 0 zeroed.  Vectorised in row chunks.
    """
    krho = np.arange(-Nrho / 2, Nrho / 2)
    Nl = osthlarge * Ntheta
    thl = np.arange(-Nl / 2, Nl / 2) / Nl * betas
    h = _h_weights(Nl)
    logcos = np.log(np.cos(thl))
    lo, hi = Nl // 2 - Ntheta // 2, Nl // 2 + Ntheta // 2
    rows = range(Nrho) if rows is None else rows
    rows = np.asarray(list(rows))
    out = np.zeros((len(rows), Ntheta), dtype=np.complex128)
    chunk = max(1, (1 << 22) // Nl)
    for c0 in range(0, len(rows), chunk):
        rr = rows[c0:c0 + chunk]
        if adjoint:
            e = 2 * np.pi * 1j * krho[rr] / rhos - a
        else:
            e = -2 * np.pi * 1j * krho[rr] / rhos - 1 - a
        fc = np.exp(e[:, None] * logcos[None, :])
        F = np.fft.fftshift(np.fft.fft(np.fft.fftshift(h[None, :] * fc, axes=1), axis=1), axes=1)
        out[c0:c0 + len(rr)] = F[:, lo:hi]
    out *= thl[1] - thl[0]
    out[rows == 0] = 0
    out[:, 0] = 0
    return out


def create_fwd(P):
    """initsfwd.py:15-66 -> (parsi int32, parsf float32, parts dict)."""
    fZ = np.fft.fftshift(fzeta(P.Ntheta, P.Nrho, 2 * P.beta, P.g - np.log(P.am), 0, 4, False))
    th = P.thsp.astype(np.float64)
    rh = P.rhosp.astype(np.float64)
    tmp1 = np.outer(np.exp(rh), np.cos(th)).flatten()
    tmp2 = np.outer(np.exp(rh), np.sin(th)).flatten()
    lp2C1, lp2C2, cids_k = [], [], []
    for k in range(NSPAN):
        phi = k * P.beta + P.beta / 2
        c1 = ((tmp1 - (1 - P.aR)) * np.cos(phi) - tmp2 * np.sin(phi)) / P.aR
        c2 = ((tmp1 - (1 - P.aR)) * np.sin(phi) + tmp2 * np.cos(phi)) / P.aR
        c2 *= -1  # "adjust for Tomopy", initsfwd.py:29
        cids = np.where((c1 ** 2 + c2 ** 2) <= 1)[0]
        cids_k.append(cids)
        lp2C1.append(c1[cids])
        lp2C2.append(c2[cids])
    # the reference keeps only the last span's cids (initsfwd.py:30); the
    # membership is rotation invariant, so every span must agree
    for k in range(NSPAN - 1):
        if not np.array_equal(cids_k[k], cids_k[-1]):
            raise RuntimeError("C->LP disk membership differs between spans")
    cids = cids_k[-1]
    s0, th0 = np.meshgrid(P.s, P.proj)
    th0 = th0.flatten()
    s0 = s0.flatten()
    pids, p2lp1, p2lp2 = [], [], []
    for k in range(NSPAN):
        pk = np.where((th0 >= k * P.beta - P.beta / 2) & (th0 < k * P.beta + P.beta / 2))[0]
        pids.append(pk)
        th00 = th0[pk] - k * P.beta
        p2lp1.append(th00)
        p2lp2.append(np.log(s0[pk] * P.aR + (1 - P.aR) * np.cos(th00)))
    t0, t1 = np.float64(P.thsp[0]), np.float64(P.thsp[-1])
    r0, r1 = np.float64(P.rhosp[0]), np.float64(P.rhosp[-1])
    for k in range(NSPAN):
        lp2C1[k] = (lp2C1[k] + 1) / 2 * (P.N - 1)
        lp2C2[k] = (lp2C2[k] + 1) / 2 * (P.N - 1)
        p2lp1[k] = (p2lp1[k] - t0) / (t1 - t0) * (P.Ntheta - 1)
        p2lp2[k] = (p2lp2[k] - r0) / (r1 - r0) * (P.Nrho - 1)
    const = np.sqrt(P.N * P.osangles / P.Nproj) * np.pi / 4 / P.aR / np.sqrt(2)
    H = P.Ntheta // 2 + 1
    fZgpu = fZ[:, :H] * const
    if P.interp_type == "cubic":
        fZgpu = fZgpu / P.B3com[:, :H]
    # savePfwdpars, initsfwd.py:96-126
    fZvec = np.zeros(fZgpu.size * 2)
    fZvec[::2] = fZgpu.real.flatten()
    fZvec[1::2] = fZgpu.imag.flatten()
    parsi = np.concatenate([[len(p) for p in pids]] + pids + [[len(cids)], cids]).astype(np.int32)
    parsf = np.concatenate(lp2C1 + lp2C2 + p2lp1 + p2lp2 + [fZvec]).astype(np.float32)
    parts = dict(fZ=fZgpu, lp2C1=lp2C1, lp2C2=lp2C2, p2lp1=p2lp1, p2lp2=p2lp2, cids=cids, pids=pids)
    return parsi, parsf, parts


def wint_core(n, t):
    """initsadj.py:176-200 without the tail ramp: weights w such that
    sum(w * g(t)) = integral of t * g(t) over [t_0, t_end] for every
    polynomial g of degree < n (degree n-1 interpolation on overlapping
    windows, each sub-interval averaged over the windows covering it).
    The j-loop is vectorised (same sums, other order)."""
    t = np.asarray(t, dtype=np.float64)
    N = len(t)
    s = np.linspace(1e-40, 1, n)
    iv = np.linalg.inv(np.exp(np.outer(np.arange(n), np.log(s))))
    tmp2 = np.arange(1, n + 2)
    u = np.diff(np.exp(np.outer(tmp2, np.log(s))) * np.tile(1.0 / tmp2[:, None], [1, n]))
    W1 = iv @ u[1:n + 1, :]
    W2 = iv @ u[0:n, :]
    p = 1 / np.concatenate((np.arange(1, n), (n - 1) * np.ones(N - 2 * (n - 1) - 1),
                            np.arange(n - 1, 0, -1)))
    J = N - n + 1
    j = np.arange(J)
    a = t[j + n - 1] - t[j]
    # P[j, k] = p[j + k], k < n-1
    Pm = p[j[:, None] + np.arange(n - 1)[None, :]]
    C = (a ** 2)[:, None] * (Pm @ W1.T) + (a * t[j])[:, None] * (Pm @ W2.T)
    w = np.zeros(N)
    for i in range(n):
        w[i:i + J] += C[:, i]
    return w


def wint(n, t):
    """initsadj.py:176-203 (MATLAB twin tests/matlab_files/wint.m:1-20): the
    core weights with the last 40 replaced by a linear ramp (initsadj.py:202)."""
    w = wint_core(n, t)
    N = len(w)
    wn = w.copy()
    wn[-40:] = w[-40] / (N - 40) * np.arange(N - 40, N)
    return wn


def take_filter(N, filter_type):
    """initsadj.py:146-173: centred length-4N frequency response (float32)."""
    osf = 4
    d = 0.5
    Ne = osf * N
    t = np.arange(0, Ne / 2 + 1) / Ne
    w = Ne * 0.5 * wint(12, t)
    if filter_type == "ramp":
        wfa = w
    elif filter_type == "shepp-logan":
        wfa = w * np.sinc(t / (2 * d)) * (t / d <= 2)
    elif filter_type == "cosine":
        wfa = w * np.cos(np.pi * t / (2 * d)) * (t / d <= 1)
    elif filter_type == "cosine2":
        wfa = w * np.cos(np.pi * t / (2 * d)) ** 2 * (t / d <= 1)
    elif filter_type == "hamming":
        wfa = w * (0.54 + 0.46 * np.cos(np.pi * t / d)) * (t / d <= 1)
    elif filter_type == "hann":
        wfa = w * (1 + np.cos(np.pi * t / d)) / 2.0 * (t / d <= 1)
    elif filter_type == "parzen":
        wfa = w * (1 - t / d) ** 3 * (t / d <= 1)
    else:
        raise ValueError("unknown filter %r" % filter_type)
    wfa = wfa * (wfa >= 0)
    out = np.concatenate((np.flipud(wfa[1:]), [2 * wfa[0]], wfa[1:]))
    return out[:-1].astype(np.float32)


def normalize_filter(filter_type):
    """The reference compares with 'none' but its tests pass 'None'
    (initsadj.py:107 vs test_adj_many.py:13); names match case-insensitively."""
    if filter_type is None:
        return "none"
    f = str(filter_type).lower()
    if f != "none" and f not in FILTERS:
        raise ValueError("unknown filter %r" % filter_type)
    return f


def create_adj(P, filter_type):
    """initsadj.py:19-116 -> (parsi int32, parsf float32, parts dict)."""
    filter_type = normalize_filter(filter_type)
    fZ = np.fft.fftshift(fzeta(P.Ntheta, P.Nrho, 2 * P.beta, P.g - np.log(P.am), 0, 4, True))
    lin = np.linspace(-1, 1, P.N)
    x1, x2 = np.meshgrid(lin, lin)
    x1 = x1.flatten()
    x2 = x2.flatten() * (-1)  # "adjust for tomopy", initsadj.py:28
    cids = np.where(x1 ** 2 + x2 ** 2 <= 1)[0]
    C2lp1, C2lp2 = [], []
    for k in range(NSPAN):
        phi = k * P.beta + P.beta / 2
        z1 = P.aR * (x1[cids] * np.cos(phi) + x2[cids] * np.sin(phi)) + (1 - P.aR)
        z2 = P.aR * (-x1[cids] * np.sin(phi) + x2[cids] * np.cos(phi))
        C2lp1.append(np.arctan2(z2, z1))
        C2lp2.append(np.log(np.sqrt(z1 ** 2 + z2 ** 2)))
    z1, z2 = np.meshgrid(P.thsp.astype(np.float64), np.exp(P.rhosp.astype(np.float64)))
    z1 = z1.flatten()
    z2 = z2.flatten()
    hb = P.beta / 2
    # the reference's z1 is the float32 lattice compared with the Python float
    # beta/2, which numpy/cupy evaluate in float32 (initsadj.py:46, 57, 64);
    # an FP64 comparison would drop or add the first theta column of a span
    z1f = np.repeat(P.thsp[None, :], P.Nrho, axis=0).flatten()
    inspan = (z1f >= np.float32(-hb)) & (z1f < np.float32(hb))
    z2n = (z2 - (1 - P.aR) * np.cos(z1)) / P.aR
    lpids = np.where(inspan & (abs(z2n) <= 1))[0]
    lp2p1 = [z1[lpids] + k * P.beta for k in range(NSPAN)]
    lp2p2 = [z2n[lpids] for _ in range(NSPAN)]
    # rho-wrapped points, initsadj.py:52-72.  Right side: rows with rho > g,
    # sampled at rho + log(am) - g; left side: the first row(s), at rho - log(am) + g.
    lz = np.log(z2)
    wr = np.where(lz > P.g)[0]
    zr = (np.exp(lz[wr] + np.log(P.am) - P.g) - (1 - P.aR) * np.cos(z1[wr])) / P.aR
    kr = np.where(inspan[wr] & (abs(zr) <= 1))[0]
    drho = np.float64(P.rhosp[1]) - np.float64(P.rhosp[0])
    wl = np.where(lz < np.log(P.am) - P.g + drho)[0]
    zl = (np.exp(lz[wl] - np.log(P.am) + P.g) - (1 - P.aR) * np.cos(z1[wl])) / P.aR
    kl = np.where(inspan[wl] & (abs(zl) <= 1))[0]
    wids = np.concatenate((wr[kr], wl[kl]))
    # the reference indexes the full theta grid with kr/kl (initsadj.py:69);
    # both wrapped sets are whole rows starting at a row boundary, so that is
    # the theta of the wrapped point itself, which is what is used here
    lp2p1w = [z1[wids] + k * P.beta for k in range(NSPAN)]
    lp2p2w = [np.concatenate((zr[kr], zl[kl])) for _ in range(NSPAN)]
    pids = [np.where((P.proj >= k * P.beta - hb) & (P.proj < k * P.beta + hb))[0]
            for k in range(NSPAN)]
    proj0 = [P.proj[pids[k][0]] for k in range(NSPAN)]
    projl = [P.proj[pids[k][-1]] - P.proj[pids[k][0]] for k in range(NSPAN)]
    projp = (P.Nproj - 1) / (proj0[NSPAN - 1] + projl[NSPAN - 1] - proj0[0])
    t0, t1 = np.float64(P.thsp[0]), np.float64(P.thsp[-1])
    r0, r1 = np.float64(P.rhosp[0]), np.float64(P.rhosp[-1])
    for k in range(NSPAN):
        lp2p1[k] = (lp2p1[k] - proj0[k]) / projl[k] * (len(pids[k]) - 1) + (proj0[k] - proj0[0]) * projp
        lp2p1w[k] = (lp2p1w[k] - proj0[k]) / projl[k] * (len(pids[k]) - 1) + (proj0[k] - proj0[0]) * projp
        lp2p2[k] = (lp2p2[k] + 1) / 2 * (P.N - 1)
        lp2p2w[k] = (lp2p2w[k] + 1) / 2 * (P.N - 1)
        C2lp1[k] = (C2lp1[k] - t0) / (t1 - t0) * (P.Ntheta - 1)
        C2lp2[k] = (C2lp2[k] - r0) / (r1 - r0) * (P.Nrho - 1)
    const = (P.N + 1) * (P.N - 1) / P.N ** 2 / 2 * np.sqrt(P.osangles * P.Nproj / P.N / 2)
    H = P.Ntheta // 2 + 1
    fZgpu = fZ[:, :H] * const
    if P.interp_type == "cubic":
        fZgpu = fZgpu / P.B3com[:, :H]
    filt = take_filter(P.N, filter_type) if filter_type != "none" else None
    # savePadjpars, initsadj.py:205-244
    fZvec = np.zeros(fZgpu.size * 2)
    fZvec[::2] = fZgpu.real.flatten()
    fZvec[1::2] = fZgpu.imag.flatten()
    parsi = [[len(cids)], [len(lpids)], [len(wids)], cids, lpids, wids]
    parsf = C2lp1 + C2lp2 + lp2p1 + lp2p2 + lp2p1w + lp2p2w + [fZvec]
    if filt is not None:
        parsf.append(filt)
        parsi.append([1])
    else:
        parsi.append([0])
    parsi = np.concatenate(parsi).astype(np.int32)
    parsf = np.concatenate(parsf).astype(np.float32)
    parts = dict(fZ=fZgpu, C2lp1=C2lp1, C2lp2=C2lp2, lp2p1=lp2p1, lp2p2=lp2p2,
                 lp2p1w=lp2p1w, lp2p2w=lp2p2w, cids=cids, lpids=lpids, wids=wids, filter=filt)
    return parsi, parsf, parts


def precompute(N0, Nproj, Nslices, cor, interp_type="cubic", filter_type="none", fwd=True):
    """lpTransform.precompute (lpTransform.py:19-31) -> dict of wire arrays."""
    P, glpars = create_gl(N0, Nproj, Nslices, cor, interp_type)
    out = dict(geom=P, glpars=glpars)
    if fwd:
        out["fwdparsi"], out["fwdparamsf"], out["fwd"] = create_fwd(P)
    out["adjparsi"], out["adjparamsf"], out["adj"] = create_adj(P, filter_type)
    return out
