"""Test inputs built independently of the library: the reference's
deterministic initializer restated in numpy (gppComplex.cpp:296-329),
numpy-seeded inputs with the benchmark distributions, and crafted
near-pole inputs that drive every occupied-band branch of
ssxt_scht_solver (gppComplex.cpp:35-56)."""
from __future__ import annotations

import numpy as np


def det_inputs(nb, nv, ncouls, npg=None, ngpown=None, nfreq=3, igp_zero=False):
    """gppComplex.cpp:296-329; ngpown = ncouls // nodes_per_group (:238)."""
    if ngpown is None:
        ngpown = ncouls // npg
    i = np.arange(nb, dtype=np.float64)[:, None]
    g = np.arange(ncouls, dtype=np.float64)[None, :]
    aq = 0.5 * (i + g) * (1 + 1j)
    k = np.arange(ngpown, dtype=np.float64)[:, None]
    eps = 0.5 * (k + g) * (1 + 1j)
    jj = np.arange(ngpown, dtype=np.int64)
    idx = ((jj if igp_zero else jj + 1) * ncouls // ngpown).astype(np.int32)
    wx = np.array([10.0 - 6.0 + 1.0 * ((iw + 1) - 2) for iw in range(nfreq)])
    wx = np.where(wx < 1e-6, 1e-6, wx)
    return {
        "inv_igp_index": idx, "wx_array": wx, "vcoul": np.arange(ncouls, dtype=np.float64),
        "aqsmtemp": aq.copy(), "aqsntemp": aq.copy(), "I_eps_array": eps.copy(), "wtilde_array": eps.copy(),
    }


def seeded_inputs(nb, ncouls, ngpown, nfreq=3, seed=0, wx_center=4.0, dw=1.0, sorted_index=True):
    """The benchmark value distributions, drawn with numpy's PCG64."""
    rng = np.random.default_rng(seed)

    def cu(shape, lo, hi, ilo, ihi):
        return rng.uniform(lo, hi, shape) + 1j * rng.uniform(ilo, ihi, shape)

    idx = rng.choice(ncouls, size=ngpown, replace=False)
    if sorted_index:
        idx = np.sort(idx)
    wx = wx_center + dw * (np.arange(nfreq) - 0.5 * (nfreq - 1))
    return {
        "inv_igp_index": idx.astype(np.int32),
        "wx_array": np.maximum(wx, 1e-6),
        "vcoul": rng.uniform(0.1, 10.0, ncouls),
        "aqsmtemp": cu((nb, ncouls), -1, 1, -1, 1),
        "aqsntemp": cu((nb, ncouls), -1, 1, -1, 1),
        "I_eps_array": cu((ngpown, ncouls), -0.5, 0.5, -0.5, 0.5),
        "wtilde_array": cu((ngpown, ncouls), 0.5, 5.0, -0.05, 0.05),
    }


def near_pole_inputs(nb, ncouls, ngpown, seed=0):
    """Seeded inputs with wx = (399.4, 4.0, 0.1) and wtilde placed on purpose
    relative to those poles so that the occupied-band solver takes every
    branch of gppComplex.cpp:35-56:
      kind 0  default seeded wtilde                     -> branch 1
      kind 1  4 + U(-0.45,0.45) + i U(-0.01,0.01)       -> |wdiff|^2 <= 0.25 at wx=4: branch 2
      kind 2  4 + U(-1e-4,1e-4) + i U(-1e-5,1e-5)       -> extreme near-pole factors at wx=4
      kind 3  U(-1e-5,1e-5) + i U(-1e-5,1e-5)           -> |delw|^2 <= 1e-6 at wx=0.1: branch 3
      kind 4  400 + i U(-0.01,0.01)                     -> |delw|^2 >= 2.5e5 at wx=399.4: branch 2
    """
    a = seeded_inputs(nb, ncouls, ngpown, 3, seed)
    rng = np.random.default_rng(seed + 1000)
    shape = a["wtilde_array"].shape
    kind = rng.integers(0, 5, shape)
    wt = a["wtilde_array"]
    u = lambda lo, hi: rng.uniform(lo, hi, shape)
    wt = np.where(kind == 1, 4.0 + u(-0.45, 0.45) + 1j * u(-0.01, 0.01), wt)
    wt = np.where(kind == 2, 4.0 + u(-1e-4, 1e-4) + 1j * u(-1e-5, 1e-5), wt)
    wt = np.where(kind == 3, u(-1e-5, 1e-5) + 1j * u(-1e-5, 1e-5), wt)
    wt = np.where(kind == 4, 400.0 + 1j * u(-0.01, 0.01), wt)
    a["wtilde_array"] = wt
    a["wx_array"] = np.array([399.4, 4.0, 0.1])
    return a


def ref_delwr(wx, wt):
    """|delw|^2 and |wdiff|^2 of the occupied-band solver in the reference's
    operation order (gppComplex.cpp:27-33: wdiff = wx - wt, rden = 1/|wdiff|^2,
    delw = wt conj(wdiff) rden), one IEEE operation at a time."""
    wtr, wti = np.real(wt), np.imag(wt)
    wdr, wdi = wx - wtr, 0.0 - wti
    cre, cim = wdr, -1.0 * wdi
    wdiffr = wdr * cre - wdi * cim
    rden = 1.0 / wdiffr
    dre = (wtr * cre - wti * cim) * rden
    dim_ = (wtr * cim + wti * cre) * rden
    return dre * dre - dim_ * (-1.0 * dim_), wdiffr


def threshold_inputs(nb, ncouls, ngpown, seed=0, spread=12):
    """Seeded inputs with wx = (399.4, 4.0, 0.1) and wtilde placed so that the
    reference's |delw|^2 (gppComplex.cpp:32) lies within a few ulps of a
    branch threshold for one of the frequencies:
      kind 0  |delw|^2 ~ limitone = 2.5e5 at wx = 399.4, |wdiff|^2 ~ 0.64 > 0.25:
              branch 1 (:35) or branch 2 (:43) by the last bits
      kind 1  |delw|^2 ~ to1 = 1e-6 at wx = 0.1, |wdiff|^2 ~ 0.01 <= 0.25:
              branch 2 (:43) or neither by the last bits
      kind 2  default seeded wtilde
    Each threshold element's modulus is scaled by (1 + k 2^-52), k uniform in
    [-spread, spread], so the decisions fall on both sides."""
    a = seeded_inputs(nb, ncouls, ngpown, 3, seed)
    rng = np.random.default_rng(seed + 2000)
    shape = a["wtilde_array"].shape
    kind = rng.integers(0, 3, shape)
    theta = rng.uniform(-np.pi, np.pi, shape)
    jitter = 1.0 + rng.integers(-spread, spread + 1, shape) * 2.0 ** -52

    def at_ratio(wx, t, near_wx):
        # wt with |wt|^2 / |wx - wt|^2 = t; near_wx: wt = wx + d e^{i theta}
        # (pole side), else wt = r e^{i theta} (small wt)
        c = np.cos(theta)
        if near_wx:
            # |wx + d e^{i th}|^2 = t d^2  ->  (t - 1) d^2 - 2 wx c d - wx^2 = 0
            dd = (2 * wx * c + np.sqrt(4 * wx * wx * c * c + 4 * (t - 1) * wx * wx)) / (2 * (t - 1))
            return wx + dd * np.exp(1j * theta)
        # r^2 = t |wx - r e^{i th}|^2  ->  (1 - t) r^2 + 2 t wx c r - t wx^2 = 0
        r = (-2 * t * wx * c + np.sqrt(4 * t * t * wx * wx * c * c + 4 * (1 - t) * t * wx * wx)) / (2 * (1 - t))
        return r * np.exp(1j * theta)

    wt = a["wtilde_array"]
    wt = np.where(kind == 0, at_ratio(399.4, 2.5e5, True) * jitter, wt)
    wt = np.where(kind == 1, at_ratio(0.1, 1e-6, False) * jitter, wt)
    a["wtilde_array"] = wt
    a["wx_array"] = np.array([399.4, 4.0, 0.1])
    a["_kind"] = kind
    return a


def per_band_wx(nb, nfreq, seed=0, levels=None, e_lk=10.0, dw=1.0):
    """Band-dependent frequencies wx[n1][iw] = e_lk - e_n1kq[n1] + dw*(iw-1),
    clamped at 1e-6 as gppComplex.cpp:326-328 clamps its one row; BerkeleyGW
    gives every band its own e_n1kq, the mini-app fixes e_n1kq = 6.  With
    ``levels`` the band energies are drawn from that many distinct values
    (degenerate bands share a row), else every band gets its own."""
    rng = np.random.default_rng(seed + 500)
    if levels is None:
        e = rng.uniform(2.0, 9.0, nb)
    else:
        e = rng.choice(rng.uniform(2.0, 9.0, levels), size=nb)
    wx = e_lk - e[:, None] + dw * (np.arange(nfreq)[None, :] - 1.0)
    return np.maximum(wx, 1e-6)


def ref_abs_ssx(wx, wt, ie, branch):
    """|ssx| before the cutoff and the cutoff 4|I_eps| of the occupied-band
    solver for branch 1 (gppComplex.cpp:38-41) or 2 (:46-50), one IEEE
    operation at a time in the reference's order (Complex.h:141-250)."""
    wr, wi = np.real(wt), np.imag(wt)
    er, ei = np.real(ie), np.imag(ie)
    w2r, w2i = wr * wr - wi * wi, 2 * wr * wi                     # square, :117
    o2r, o2i = w2r * er - w2i * ei, w2r * ei + w2i * er           # Omega2, :118
    if branch == 1:
        cdend = wx * wx                                           # pow(wx, 2), :38
        rden = 1.00 / (cdend * cdend)                             # :39-40
        sr, si = o2r * cdend * rden, o2i * cdend * rden           # mult(a, b, c), :41
    else:
        cr, ci = wx - wr, 0 - wi                                  # wdiff = cden, :27-29
        rden = 1 / (cr * cr - ci * (-1 * ci))                     # :30
        dr = (wr * cr - wi * (-1 * ci)) * rden                    # delw, :31
        di = (wr * (-1 * ci) + wi * cr) * rden
        pr, pi_ = 0.50 + dr, 0 + di                               # doublePlus(0.5, delw)
        c2r, c2i = (w2r * pr - w2i * pi_) * 4.00, (w2r * pi_ + w2i * pr) * 4.00   # :46
        rd = 1.00 / (c2r * c2r - c2i * (-1 * c2i))                # :48-49
        xr = (-o2r) * c2r - (-o2i) * (-1 * c2i)                   # -Omega2 * conj(cden)
        xi = (-o2r) * (-1 * c2i) + (-o2i) * c2r
        yr, yi = dr * rd, di * rd                                 # delw * rden
        sr, si = xr * yr - xi * yi, xr * yi + xi * yr             # :50
    return np.sqrt(sr * sr + si * si), np.sqrt(er * er + ei * ei) * 4.0


def ssx_cutoff_inputs(nb, ncouls, ngpown, seed=0, spread=12):
    """Seeded inputs with wx = (-0.6, -0.2, 3.0) and wtilde placed so that the
    occupied solver's |ssx| lies within a few ulps of the cutoff 4|I_eps| of
    gppComplex.cpp:58-59 (live only at negative frequencies):
      kind 0  branch 1 at wx = -0.6 with |wt|^2 / wx^2 ~ 4 (|wt| ~ 1.2)
      kind 1  branch 2 at wx = -0.2 with |wt| / |wx + wt| ~ 8
      kind 2  default seeded wtilde
    Each placed modulus is scaled by (1 + k 2^-52), k uniform in
    [-spread, spread], so the cutoff falls on both sides."""
    a = seeded_inputs(nb, ncouls, ngpown, 3, seed)
    rng = np.random.default_rng(seed + 3000)
    shape = a["wtilde_array"].shape
    kind = rng.integers(0, 3, shape)
    jitter = 1.0 + rng.integers(-spread, spread + 1, shape) * 2.0 ** -52
    th0 = rng.uniform(-2.5, 2.5, shape)                 # |-0.6 - wt|^2 = 1.8 + 1.44 cos(th) > 0.25
    wt0 = 1.2 * np.exp(1j * th0) * jitter
    wx1, t = -0.2, 8.0
    th1 = rng.uniform(0.02, 0.11, shape) * rng.choice([-1.0, 1.0], shape)
    aa, bb, cc = 1 - 1 / t ** 2, 2 * wx1 * np.cos(th1), wx1 * wx1
    r1 = (-bb + np.sqrt(bb * bb - 4 * aa * cc)) / (2 * aa)   # |wx + r e^{ith}| = r / t
    wt1 = r1 * np.exp(1j * th1) * jitter
    wt = a["wtilde_array"]
    wt = np.where(kind == 0, wt0, wt)
    wt = np.where(kind == 1, wt1, wt)
    a["wtilde_array"] = wt
    a["wx_array"] = np.array([-0.6, -0.2, 3.0])
    a["_kind"] = kind
    return a
