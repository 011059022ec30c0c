"""Pins the CPU oracle (oracle/gpp_oracle.c) before anything is checked
against it:

* the README golden achtemp triples (the reference's only published
  vectors, README.md:25-42, inputs testSmallJob/testMedJob/testBigJob.pbs:17);
* the closed-form integer asxtemp at the deterministic init;
* invariants of the deterministic init;
* exact-rational brute force on tiny shapes (seeded and near-pole too);
* hand-derived closed forms for each occupied-band branch, checked against
  numpy complex128;
* bitwise agreement with the compiled reference (oracle/_ref) where it was
  built, and with the fixtures the reference produced (tests/golden/).
"""
import os

import numpy as np
import pytest

import exact_gpp
import gpp_inputs

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# README.md:27-41 (achtemp[0..2], 15 significant digits)
README_GOLDEN = {
    800: [(-2.33321481494884e+17, -2.30534726978554e+17),
          (-2.34252423351125e+17, -2.30513076345906e+17),
          (-2.35188613227291e+17, -2.30484574707011e+17)],
    1600: [(-1.00387366710739e+19, -9.9815856421254e+18),
           (-1.00578107032611e+19, -9.98139840041474e+18),
           (-1.00769357238471e+19, -9.98115517772354e+18)],
    3200: [(-5.13842445834172e+20, -5.12423661308937e+20),
           (-5.14315659139572e+20, -5.12421580271217e+20),
           (-5.14789457600568e+20, -5.12418893702435e+20)],
}


def rel_err(a, b):
    """Largest per-component relative error of a against b."""
    a = np.asarray(a, np.complex128).ravel()
    b = np.asarray(b, np.complex128).ravel()
    ar, ai, br, bi = a.real, a.imag, b.real, b.imag
    floor = 1e-300 + 1e-30 * np.max(np.abs(np.concatenate([br, bi])))
    return float(max(np.max(np.abs(ar - br) / np.maximum(np.abs(br), floor)),
                     np.max(np.abs(ai - bi) / np.maximum(np.abs(bi), floor))))


@pytest.mark.parametrize("ncouls", [800, 1600, 3200])
def test_readme_golden_triples(oracle, ncouls):
    arrs = gpp_inputs.det_inputs(512, 2, ncouls, npg=20)
    out = oracle.sigma(arrs, nvband=2, threads=8)
    gold = np.array([complex(r, i) for r, i in README_GOLDEN[ncouls]])
    # 15 printed significant digits -> half an ulp of the 15th digit
    assert rel_err(out["achtemp"], gold) < 6e-15


def _asx_closed_form(nb, nv, ncouls, ngpown, wx, occ=1.0):
    """asxtemp[iw] = occ(-1+i)/(8 wx^2) * sum_{n1<nv,j} igp(n1+igp) sum_ig (n1+ig)(j+ig)^3
    at the deterministic init (every occupied term is in branch 1 there)."""
    ig = np.arange(ncouls, dtype=object)
    total = 0
    for n1 in range(nv):
        for j in range(ngpown):
            igp = min((j + 1) * ncouls // ngpown, ncouls - 1)
            s = int(sum((n1 + ig) * (j + ig) ** 3))
            total += igp * (n1 + igp) * s
    return np.array([occ * complex(-total, total) / (8 * w * w) for w in wx])


@pytest.mark.parametrize("shape", [(4, 2, 40, 5), (8, 3, 64, 4), (512, 2, 800, 20)])
def test_asxtemp_closed_form(oracle, shape):
    nb, nv, ncouls, npg = shape
    arrs = gpp_inputs.det_inputs(nb, nv, ncouls, npg=npg)
    out = oracle.sigma(arrs, nvband=nv, threads=8)
    ref = _asx_closed_form(nb, nv, ncouls, ncouls // npg, arrs["wx_array"])
    assert rel_err(out["asxtemp"], ref) < 1e-14


def test_det_init_invariants(oracle):
    arrs = gpp_inputs.det_inputs(64, 2, 256, npg=8)
    out = oracle.sigma(arrs, nvband=2, threads=8)
    wx = arrs["wx_array"]
    asx = out["asxtemp"]
    # re(asx) = -im(asx), asx*wx^2 independent of iw
    np.testing.assert_allclose(asx.real, -asx.imag, rtol=1e-14)
    np.testing.assert_allclose(asx * wx ** 2, asx[0] * wx[0] ** 2, rtol=1e-14)
    # sum of acht_n1_loc = achtemp[2]
    np.testing.assert_allclose(out["acht_n1_loc"].sum(), out["achtemp"][2], rtol=1e-12)
    # achstemp has re = im
    assert abs(out["achstemp"].real - out["achstemp"].imag) <= 1e-14 * abs(out["achstemp"].real)
    # every occupied term is in branch 1 at this init: achtemp does not depend on nvband
    out0 = oracle.sigma(arrs, nvband=0, threads=8)
    np.testing.assert_array_equal(out0["achtemp"], out["achtemp"])


def _check_exact(oracle, arrs, nv, tol=1e-13):
    ex = exact_gpp.sigma_exact(arrs, nv)
    out = oracle.sigma(arrs, nvband=nv)
    for k in ("achtemp", "asxtemp"):
        for iw in range(len(arrs["wx_array"])):
            e = exact_gpp.to_complex(ex[k][iw])
            scale = ex[k + "_mag"][iw] + 1e-300
            assert abs(out[k][iw] - e) <= tol * scale, (k, iw, out[k][iw], e)
    e = exact_gpp.to_complex(ex["achstemp"])
    assert abs(out["achstemp"] - e) <= tol * (ex["achstemp_mag"] + 1e-300)
    for n1 in range(arrs["aqsmtemp"].shape[0]):
        e = exact_gpp.to_complex(ex["acht_n1_loc"][n1])
        assert abs(out["acht_n1_loc"][n1] - e) <= tol * (ex["achtemp_mag"][-1] + ex["achtemp_mag"][min(2, len(arrs["wx_array"]) - 1)])


def test_threshold_inputs_straddle_thresholds():
    """The threshold inputs the factored-form tests use put the reference's
    |delw|^2 within 2^-40 of limitone and of to1 for thousands of elements,
    on both sides of each threshold, and keep |wdiff|^2 on the side of 0.25
    that makes the |delw|^2 test decide (gppComplex.cpp:35, :43)."""
    a = gpp_inputs.threshold_inputs(24, 2000, 60, seed=5)
    kind, wt = a["_kind"], a["wtilde_array"]
    for k, wx, t, above_quarter in ((0, 399.4, 2.5e5, True), (1, 0.1, 1e-6, False)):
        delwr, wdiffr = gpp_inputs.ref_delwr(wx, wt[kind == k])
        band = np.abs(delwr / t - 1) < 2.0 ** -40
        assert (delwr[band] < t).sum() > 1000 and (delwr[band] > t).sum() > 1000
        assert np.all((wdiffr > 0.25) == above_quarter)


def test_ssx_cutoff_inputs_straddle_cutoff(oracle):
    """The negative-frequency inputs the factored-form tests use put the
    reference's |ssx| within 2^-40 of the cutoff 4|I_eps| (gppComplex.cpp:
    58-59) for thousands of elements on each side, in branch 1 (wx = -0.6)
    and branch 2 (wx = -0.2); the numpy restatement of the reference's ssx
    agrees with the oracle's solver, and a branch-1 ssx rounded as
    Omega2 * (wx^2 * (1/wx^4)) instead of (Omega2 * wx^2) * (1/wx^4) decides
    hundreds of them the other way."""
    a = gpp_inputs.ssx_cutoff_inputs(24, 2000, 60, seed=5)
    kind, wt, ie = a["_kind"], a["wtilde_array"], a["I_eps_array"]
    for k, wx, br in ((0, -0.6, 1), (1, -0.2, 2)):
        s, cut = gpp_inputs.ref_abs_ssx(wx, wt[kind == k], ie[kind == k], br)
        band = np.abs(s / cut - 1) < 2.0 ** -40
        assert (s[band] > cut[band]).sum() > 5000 and (s[band] <= cut[band]).sum() > 5000
        delwr, wdiffr = gpp_inputs.ref_delwr(wx, wt[kind == k])
        if br == 1:
            assert np.all((wdiffr > 0.25) & (delwr < 2.5e5))
        else:
            assert np.all(((wdiffr <= 0.25) | (delwr >= 2.5e5)) & (delwr > 1e-6))
        for i in np.flatnonzero((kind == k).ravel())[:200]:
            w, e = wt.ravel()[i], ie.ravel()[i]
            got_br, _, ssx = oracle.pole_terms(wx, w, e)
            s1, c1 = gpp_inputs.ref_abs_ssx(wx, np.array([w]), np.array([e]), br)
            assert got_br == br and (ssx == 0) == (s1[0] > c1[0])
        if br == 1:
            wr, wi, er, ei = wt[kind == k].real, wt[kind == k].imag, ie[kind == k].real, ie[kind == k].imag
            w2r, w2i = wr * wr - wi * wi, 2 * wr * wi
            o2r, o2i = w2r * er - w2i * ei, w2r * ei + w2i * er
            c = (wx * wx) * (1.00 / ((wx * wx) * (wx * wx)))
            alt = np.sqrt((o2r * c) ** 2 + (o2i * c) ** 2)
            assert ((alt > cut) != (s > cut)).sum() > 100


def test_exact_rational_det(oracle):
    _check_exact(oracle, gpp_inputs.det_inputs(4, 2, 6, ngpown=3), 2)


@pytest.mark.parametrize("seed", [1, 2])
def test_exact_rational_seeded(oracle, seed):
    _check_exact(oracle, gpp_inputs.seeded_inputs(4, 7, 3, nfreq=3, seed=seed), 2)


def test_exact_rational_near_pole(oracle):
    _check_exact(oracle, gpp_inputs.near_pole_inputs(3, 9, 4, seed=5), 2)


def test_exact_rational_nfreq_and_clamp(oracle):
    a = gpp_inputs.seeded_inputs(3, 5, 3, nfreq=5, seed=9)
    a["inv_igp_index"] = np.array([4, 7, 100], np.int32)  # igp clamps to ncouls-1
    _check_exact(oracle, a, 1)


# ---- occupied-band branches: closed forms, textbook complex arithmetic -----

def _np_pole(wx, wt, ie):
    """ssxt_scht_solver by textbook formulas (numpy complex128)."""
    wdiff = wx - wt
    delw = wt / wdiff
    om2 = wt ** 2 * ie
    if abs(wdiff) ** 2 > 0.25 and abs(delw) ** 2 < 1.0 / (1e-6 * 4.0):
        sch, ssx, br = delw * ie, om2 / wx ** 2, 1
    elif abs(delw) ** 2 > 1e-6:
        cden = 4 * wt ** 2 * (0.5 + delw)
        sch, ssx, br = 0j, -om2 * delw / cden, 2
    else:
        sch, ssx, br = 0j, 0j, 3
    if wx < 0 and abs(ssx) > 4 * abs(ie):
        ssx = 0j
    return br, sch, ssx


POLE_CASES = [
    # (wx, wtilde, I_eps, expected branch)
    (4.0, 1.0 + 0.1j, 0.3 - 0.2j, 1),         # far from the pole
    (4.0, 3.9 + 0.01j, 0.3 - 0.2j, 2),        # |wdiff|^2 = 0.0101 <= 0.25
    (399.4, 400.0 + 0.003j, -0.1 + 0.4j, 2),  # |wdiff|^2 > 0.25 but |delw|^2 ~ 4.4e5 >= 2.5e5
    (0.1, 1e-5 - 2e-6j, 0.2 + 0.2j, 3),       # |wdiff|^2 <= 0.25 and |delw|^2 <= 1e-6
    (4.0, 0j, 0.5 + 0.5j, 1),                 # wtilde = 0: delw = 0
    (-1.0, 0.2 + 0j, 1e-3 + 0j, 1),           # wx < 0, |ssx| below the cutoff: kept
    (-1.0, 3.0 + 0j, 1e-3 + 1e-3j, 1),        # wx < 0, |ssx| = 9|I_eps| > 4|I_eps|: zeroed
]


@pytest.mark.parametrize("wx,wt,ie,branch", POLE_CASES)
def test_pole_branches_closed_form(oracle, wx, wt, ie, branch):
    br, sch, ssx = oracle.pole_terms(wx, wt, ie)
    nbr, nsch, nssx = _np_pole(wx, wt, ie)
    ebr, esch, essx = exact_gpp.pole_terms_exact(exact_gpp.Fr(wx), exact_gpp._c(wt), exact_gpp._c(ie))
    assert br == branch == nbr == ebr
    for got, want, ex in ((sch, nsch, esch), (ssx, nssx, essx)):
        ex = exact_gpp.to_complex(ex)
        assert abs(got - want) <= 1e-14 * max(abs(want), 1e-300)
        assert abs(got - ex) <= 4e-16 * 8 * max(abs(ex), 1e-300)
    if wx < 0 and branch == 1 and abs(wt) > 1:
        assert ssx == 0


def test_pole_exactly_at_pole(oracle):
    # wtilde == wx: rden = 1/0 = inf and delw = 0*inf = NaN, so both branch
    # tests are false and the occupied solver falls through to branch 3
    # (sch = ssx = 0); the unoccupied term a*(delw*I_eps) is NaN.
    br, sch, ssx = oracle.pole_terms(4.0, 4.0 + 0j, 1 + 1j)
    assert br == 3 and sch == 0 and ssx == 0
    if oracle.ref_available():
        assert oracle.pole_terms(4.0, 4.0 + 0j, 1 + 1j, use_ref=True)[1:] == (sch, ssx)
    a = gpp_inputs.seeded_inputs(2, 8, 2, seed=1)
    a["wtilde_array"][1, 3] = a["wx_array"][1]
    out = oracle.sigma(a, nvband=1)
    assert np.isnan(out["achtemp"][1].real) and not np.isnan(out["achtemp"][0].real)


def test_oracle_rejects_negative_index(oracle):
    a = gpp_inputs.seeded_inputs(2, 8, 2, seed=1)
    a["inv_igp_index"] = np.array([1, -3], np.int32)
    with pytest.raises(ValueError):
        oracle.sigma(a, nvband=1)


def test_openmp_is_bitwise_sequential(oracle):
    a = gpp_inputs.near_pole_inputs(6, 300, 20, seed=3)
    s = oracle.sigma(a, nvband=2, threads=1)
    p = oracle.sigma(a, nvband=2, threads=8)
    for k in ("achtemp", "asxtemp", "acht_n1_loc"):
        np.testing.assert_array_equal(s[k], p[k])
    assert s["achstemp"] == p["achstemp"]


def test_compensated_oracle_close(oracle):
    a = gpp_inputs.seeded_inputs(8, 500, 25, seed=4)
    s = oracle.sigma(a, nvband=2, threads=8)
    d = oracle.sigma_dd(a, nvband=2, threads=8)
    for k in ("achtemp", "asxtemp"):
        np.testing.assert_allclose(s[k], d[k], rtol=1e-11)


# ---- the compensated (double-double) oracle ---------------------------------
#
# oracle_sigma_dd sums the reference's rounded terms in double-double, so it
# is the reference's arithmetic without the summation error of its
# left-to-right sums (gppComplex.cpp:125-126, :146, :363, :379).  Its helpers
# are checked against exact rational arithmetic, and the whole loop against
# the exact-rational brute force on inputs where the sequential sum is far
# off, so a TwoSum that loses its low part fails.

def _fr_sum(xs):
    return sum((exact_gpp.Fr(float(x)) for x in xs), exact_gpp.Fr(0))


@pytest.mark.parametrize("xs,exact", [
    ([2.0 ** 60, 1.0, -2.0 ** 60], 1.0),                 # the 1 is below the ulp of 2^60
    ([1.0] + [2.0 ** -60] * 1000, 1.0 + 1000 * 2.0 ** -60),
    ([2.0 ** 80, 3.0, -(2.0 ** 80), -1.5, 2.0 ** -40], 1.5 + 2.0 ** -40),
])
def test_dd_sum_exact_known(oracle, xs, exact):
    hi, lo = oracle.dd_sum(xs)
    assert exact_gpp.Fr(hi) + exact_gpp.Fr(lo) == _fr_sum(xs)
    assert hi + lo == exact
    # the plain left-to-right double sum loses the small part
    plain = 0.0
    for x in xs:
        plain += x
    assert plain != hi + lo


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_dd_sum_random_vs_exact(oracle, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(4000) * 2.0 ** rng.integers(-30, 30, 4000)
    x = np.concatenate([x, -x[::-1] * (1 + 2.0 ** -30)])  # heavy cancellation
    hi, lo = oracle.dd_sum(x)
    ex = _fr_sum(x)
    scale = float(np.abs(x).sum())
    err = abs(float(exact_gpp.Fr(hi) + exact_gpp.Fr(lo) - ex))
    plain = float(np.cumsum(x)[-1])
    assert err <= 2.0 ** -100 * len(x) * scale
    assert abs(plain - float(ex)) > 1e6 * max(err, 2.0 ** -200 * scale)


def test_dd_sum_dd_and_mul_d_vs_exact(oracle):
    rng = np.random.default_rng(7)
    F = exact_gpp.Fr
    pairs = []
    for _ in range(500):
        h = float(rng.standard_normal() * 2.0 ** int(rng.integers(-20, 40)))
        l = float(h * 2.0 ** -53 * rng.uniform(-0.5, 0.5))  # |lo| <= ulp(hi)/2
        pairs.append((h, l))
    pairs += [(-h, -l * (1 - 2.0 ** -20)) for h, l in pairs[::-1]]
    hi, lo = oracle.dd_sum_dd(pairs)
    ex = sum((F(h) + F(l) for h, l in pairs), F(0))
    scale = sum(abs(h) for h, _ in pairs)
    assert abs(float(F(hi) + F(lo) - ex)) <= 2.0 ** -100 * len(pairs) * scale
    # dropping every pair's lo part is wrong by far more than that
    assert abs(float(sum((F(h) for h, _ in pairs), F(0)) - ex)) > 1e3 * 2.0 ** -100 * len(pairs) * scale
    for h, l in pairs[:200]:
        for b in (0.5 * 3.7, 1.5, float(rng.uniform(0.05, 5.0))):
            rh, rl = oracle.dd_mul_d(h, l, b)
            want = (F(h) + F(l)) * F(b)
            assert abs(float(F(rh) + F(rl) - want)) <= 2.0 ** -104 * abs(h * b)
    # the low part is carried: (2^60, 1) * 1.5 = 1.5 * 2^60 + 1.5 exactly
    rh, rl = oracle.dd_mul_d(2.0 ** 60, 1.0, 1.5)
    assert F(rh) + F(rl) == F(2) ** 60 * F(3, 2) + F(3, 2)


def _dyadic_cancelling_inputs():
    """Inputs whose every term the reference computes exactly (dyadic values,
    wtilde = 1 and wx = 2 or 3 so |wdiff|^2 = 1 or 4 and delw = 1 or 1/2;
    at wx = 2 branch 1 gives ssx = I_eps/4 exactly), with terms of 2^60 that
    cancel inside a column's G sum (band 0) or across bands (1 and 2), so
    the reference's left-to-right sums lose the small terms entirely."""
    nb, ncouls, ngpown = 3, 8, 2
    big = 2.0 ** 60
    an = np.zeros((nb, ncouls), np.complex128)
    an[0] = [big * (1 + 1j), 1 + 2j, -big * (1 + 1j), 3 - 1j, 2.0 ** -30, 0, 0.25j, -0.5]
    # bands 1 and 2 leave +-2^60 in every column's G sum; they cancel only
    # across bands, after the scaling by 0.5*vcoul (double-double times double)
    an[1] = [big, 0.75 - 1j, big * 1j, 5.0, -big * 1j, 2.0 ** -20, 1j, 0]
    an[2] = [-big, 1.5, 0.5 - 0.5j, 7j, -2.0, 2.0 ** -10, 1, 0]
    return {
        "inv_igp_index": np.array([0, 1], np.int32),
        "wx_array": np.array([2.0, 3.0]),
        "vcoul": np.array([3.0, 2.0] + [1.0] * (ncouls - 2)),   # 0.5*vcoul = 1.5 and 1
        "aqsmtemp": np.ones((nb, ncouls), np.complex128),
        "aqsntemp": an,
        "I_eps_array": np.ones((ngpown, ncouls), np.complex128) * (1 + 0j),
        "wtilde_array": np.ones((ngpown, ncouls), np.complex128),
    }


def test_compensated_oracle_exact_on_dyadic_cancellation(oracle):
    a = _dyadic_cancelling_inputs()
    ex = exact_gpp.sigma_exact(a, nv=1)
    d = oracle.sigma_dd(a, nvband=1, all_outputs=True)
    s = oracle.sigma(a, nvband=1)
    # acht_n1_loc (frequency 1 of 2 here: wx = 3, exact terms) of every band
    for n1 in range(3):
        assert d["acht_n1_loc"][n1] == exact_gpp.to_complex(ex["acht_n1_loc"][n1]), n1
    worst_seq = 0.0
    checked = 0
    for k, iws in (("achtemp", (0, 1)), ("asxtemp", (0,))):  # asxtemp at wx=3 has inexact 1/81 terms
        for iw in iws:
            e = exact_gpp.to_complex(ex[k][iw])  # correctly rounded exact value
            assert d[k][iw] == e, (k, iw, d[k][iw], e)  # the dd sum is exact, then rounded once
            worst_seq = max(worst_seq, abs(s[k][iw] - e) / abs(e))
            checked += 1
    assert checked == 3
    assert worst_seq > 1e-3  # the sequential sums lose the small terms
    # asxtemp at wx = 3: terms carry the reference's rounding of 1/81, so dd
    # is within the terms' own rounding of the exact value
    e = exact_gpp.to_complex(ex["asxtemp"][1])
    assert abs(d["asxtemp"][1] - e) <= 2e-16 * ex["asxtemp_mag"][1]


@pytest.mark.parametrize("case", ["seeded", "near_pole", "cancel"])
def test_compensated_oracle_vs_exact_rational(oracle, case):
    if case == "seeded":
        a, nv = gpp_inputs.seeded_inputs(4, 7, 3, nfreq=3, seed=11), 2
    elif case == "near_pole":
        a, nv = gpp_inputs.near_pole_inputs(3, 9, 4, seed=6), 2
    else:
        # each band's row followed by its negation times (1 + 2^-30): the G
        # sums cancel to about 2^-30 of their terms
        a, nv = gpp_inputs.seeded_inputs(3, 12, 2, nfreq=2, seed=3), 1
        for key in ("aqsntemp",):
            h = a[key][:, :6]
            a[key][:, 6:] = -h * (1 + 2.0 ** -30)
        a["I_eps_array"][:, 6:] = a["I_eps_array"][:, :6]
        a["wtilde_array"][:, 6:] = a["wtilde_array"][:, :6]
    ex = exact_gpp.sigma_exact(a, nv)
    d = oracle.sigma_dd(a, nvband=nv, all_outputs=True)
    d2 = oracle.sigma_dd(a, nvband=nv, threads=4)
    for k in ("achtemp", "asxtemp"):
        np.testing.assert_array_equal(d[k], d2[k])
        for iw in range(len(a["wx_array"])):
            e = exact_gpp.to_complex(ex[k][iw])
            scale = ex[k + "_mag"][iw] + 1e-300
            assert abs(d[k][iw] - e) <= 2e-16 * scale, (case, k, iw, d[k][iw], e)
    for n1 in range(a["aqsmtemp"].shape[0]):
        e = exact_gpp.to_complex(ex["acht_n1_loc"][n1])
        assert abs(d["acht_n1_loc"][n1] - e) <= 2e-16 * (ex["acht_n1_loc_mag"][n1] + 1e-300), (case, n1)
    e = exact_gpp.to_complex(ex["achstemp"])
    assert abs(d["achstemp"] - e) <= 2e-16 * (ex["achstemp_mag"] + 1e-300)


# ---- the compiled reference --------------------------------------------------

def _ref_or_skip(oracle):
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built (reference sources absent)")


@pytest.mark.parametrize("case", ["det", "seeded", "near_pole", "nfreq5"])
def test_oracle_bitwise_equals_compiled_reference(oracle, case):
    _ref_or_skip(oracle)
    if case == "det":
        a, nv = gpp_inputs.det_inputs(16, 2, 100, npg=5), 2
    elif case == "seeded":
        a, nv = gpp_inputs.seeded_inputs(10, 200, 30, seed=11), 3
    elif case == "near_pole":
        a, nv = gpp_inputs.near_pole_inputs(6, 256, 16, seed=12), 4
    else:
        a, nv = gpp_inputs.seeded_inputs(5, 100, 10, nfreq=5, seed=13), 2
    o = oracle.sigma(a, nvband=nv)
    r = oracle.sigma(a, nvband=nv, use_ref=True)
    for k in ("achtemp", "asxtemp", "acht_n1_loc"):
        np.testing.assert_array_equal(o[k], r[k], err_msg=k)
    assert o["achstemp"] == r["achstemp"]


@pytest.mark.parametrize("wx,wt,ie,branch", POLE_CASES)
def test_pole_terms_bitwise_equal_compiled_reference(oracle, wx, wt, ie, branch):
    _ref_or_skip(oracle)
    _, sch, ssx = oracle.pole_terms(wx, wt, ie)
    _, rsch, rssx = oracle.pole_terms(wx, wt, ie, use_ref=True)
    assert (sch, ssx) == (rsch, rssx)


def test_oracle_matches_reference_fixtures(oracle):
    """Fixtures written by tests/golden/make_golden.py from the compiled
    reference (oracle/_ref) -- they hold inputs and the reference's outputs."""
    path = os.path.join(GOLDEN_DIR, "ref_sigma_vectors.npz")
    if not os.path.exists(path):
        pytest.skip("fixtures not generated")
    z = np.load(path)
    names = sorted({k.split("__")[0] for k in z.files})
    assert names
    for name in names:
        a = {k: z[f"{name}__{k}"] for k in ("inv_igp_index", "wx_array", "vcoul", "aqsmtemp", "aqsntemp",
                                            "I_eps_array", "wtilde_array")}
        nv = int(z[f"{name}__nvband"])
        o = oracle.sigma(a, nvband=nv)
        for k in ("achtemp", "asxtemp", "acht_n1_loc"):
            np.testing.assert_array_equal(o[k], z[f"{name}__out_{k}"], err_msg=f"{name}:{k}")
        assert o["achstemp"] == z[f"{name}__out_achstemp"]


def test_pole_terms_match_reference_fixtures(oracle):
    path = os.path.join(GOLDEN_DIR, "ref_pole_vectors.npz")
    if not os.path.exists(path):
        pytest.skip("fixtures not generated")
    z = np.load(path)
    wx, wt, ie, sch, ssx = z["wx"], z["wtilde"], z["ieps"], z["sch"], z["ssx"]
    for i in range(len(wx)):
        _, s1, s2 = oracle.pole_terms(float(wx[i]), complex(wt[i]), complex(ie[i]))
        assert s1 == sch[i] or (np.isnan(s1.real) and np.isnan(sch[i].real)), i
        assert s2 == ssx[i] or (np.isnan(s2.real) and np.isnan(ssx[i].real)), i


# ---- band-dependent frequencies (oracle_sigma_wxb) --------------------------

def _one_band(arrs, b, nv):
    sub = dict(arrs)
    sub["aqsmtemp"] = arrs["aqsmtemp"][b:b + 1]
    sub["aqsntemp"] = arrs["aqsntemp"][b:b + 1]
    sub["wx_array"] = np.ascontiguousarray(arrs["wx_array"][b])
    return sub, (1 if b < nv else 0)


def test_per_band_wx_equal_rows_bitwise(oracle):
    a = gpp_inputs.seeded_inputs(9, 200, 11, nfreq=4, seed=41)
    shared = oracle.sigma(a, nvband=3)
    b = dict(a)
    b["wx_array"] = np.tile(a["wx_array"], (9, 1))
    rows = oracle.sigma(b, nvband=3, threads=4)
    for k in ("achtemp", "asxtemp", "acht_n1_loc"):
        np.testing.assert_array_equal(shared[k], rows[k], err_msg=k)
    assert shared["achstemp"] == rows["achstemp"]


@pytest.mark.parametrize("use_ref", [False, True])
def test_per_band_wx_decomposes_by_band(oracle, use_ref):
    """Each band's acht_n1_loc depends on that band's row only: it must equal,
    bit for bit, a one-band run with that row as the shared wx_array (run by
    the oracle, or by the compiled reference harness); the sums over bands
    equal the sum of the one-band runs to rounding.  Band 1 (occupied) gets
    negative frequencies, so the ssx cutoff of gppComplex.cpp:58-59 is live."""
    if use_ref:
        _ref_or_skip(oracle)
    nb, nv = 7, 3
    a = gpp_inputs.seeded_inputs(nb, 160, 10, nfreq=4, seed=42)
    a["wx_array"] = gpp_inputs.per_band_wx(nb, 4, seed=42)
    a["wx_array"][1] = [-3.0, -1.0, 0.5, 2.0]
    assert len({tuple(r) for r in a["wx_array"]}) == nb
    full = oracle.sigma(a, nvband=nv, threads=3)
    tot = {k: 0 for k in ("achtemp", "asxtemp", "achstemp")}
    for b in range(nb):
        sub, snv = _one_band(a, b, nv)
        one = oracle.sigma(sub, nvband=snv, use_ref=use_ref)
        assert full["acht_n1_loc"][b] == one["acht_n1_loc"][0], b
        for k in tot:
            tot[k] = tot[k] + one[k]
    for k in tot:
        assert rel_err(full[k], tot[k]) <= 1e-13, k
    # a row-0 or transposed read of wx would change these:
    shared = oracle.sigma(dict(a, wx_array=np.ascontiguousarray(a["wx_array"][0])), nvband=nv)
    assert not np.array_equal(shared["acht_n1_loc"][1:], full["acht_n1_loc"][1:])


def test_per_band_wx_bad_shape_rejected(oracle):
    a = gpp_inputs.seeded_inputs(4, 50, 3, seed=43)
    a["wx_array"] = np.ones((3, 3))
    with pytest.raises(ValueError):
        oracle.sigma(a, nvband=1)
