"""Band-dependent frequencies (gpp_set_band_frequencies).  The direct forms
read a frequency row per band; the factored forms factor each group of bands
that shares a row ("bucket") and fall back to the direct form beyond
GPP_OPT_MAX_BUCKETS groups.  The checker is the oracle's per-band loop
(oracle_sigma_wxb), itself pinned in test_oracle.py by the one-band
decomposition against the oracle and the compiled reference.  Bounds as for
the shared row: 1e-12 per component for the direct forms, 1e-10 for the
factored ones."""
import numpy as np
import pytest

import gpp_inputs
from test_gpu_parity import assert_close, problem_of

pytestmark = pytest.mark.gpu

KEYS_ALL = ("achtemp", "asxtemp", "acht_n1_loc", "achstemp")
DIRECT = ["direct", "direct-strict"]
FACTORED = ["factored-f64", "factored-ozaki", "factored-ozaki-i8"]


def _tol(form):
    return 1e-12 if form.startswith("direct") else 1e-10


def _inputs(nb, nv, ncouls, ngpown, nfreq, seed, levels):
    a = gpp_inputs.seeded_inputs(nb, ncouls, ngpown, nfreq=nfreq, seed=seed)
    a["wx_array"] = gpp_inputs.per_band_wx(nb, nfreq, seed=seed, levels=levels)
    return a


@pytest.mark.parametrize("form", DIRECT + FACTORED)
def test_distinct_rows_vs_oracle(ctx, oracle, form):
    """Every band its own row (40 > 32 groups): the factored forms run the
    direct form, bit for bit."""
    a = _inputs(40, 5, 700, 37, 4, seed=51, levels=None)
    ctx.upload(problem_of(a, 5), a)
    assert ctx.band_buckets() == 40
    out = ctx.sigma(form, acht_n1_loc=True)
    if form in FACTORED:
        assert ctx.last_form() == "direct"
        ref_direct = ctx.sigma("direct")
        for k in KEYS_ALL:
            np.testing.assert_array_equal(out[k], ref_direct[k])
    else:
        assert ctx.last_form() == form
    assert_close(out, oracle.sigma(a, nvband=5, threads=8), 1e-12, KEYS_ALL)


@pytest.mark.parametrize("levels", [2, 5])
@pytest.mark.parametrize("form", DIRECT + FACTORED)
def test_bucketed_vs_oracle(ctx, oracle, form, levels):
    a = _inputs(61, 6, 900, 45, 5, seed=52 + levels, levels=levels)
    ctx.upload(problem_of(a, 6), a)
    nbk = ctx.band_buckets()
    assert 2 <= nbk <= levels
    out = ctx.sigma(form, acht_n1_loc=True)
    assert ctx.last_form() == form
    assert_close(out, oracle.sigma(a, nvband=6, threads=8), _tol(form), KEYS_ALL)


@pytest.mark.parametrize("form", DIRECT + FACTORED)
def test_near_pole_rows(ctx, oracle, form):
    """Three rows that are permutations of (399.4, 4.0, 0.1) over near-pole
    wtilde: every occupied-band branch occurs at every frequency position."""
    a = gpp_inputs.near_pole_inputs(30, 800, 40, seed=53)
    perms = np.array([[399.4, 4.0, 0.1], [0.1, 399.4, 4.0], [4.0, 0.1, 399.4]])
    a["wx_array"] = perms[np.random.default_rng(53).integers(0, 3, 30)]
    ctx.upload(problem_of(a, 7), a)
    assert ctx.band_buckets() == 3
    out = ctx.sigma(form, acht_n1_loc=True)
    assert ctx.last_form() == form
    assert_close(out, oracle.sigma(a, nvband=7, threads=8), _tol(form), KEYS_ALL)


@pytest.mark.parametrize("form", DIRECT + FACTORED)
def test_equal_rows_bitwise_shared(ctx, form):
    """Identical rows give exactly the shared-row results (one bucket is the
    plain factored form; the direct forms read the same values)."""
    a = gpp_inputs.seeded_inputs(20, 500, 30, nfreq=3, seed=54)
    p = problem_of(a, 3)
    ctx.upload(p, a)
    shared = ctx.sigma(form, acht_n1_loc=True)
    ctx.set_band_frequencies(np.tile(a["wx_array"], (20, 1)))
    assert ctx.band_buckets() == 1
    rows = ctx.sigma(form, acht_n1_loc=True)
    assert ctx.last_form() == form
    for k in KEYS_ALL:
        np.testing.assert_array_equal(shared[k], rows[k])


def test_max_buckets_option_and_reset(ctx):
    import pygpp
    a = _inputs(30, 4, 400, 20, 3, seed=55, levels=5)
    ctx.upload(problem_of(a, 4), a)
    nbk = ctx.band_buckets()
    assert nbk > 2
    try:
        ctx.set_option(pygpp.GPP_OPT_MAX_BUCKETS, nbk - 1)
        fell_back = ctx.sigma("factored-f64", acht_n1_loc=True)
        assert ctx.last_form() == "direct"
        ctx.set_option(pygpp.GPP_OPT_MAX_BUCKETS, nbk)
        factored = ctx.sigma("factored-f64", acht_n1_loc=True)
        assert ctx.last_form() == "factored-f64"
        assert_close(factored, fell_back, 1e-10, KEYS_ALL)
        with pytest.raises(pygpp.GppError):
            ctx.set_option(pygpp.GPP_OPT_MAX_BUCKETS, 0)
    finally:
        ctx.set_option(pygpp.GPP_OPT_MAX_BUCKETS, 32)
    # clearing restores the shared row (the first row, as upload sent it)
    ctx.set_band_frequencies(None)
    assert ctx.band_buckets() == 1
    cleared = ctx.sigma("direct")
    b = dict(a, wx_array=np.ascontiguousarray(a["wx_array"][0]))
    ctx.upload(problem_of(b, 4), b)
    assert ctx.band_buckets() == 1
    shared = ctx.sigma("direct")
    for k in ("achtemp", "asxtemp", "achstemp"):
        np.testing.assert_array_equal(cleared[k], shared[k])
    # a fresh upload drops band frequencies set before it
    ctx.set_band_frequencies(a["wx_array"])
    ctx.upload(problem_of(b, 4), b)
    assert ctx.band_buckets() == 1


def test_bucketed_benchmark_shape_subset(ctx, oracle):
    """24 bands at ncouls 32768, ngpown 1638, three energy levels."""
    import pygpp
    p = pygpp.Problem(24, 2, 32768, 1638)
    a = pygpp.fill_seeded(p, seed=5)
    a["wx_array"] = gpp_inputs.per_band_wx(24, 3, seed=5, levels=3)
    ctx.upload(p, a)
    ref = oracle.sigma(a, nvband=2, threads=16)
    for form in ("direct",) + tuple(FACTORED):
        out = ctx.sigma(form, acht_n1_loc=True)
        assert ctx.last_form() == form
        assert_close(out, ref, _tol(form), KEYS_ALL)
