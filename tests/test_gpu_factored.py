"""Parity of the factored forms (band sum through the complex Gram product on
MFMA, then the FP64 pole-factor reduce) with the FP64 oracle.  Bound: 1e-10
relative per component (components below 1e-6 of their modulus are compared
relative to the modulus).  The FP64-MFMA form is expected near 1e-14."""
import numpy as np
import pytest

import gpp_inputs
from test_gpu_parity import (assert_close, assert_close_conditioned, comp_rel_err, fused_reduce, problem_of, run_gpu,
                             separate_reduce)

pytestmark = pytest.mark.gpu

TOL_FACTORED = 1e-10
KEYS = ("achtemp", "asxtemp", "achstemp")
FACTORED = ["factored-f64", "factored-ozaki", "factored-ozaki-i8"]

SHAPES = [
    (16, 2, 1000, 37, 3),
    (13, 3, 777, 21, 3),
    (9, 0, 130, 9, 3),
    (5, 5, 200, 6, 3),
    (11, 2, 65, 5, 2),
    (10, 2, 300, 10, 5),
    (6, 2, 256, 8, 16),
    (130, 4, 300, 140, 3),   # K not a multiple of the band stage; several Gram tiles
    (1, 0, 1, 1, 3),
]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("form", FACTORED)
def test_factored_seeded_vs_oracle(ctx, oracle, shape, form):
    nb, nv, ncouls, ngpown, nfreq = shape
    arrs = gpp_inputs.seeded_inputs(nb, ncouls, ngpown, nfreq=nfreq, seed=sum(shape) + 7,
                                    dw=0.25 if nfreq > 5 else 1.0)
    assert_close(run_gpu(ctx, arrs, nv, form), oracle.sigma(arrs, nvband=nv, threads=8), TOL_FACTORED, KEYS)


@pytest.mark.parametrize("form", FACTORED)
def test_factored_near_pole(ctx, oracle, form):
    arrs = gpp_inputs.near_pole_inputs(40, 1500, 70, seed=3)
    assert_close(run_gpu(ctx, arrs, 4, form), oracle.sigma(arrs, nvband=4, threads=8), TOL_FACTORED, KEYS)


@pytest.mark.parametrize("form", FACTORED)
def test_factored_branch_thresholds(ctx, oracle, form):
    """Two thirds of the elements put the reference's |delw|^2 within a few
    ulps of limitone (wx = 399.4) or to1 (wx = 0.1), on both sides.  The
    factored forms decide gppComplex.cpp:35 and :43 from an estimate of
    |delw|^2 away from the thresholds and from the reference's own rounding
    near them: every decision, and so every output, matches the oracle (a
    single wrong decision moves an output by far more than 1e-10)."""
    arrs = gpp_inputs.threshold_inputs(24, 2000, 60, seed=5)
    arrs.pop("_kind")
    ref = oracle.sigma(arrs, nvband=4, threads=8)
    assert_close(run_gpu(ctx, arrs, 4, form), ref, TOL_FACTORED, KEYS)
    assert_close(ctx.sigma(form, acht_n1_loc=True), ref, TOL_FACTORED, KEYS + ("acht_n1_loc",))


@pytest.mark.parametrize("form", FACTORED + ["direct", "direct-strict"])
def test_ssx_cutoff_straddle(ctx, oracle, form):
    """Two thirds of the elements put the reference's |ssx| within a few ulps
    of the cutoff 4|I_eps| of gppComplex.cpp:58-59 at a negative frequency, on
    both sides, in branch 1 (wx = -0.6) and branch 2 (wx = -0.2).  The
    factored reduce decides branch 1 on the reference's own rounding of ssx
    and branch 2 on its estimate outside a 2^-40 band (the reference's ssx
    recomputed inside it): every decision, and so every output, matches the
    oracle (one wrong decision moves asxtemp by far more than 1e-10)."""
    arrs = gpp_inputs.ssx_cutoff_inputs(24, 2000, 60, seed=5)
    arrs.pop("_kind")
    ref = oracle.sigma(arrs, nvband=4, threads=8)
    tol = 1e-12 if form.startswith("direct") else TOL_FACTORED
    assert_close(run_gpu(ctx, arrs, 4, form), ref, tol, KEYS)
    if not form.startswith("direct"):
        assert_close(ctx.sigma(form, acht_n1_loc=True), ref, tol, KEYS + ("acht_n1_loc",))


NEG_WX = {"nfreq3": [-2.0, -0.5, 3.0], "nfreq4": [-2.0, -0.5, 1.0, 3.0],
          "nfreq5": [-3.0, -0.7, 0.4, 2.0, 4.0]}


@pytest.mark.parametrize("wx", list(NEG_WX), ids=list(NEG_WX))
@pytest.mark.parametrize("form", FACTORED)
def test_factored_negative_frequencies(ctx, oracle, form, wx):
    """Negative frequencies make the ssx cutoff of gppComplex.cpp:58-59 live
    (|ssx| > 4 |I_eps| is dropped): per-frequency P (passes of up to 3
    frequencies) and the products folded once per element (passes of 4 and
    8) both take their explicit-ssx branch there."""
    arrs = gpp_inputs.seeded_inputs(20, 1500, 60, nfreq=len(NEG_WX[wx]), seed=61)
    arrs["wx_array"] = np.array(NEG_WX[wx])
    # branch-1 elements at wx = -0.5 with |wt| > 1 have |ssx| = |wt|^2 |I_eps| / wx^2 > 4 |I_eps|
    wt = arrs["wtilde_array"]
    assert np.count_nonzero((np.abs(wt) > 1.0) & (np.abs(-0.5 - wt) > 0.5)) > 1000
    ref = oracle.sigma(arrs, nvband=4, threads=8)
    assert_close(run_gpu(ctx, arrs, 4, form), ref, TOL_FACTORED, KEYS)
    assert_close(ctx.sigma(form, acht_n1_loc=True), ref, TOL_FACTORED, KEYS + ("acht_n1_loc",))


@pytest.mark.parametrize("ncouls", [800, 3200])
@pytest.mark.parametrize("form", FACTORED)
def test_factored_readme_golden(ctx, oracle, ncouls, form):
    from test_oracle import README_GOLDEN
    arrs = gpp_inputs.det_inputs(512, 2, ncouls, npg=20)
    out = run_gpu(ctx, arrs, 2, form)
    gold = np.array([complex(r, i) for r, i in README_GOLDEN[ncouls]])
    assert comp_rel_err(out["achtemp"], gold) <= TOL_FACTORED
    assert_close(out, oracle.sigma(arrs, nvband=2, threads=8), TOL_FACTORED, KEYS)


ACHT_SHAPES = SHAPES + [
    (300, 9, 1500, 300, 3),   # several column tiles of both band ranges, 40+ occupied-tile columns
    (40, 40, 600, 30, 3),     # every band occupied: no unoccupied launch
    (70, 0, 2100, 129, 4),    # no occupied band; j tiles with one live row
]


@pytest.mark.parametrize("shape", ACHT_SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("form", FACTORED)
def test_factored_acht_n1_loc_vs_oracle(ctx, oracle, shape, form):
    """acht_n1_loc from the second product (pole factor of frequency 2, branch
    mask for the occupied bands) against the oracle's per-band sums; asking for
    it leaves the other outputs bit-identical to those of the separate pole
    reduce (the stages it runs)."""
    nb, nv, ncouls, ngpown, nfreq = shape
    arrs = gpp_inputs.seeded_inputs(nb, ncouls, ngpown, nfreq=nfreq, seed=sum(shape) + 11,
                                    dw=0.25 if nfreq > 5 else 1.0)
    ctx.upload(problem_of(arrs, nv), arrs)
    out = ctx.sigma(form, acht_n1_loc=True)
    assert_close(out, oracle.sigma(arrs, nvband=nv, threads=8), TOL_FACTORED, KEYS + ("acht_n1_loc",))
    with separate_reduce(ctx):
        plain = ctx.sigma(form)
    for k in KEYS:
        np.testing.assert_array_equal(out[k], plain[k])


@pytest.mark.parametrize("form", FACTORED)
def test_factored_acht_n1_loc_near_pole(ctx, oracle, form):
    """Occupied-band branch decisions at frequency 2 straddle their thresholds."""
    arrs = gpp_inputs.near_pole_inputs(40, 1500, 70, seed=5)
    ctx.upload(problem_of(arrs, 6), arrs)
    ref = oracle.sigma(arrs, nvband=6, threads=8)
    assert_close(ctx.sigma(form, acht_n1_loc=True), ref, TOL_FACTORED, KEYS + ("acht_n1_loc",))


def test_factored_full_c1_band_subset(ctx, oracle):
    import pygpp
    p = pygpp.Problem(24, 2, 32768, 1638)
    arrs = pygpp.fill_seeded(p, seed=3)
    ctx.upload(p, arrs)
    assert_close(ctx.sigma("factored-f64"), oracle.sigma(arrs, nvband=2, threads=16), TOL_FACTORED, KEYS)


def test_factored_full_c1_vs_direct_strict(ctx):
    """Full benchmark size: every factored form against the strict direct
    kernel, the Ozaki forms also at 5 slices (39-bit digits)."""
    import pygpp
    p = pygpp.Problem(512, 2, 32768, 1638)
    ctx.generate_seeded(p, seed=1)
    ref = ctx.sigma("direct-strict")
    for form in FACTORED:
        out = ctx.sigma(form)
        assert_close(out, ref, TOL_FACTORED, KEYS)
        if form == "factored-f64":
            assert_close(ctx.sigma(form, acht_n1_loc=True), ref, TOL_FACTORED, KEYS + ("acht_n1_loc",))
        again = ctx.sigma(form)
        for k in KEYS:
            np.testing.assert_array_equal(out[k], again[k])
    try:
        ctx.set_option(pygpp.GPP_OPT_OZAKI_SLICES, 5)
        for form in ("factored-ozaki", "factored-ozaki-i8"):
            assert_close(ctx.sigma(form), ref, TOL_FACTORED, KEYS)
    finally:
        ctx.set_option(pygpp.GPP_OPT_OZAKI_SLICES, 6)


@pytest.mark.parametrize("form", ["factored-ozaki", "factored-ozaki-i8"])
def test_ozaki_error_falls_with_slices(ctx, oracle, form):
    """Each slice after the first carries about 8 more bits (base-254 digits),
    so the error must fall by roughly 2^8 per slice; FP16 at 6 slices keeps
    7-bit digits (42 bits against 39 at 5), a fall of about 2^3.  5 and 6
    slices must meet the 1e-10 bound."""
    import pygpp
    arrs = gpp_inputs.seeded_inputs(40, 1200, 50, seed=21)
    ctx.upload(problem_of(arrs, 2), arrs)
    ref = oracle.sigma(arrs, nvband=2, threads=8)
    errs = []
    try:
        for s in (3, 4, 5, 6):
            ctx.set_option(pygpp.GPP_OPT_OZAKI_SLICES, s)
            errs.append(comp_rel_err(ctx.sigma(form)["achtemp"], ref["achtemp"]))
    finally:
        ctx.set_option(pygpp.GPP_OPT_OZAKI_SLICES, 6)
    assert errs[2] <= TOL_FACTORED and errs[3] <= TOL_FACTORED, errs
    falls = [a / b for a, b in zip(errs, errs[1:])]
    assert falls[0] > 32 and falls[1] > 32, errs
    assert falls[2] > (2 if form == "factored-ozaki" else 32), errs


@pytest.mark.parametrize("nb", [30, 300])
def test_ozaki_layouts_and_digit_types_agree(ctx, oracle, nb):
    """All block layouts and tile orders compute the same integer sums
    (bit-identical outputs).  At 5 slices FP16 and INT8 use the same digits
    and sum levels 0-3 exactly; they differ in FP16's rounding of the higher
    levels and where FP64 rounding of the sums happens (within 1e-12).  300
    bands: several band steps per tile and two FP16 chunks."""
    import pygpp
    arrs = gpp_inputs.near_pole_inputs(nb, 700, 70, seed=22)
    ctx.upload(problem_of(arrs, 3), arrs)
    out = {}
    try:
        for lay, group in ((0, 0), (1, 0), (2, 0), (2, 1), (2, 5), (3, 0), (3, 1), (4, 0), (4, 1)):
            ctx.set_option(pygpp.GPP_OPT_OZAKI_LAYOUT, lay)
            ctx.set_option(pygpp.GPP_OPT_OZAKI_GROUP, group)
            for form in ("factored-ozaki", "factored-ozaki-i8"):
                out[(form, lay, group)] = ctx.sigma(form)
    finally:
        ctx.set_option(pygpp.GPP_OPT_OZAKI_LAYOUT, 4)
        ctx.set_option(pygpp.GPP_OPT_OZAKI_GROUP, 0)
    for form in ("factored-ozaki", "factored-ozaki-i8"):
        for key in [(0, 0), (1, 0), (2, 1), (2, 5), (3, 0), (3, 1), (4, 0), (4, 1)]:
            for k in KEYS:
                np.testing.assert_array_equal(out[(form,) + key][k], out[(form, 2, 0)][k])
    ref = oracle.sigma(arrs, nvband=3, threads=8)
    assert_close(out[("factored-ozaki-i8", 2, 0)], ref, TOL_FACTORED, KEYS)
    assert_close(out[("factored-ozaki", 2, 0)], ref, TOL_FACTORED, KEYS)
    five = {}
    try:
        ctx.set_option(pygpp.GPP_OPT_OZAKI_SLICES, 5)
        for form in ("factored-ozaki", "factored-ozaki-i8"):
            five[form] = ctx.sigma(form)
    finally:
        ctx.set_option(pygpp.GPP_OPT_OZAKI_SLICES, 6)
    assert_close(five["factored-ozaki"], five["factored-ozaki-i8"], 1e-12, KEYS)
    assert_close(five["factored-ozaki-i8"], ref, TOL_FACTORED, KEYS)


def test_ozaki_long_band_range(ctx, oracle):
    """33000 unoccupied bands: the INT8 form flushes its exact INT32 level
    sums to FP64 every 8192 bands (five chunks), the FP16 form every 256;
    both within 1e-10 of the oracle, every layout bit-identical."""
    import pygpp
    arrs = gpp_inputs.seeded_inputs(33002, 96, 6, nfreq=3, seed=41)
    ctx.upload(problem_of(arrs, 2), arrs)
    ref = oracle.sigma(arrs, nvband=2, threads=8)
    out = {}
    for lay in (3, 1, 4):
        ctx.set_option(pygpp.GPP_OPT_OZAKI_LAYOUT, lay)
        try:
            for form in ("factored-ozaki-i8", "factored-ozaki"):
                out[(form, lay)] = ctx.sigma(form)
        finally:
            ctx.set_option(pygpp.GPP_OPT_OZAKI_LAYOUT, 4)
    for form in ("factored-ozaki-i8", "factored-ozaki"):
        assert_close(out[(form, 3)], ref, TOL_FACTORED, KEYS)
        for k in KEYS:
            np.testing.assert_array_equal(out[(form, 1)][k], out[(form, 3)][k])
            np.testing.assert_array_equal(out[(form, 4)][k], out[(form, 3)][k])


def test_ozaki_option_validation(ctx):
    import pygpp
    for bad in (2, 7):
        with pytest.raises(pygpp.GppError):
            ctx.set_option(pygpp.GPP_OPT_OZAKI_SLICES, bad)
    for opt, bad in ((pygpp.GPP_OPT_OZAKI_LAYOUT, 5), (pygpp.GPP_OPT_OZAKI_LAYOUT, -1),
                     (pygpp.GPP_OPT_OZAKI_GROUP, -1)):
        with pytest.raises(pygpp.GppError):
            ctx.set_option(opt, bad)
    assert ctx.get_option(pygpp.GPP_OPT_OZAKI_LAYOUT) == 4
    assert ctx.get_option(pygpp.GPP_OPT_OZAKI_GROUP) == 0
    assert ctx.get_option(pygpp.GPP_OPT_OZAKI_SLICES) == 6


def test_wide_passes_negative_frequencies(ctx, oracle):
    """20 frequencies, some negative (the ssx cutoff of gppComplex.cpp:58-59
    live): the automatic pass width keeps 8 there; passes of 16 asked for
    explicitly run the negative-frequency kernel at that width.  Both within
    1e-10 of the oracle and bitwise equal (the folding depends on the
    frequency count only)."""
    import pygpp
    arrs = gpp_inputs.seeded_inputs(20, 1500, 60, nfreq=20, seed=67, dw=0.5)
    arrs["wx_array"] = np.linspace(-2.5, 4.5, 20)
    ctx.upload(problem_of(arrs, 3), arrs)
    ref = oracle.sigma(arrs, nvband=3, threads=8)
    out = {}
    try:
        for fc in (0, 8, 16):
            ctx.set_option(pygpp.GPP_OPT_REDUCE_FREQS, fc)
            out[fc] = ctx.sigma("factored-f64")
            assert_close(out[fc], ref, TOL_FACTORED, KEYS)
    finally:
        ctx.set_option(pygpp.GPP_OPT_REDUCE_FREQS, 0)
    for k in KEYS:
        np.testing.assert_array_equal(out[16][k], out[8][k], err_msg=k)
        np.testing.assert_array_equal(out[0][k], out[8][k], err_msg=k)


def test_acht_two_occupied_bands_schedules(ctx, oracle):
    """acht_n1_loc with two occupied bands, whose terms the p2-storing reduce
    sums itself (the other counts stream the branch mask through
    k_acht_occ): against the oracle with 12 G-vector splits per column and
    one, over pass and block widths that put frequency 2 in passes of 4, 8
    and 16 and in blocks of 4 and 8 frequencies, bit-identical across the
    widths within a chunk; then with a negative frequency (the generic
    kernel, mask and k_acht_occ)."""
    import pygpp
    keys = KEYS + ("acht_n1_loc",)
    arrs = gpp_inputs.seeded_inputs(24, 3000, 40, nfreq=20, seed=91, dw=0.3)
    ctx.upload(problem_of(arrs, 2), arrs)
    ref = oracle.sigma(arrs, nvband=2, threads=8)
    try:
        for chunk in (256, 8192):
            ctx.set_option(pygpp.GPP_OPT_REDUCE_CHUNK, chunk)
            base = None
            for fc, fpb in ((4, 0), (8, 0), (16, 0), (8, 4), (4, 8)):
                ctx.set_option(pygpp.GPP_OPT_REDUCE_FREQS, fc)
                ctx.set_option(pygpp.GPP_OPT_REDUCE_FREQS_PER_BLOCK, fpb)
                out = ctx.sigma("factored-f64", acht_n1_loc=True)
                assert_close(out, ref, TOL_FACTORED, keys)
                if base is None:
                    base = out
                for k in keys:
                    np.testing.assert_array_equal(out[k], base[k], err_msg=f"{chunk} {fc} {fpb} {k}")
    finally:
        ctx.set_option(pygpp.GPP_OPT_REDUCE_CHUNK, 8192)
        ctx.set_option(pygpp.GPP_OPT_REDUCE_FREQS, 0)
        ctx.set_option(pygpp.GPP_OPT_REDUCE_FREQS_PER_BLOCK, 0)
    arrs["wx_array"] = np.linspace(-1.5, 3.5, 20)
    ctx.upload(problem_of(arrs, 2), arrs)
    assert_close(ctx.sigma("factored-f64", acht_n1_loc=True), oracle.sigma(arrs, nvband=2, threads=8),
                 TOL_FACTORED, keys)


def test_reduce_freqs_option(ctx):
    """Passes of 4, 8 or 16 frequencies, or 0 (the default): 16 from 16
    frequencies up, else 8."""
    import pygpp
    assert ctx.get_option(pygpp.GPP_OPT_REDUCE_FREQS) == 0
    for bad in (-1, 1, 12, 32):
        with pytest.raises(pygpp.GppError):
            ctx.set_option(pygpp.GPP_OPT_REDUCE_FREQS, bad)
    assert ctx.get_option(pygpp.GPP_OPT_REDUCE_FREQS) == 0


ALL_FORMS = ["direct", "factored-f64", "factored-ozaki", "factored-ozaki-i8"]


@pytest.mark.parametrize("nfreq", [4, 11, 13, 18, 20, 28])
def test_reduce_schedules_agree(ctx, oracle, nfreq):
    """Frequencies per pass (4; 8: passes of 8 while more than 8 remain,
    then one of 8, 6, 4 or 2, 13 = 8 + 6, 18 = 8 + 8 + 2; 16: passes of 16
    while more than 12 remain, then one of 8 and one of 4, 28 = 16 + 8 + 4)
    and per block (all, 4, 8) regroup the same per-frequency sums:
    bit-identical for one G-range size; every schedule, at G ranges of 256
    and 8192, within 1e-10 of the oracle."""
    import pygpp
    # every frequency positive (negative ones run at most 8 per pass)
    arrs = gpp_inputs.seeded_inputs(30, 3000, 64, nfreq=nfreq, seed=70 + nfreq, dw=0.3 if nfreq <= 20 else 0.25)
    assert arrs["wx_array"].min() > 0
    ctx.upload(problem_of(arrs, 4), arrs)
    ref = oracle.sigma(arrs, nvband=4, threads=8)
    try:
        for chunk in (256, 8192):
            ctx.set_option(pygpp.GPP_OPT_REDUCE_CHUNK, chunk)
            base = None
            for fc in (4, 8, 16):
                for fpb in (0, 4, 8):
                    ctx.set_option(pygpp.GPP_OPT_REDUCE_FREQS, fc)
                    ctx.set_option(pygpp.GPP_OPT_REDUCE_FREQS_PER_BLOCK, fpb)
                    out = ctx.sigma("factored-f64")
                    assert_close(out, ref, TOL_FACTORED, KEYS)
                    if base is None:
                        base = out
                    for k in KEYS:
                        np.testing.assert_array_equal(out[k], base[k], err_msg=f"{chunk} {fc} {fpb} {k}")
        ach = ctx.sigma("factored-f64", acht_n1_loc=True)
        assert_close(ach, ref, TOL_FACTORED, KEYS + ("acht_n1_loc",))
    finally:
        ctx.set_option(pygpp.GPP_OPT_REDUCE_CHUNK, 8192)
        ctx.set_option(pygpp.GPP_OPT_REDUCE_FREQS, 0)
        ctx.set_option(pygpp.GPP_OPT_REDUCE_FREQS_PER_BLOCK, 0)


GRAM_TILE_SHAPES = [
    (37, 2, 12000, 300, 3),   # K = 35: a partial last band step
    (6, 2, 40000, 200, 3),    # 4 bands, fewer than one band step
    (5, 5, 40000, 130, 3),    # no unoccupied band: Mu is zero
    (130, 4, 300, 140, 3),    # partial tiles in both directions
]


@pytest.mark.parametrize("shape", GRAM_TILE_SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_gram_tiles_agree(ctx, oracle, shape):
    """The four FP64 Gram tilings accumulate every Mu entry over the bands in
    the same order: bit-identical outputs (with and without acht_n1_loc; the
    pole reduce as its own stage), within 1e-10 of the oracle, as is the
    128 x 128 tile with the reduce in its epilogue."""
    import pygpp
    nb, nv, ncouls, ngpown, nfreq = shape
    arrs = gpp_inputs.seeded_inputs(nb, ncouls, ngpown, nfreq=nfreq, seed=nb + ngpown)
    ctx.upload(problem_of(arrs, nv), arrs)
    out = {}
    try:
        with separate_reduce(ctx):
            for tile in range(4):
                ctx.set_option(pygpp.GPP_OPT_GRAM_TILE, tile)
                out[tile] = ctx.sigma("factored-f64")
                out[(tile, "acht")] = ctx.sigma("factored-f64", acht_n1_loc=True)
    finally:
        ctx.set_option(pygpp.GPP_OPT_GRAM_TILE, 0)
    with fused_reduce(ctx, True):
        assert_close(ctx.sigma("factored-f64"), oracle.sigma(arrs, nvband=nv, threads=8), TOL_FACTORED, KEYS)
    for tile in range(1, 4):
        for k in KEYS:
            np.testing.assert_array_equal(out[tile][k], out[0][k], err_msg=f"tile {tile} {k}")
        for k in KEYS + ("acht_n1_loc",):
            np.testing.assert_array_equal(out[(tile, "acht")][k], out[(0, "acht")][k], err_msg=f"tile {tile} {k}")
    assert_close(out[3], oracle.sigma(arrs, nvband=nv, threads=8), TOL_FACTORED, KEYS)


def test_c0_cpu_config(ctx, oracle):
    """The CPU-runnable configuration: 512 2 4096 10 (ngpown 409), the
    reference's deterministic init, every form against the oracle; also the
    ngpown=410 and zero-based index variants."""
    for ngpown, zero in ((409, False), (410, True)):
        arrs = gpp_inputs.det_inputs(512, 2, 4096, ngpown=ngpown, igp_zero=zero)
        ref = oracle.sigma(arrs, nvband=2, threads=16)
        ctx.upload(problem_of(arrs, 2), arrs)
        for form in ALL_FORMS:
            tol = 1e-12 if form == "direct" else TOL_FACTORED
            assert_close(ctx.sigma(form, acht_n1_loc=True), ref, tol, KEYS + ("acht_n1_loc",))


def test_c3_shape_subset_vs_oracle(ctx, oracle):
    """The large single-instance shape (ncouls 65536, nfreq 16 on the
    4 + 0.25 (iw - 7.5) grid, 8 occupied bands) on a subset of bands and
    columns small enough for the oracle.  Some achtemp components cancel
    strongly here (the reference's own left-to-right sum is 2.3e-12 off the
    exact sum of its terms on one of them), so parity is judged against the
    compensated oracle and the sequential oracle's measured error."""
    import pygpp
    p = pygpp.Problem(12, 8, 65536, 48, nfreq=16)
    arrs = pygpp.fill_seeded(p, seed=33, wx_center=4.0, dw=0.25)
    seq = oracle.sigma(arrs, nvband=8, threads=16)
    dd = oracle.sigma_dd(arrs, nvband=8, threads=16)
    ctx.upload(p, arrs)
    for form in ALL_FORMS:
        tol = 1e-12 if form == "direct" else TOL_FACTORED
        out = ctx.sigma(form, acht_n1_loc=True)
        assert_close_conditioned(out, seq, dd, tol)
        assert_close(out, seq, tol, ("achstemp", "acht_n1_loc"))


def test_c3_full_size_forms_agree(ctx):
    """Full large instance (1536, 8, 65536, 3276, nfreq 16, ~10 GB resident):
    the factored forms against the direct FP64 kernel."""
    import pygpp
    p = pygpp.Problem(1536, 8, 65536, 3276, nfreq=16)
    ctx.generate_seeded(p, seed=3, wx_center=4.0, dw=0.25)
    ref = ctx.sigma("direct")
    for form in ("factored-f64", "factored-ozaki-i8"):
        assert_close(ctx.sigma(form), ref, TOL_FACTORED, KEYS)
