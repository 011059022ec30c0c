"""The FP64 Gram product with the pole reduce in its epilogue
(GPP_OPT_FUSED_REDUCE = 1: factored-f64 with at most three frequencies, none
negative, at most 4 occupied bands, the 128x128 tile and acht_n1_loc not
requested; no G'xG workspace):
against the oracle (gppComplex.cpp:335-383 restated) within 1e-10 per
component, against the separate Gram + reduce path within 1e-12, and bitwise
repeatable.  Shapes cover partial tiles in both directions (columns and
G-vectors not multiples of 128), 0 to 4 occupied bands (the Mo loop), no
unoccupied band (Mu = 0), 1 to 3 frequencies, and inputs that drive every
branch of the occupied solver (near_pole_inputs)."""
import numpy as np
import pytest

import gpp_inputs
from test_gpu_parity import assert_close, fused_reduce, problem_of

pytestmark = pytest.mark.gpu

TOL_ORACLE = 1e-10
TOL_SEPARATE = 1e-12
KEYS = ("achtemp", "asxtemp", "achstemp")

SHAPES = [  # nb, nv, ncouls, ngpown, nfreq
    (40, 2, 3000, 300, 3),    # partial tiles: 300 = 2*128 + 44 columns, 3000 = 23*128 + 56 G-vectors
    (37, 0, 1000, 129, 3),    # no occupied band
    (20, 4, 2000, 140, 2),    # four occupied bands
    (4, 4, 700, 100, 3),      # no unoccupied band: Mu = 0
    (30, 3, 128, 128, 1),     # one whole tile
    (50, 3, 5000, 257, 3),    # three occupied bands
]


def stages(ctx):
    return set(ctx.last_timing())


def evaluate(ctx, fused):
    with fused_reduce(ctx, fused):
        out = ctx.sigma("factored-f64")
        return out, stages(ctx)


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_fused_vs_oracle_and_separate(ctx, oracle, shape):
    nb, nv, ncouls, ngpown, nfreq = shape
    arrs = gpp_inputs.seeded_inputs(nb, ncouls, ngpown, nfreq=nfreq, seed=sum(shape), dw=0.7)
    ctx.upload(problem_of(arrs, nv), arrs)
    fused, names = evaluate(ctx, True)
    assert "factored_gram_pole" in names and "factored_pole_reduce" not in names, names
    sep, names = evaluate(ctx, False)
    assert "factored_pole_reduce" in names and "factored_gram_pole" not in names, names
    ref = oracle.sigma(arrs, nvband=nv, threads=8)
    assert_close(fused, ref, TOL_ORACLE, KEYS)
    assert_close(sep, ref, TOL_ORACLE, KEYS)
    assert_close(fused, sep, TOL_SEPARATE, KEYS)
    again, _ = evaluate(ctx, True)
    for k in KEYS:
        np.testing.assert_array_equal(again[k], fused[k], err_msg=k)


def test_fused_every_branch(ctx, oracle):
    """wx = (399.4, 4.0, 0.1) against wtilde placed at and around the poles:
    branch 1, branch 2, extreme near-pole factors and the 2^-40 bands of the
    |delw|^2 tests, decided as the reference decides them."""
    arrs = gpp_inputs.near_pole_inputs(24, 3000, 150, seed=3)
    ctx.upload(problem_of(arrs, 4), arrs)
    fused, names = evaluate(ctx, True)
    assert "factored_gram_pole" in names
    sep, _ = evaluate(ctx, False)
    ref = oracle.sigma(arrs, nvband=4, threads=8)
    assert_close(fused, ref, TOL_ORACLE, KEYS)
    assert_close(fused, sep, TOL_SEPARATE, KEYS)


def test_fused_not_taken_outside_its_scope(ctx):
    """Four frequencies, a negative frequency, five occupied bands, another
    Gram tile, or an acht_n1_loc request run the separate reduce."""
    with fused_reduce(ctx, True):
        check_scope(ctx)


def check_scope(ctx):
    import pygpp
    arrs = gpp_inputs.seeded_inputs(12, 600, 40, nfreq=4, seed=5)
    ctx.upload(problem_of(arrs, 2), arrs)
    ctx.sigma("factored-f64")
    assert "factored_pole_reduce" in stages(ctx)
    arrs = gpp_inputs.seeded_inputs(12, 600, 40, nfreq=3, seed=5)
    arrs["wx_array"] = np.array([-1.5, 0.5, 2.5])
    ctx.upload(problem_of(arrs, 2), arrs)
    ctx.sigma("factored-f64")
    assert "factored_pole_reduce" in stages(ctx)
    arrs = gpp_inputs.seeded_inputs(12, 600, 40, nfreq=3, seed=5)
    ctx.upload(problem_of(arrs, 5), arrs)
    ctx.sigma("factored-f64")
    assert "factored_pole_reduce" in stages(ctx)
    ctx.upload(problem_of(arrs, 2), arrs)
    ctx.sigma("factored-f64", acht_n1_loc=True)
    assert "factored_pole_reduce" in stages(ctx)
    ctx.set_option(pygpp.GPP_OPT_GRAM_TILE, 1)
    try:
        ctx.sigma("factored-f64")
        assert "factored_pole_reduce" in stages(ctx)
    finally:
        ctx.set_option(pygpp.GPP_OPT_GRAM_TILE, 0)
    ctx.sigma("factored-f64")
    assert "factored_gram_pole" in stages(ctx)


def test_fused_option_values(ctx):
    """Off by default; 0 and 1 only."""
    import pygpp
    assert ctx.get_option(pygpp.GPP_OPT_FUSED_REDUCE) == 0
    with pytest.raises(pygpp.GppError):
        ctx.set_option(pygpp.GPP_OPT_FUSED_REDUCE, 2)
    assert ctx.get_option(pygpp.GPP_OPT_FUSED_REDUCE) == 0
