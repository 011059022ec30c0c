"""Enqueued evaluations (gpp_sigma_async + gpp_ctx_wait) and the timing ring
(gpp_ctx_timing_summary): back-to-back evaluations give the synchronous
results bit for bit, the last one's outputs are returned, and misuse is
reported with a status."""
import numpy as np
import pytest

import gpp_inputs
from test_gpu_parity import problem_of

pytestmark = pytest.mark.gpu

KEYS = ("achtemp", "asxtemp", "achstemp")


def test_async_back_to_back_matches_sync(ctx):
    a = gpp_inputs.seeded_inputs(40, 2000, 50, nfreq=3, seed=81)
    ctx.upload(problem_of(a, 3), a)
    sync = {f: ctx.sigma(f, acht_n1_loc=True) for f in ("direct", "factored-f64", "factored-ozaki-i8")}
    for f in ("direct", "factored-ozaki-i8", "factored-f64"):
        ctx.sigma_async(f, acht_n1_loc=True)
    last = ctx.wait(acht_n1_loc=True)
    for k in KEYS + ("acht_n1_loc",):
        np.testing.assert_array_equal(last[k], sync["factored-f64"][k])
    for _ in range(5):
        ctx.sigma_async("direct")
    out = ctx.wait(acht_n1_loc=True)  # the direct form always computes acht_n1_loc
    for k in KEYS + ("acht_n1_loc",):
        np.testing.assert_array_equal(out[k], sync["direct"][k])
    t = ctx.timing_summary(5)
    assert t["evaluations"] == 5 and t["total"] > 0
    assert t["direct_unoccupied"] > 0
    # only the trailing evaluations of the same kind are averaged
    ctx.sigma_async("factored-f64")
    ctx.wait()
    assert ctx.timing_summary(32)["evaluations"] == 1


def test_async_misuse(ctx):
    import pygpp
    a = gpp_inputs.seeded_inputs(8, 300, 10, seed=82)
    ctx.upload(problem_of(a, 2), a)
    with pytest.raises(pygpp.GppError) as e:
        ctx.wait()  # nothing enqueued on these inputs
    assert e.value.status == pygpp.GPP_ERR_NO_INPUTS
    ctx.sigma_async("factored-f64", acht_n1_loc=False)
    with pytest.raises(pygpp.GppError) as e:
        ctx.wait(acht_n1_loc=True)
    assert e.value.status == pygpp.GPP_ERR_UNSUPPORTED
    out = ctx.wait()
    ref = ctx.sigma("factored-f64")
    for k in KEYS:
        np.testing.assert_array_equal(out[k], ref[k])
    with pytest.raises(pygpp.GppError):
        ctx.timing_summary(0)
