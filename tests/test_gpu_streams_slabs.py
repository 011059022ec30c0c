"""The caller-stream variant of the ABI (gpp_ctx_set_stream) and G'-column
slabs (gpp_ctx_set_columns), through the C ABI.

A host application that owns its streams (BerkeleyGW would) must be able to
order the library's work after its own and its own after the library's,
without synchronising; the reference's launchers leave that ordering to the
caller (gppCudaVer.cpp:381-395).  A slab is one rank's share of the
reference's MPI decomposition (gppMPIOpenMP3.cpp:172, 255-271)."""
import numpy as np
import pytest

import gpp_inputs
import hiprt
from test_gpu_parity import assert_close, problem_of

pytestmark = pytest.mark.gpu

BIG = ("aqsmtemp", "aqsntemp", "I_eps_array", "wtilde_array")


@pytest.mark.parametrize("form", ["direct", "factored-f64"])
def test_caller_stream_orders_library_work(ctx, form):
    """Inputs A are bound; on the caller's stream the caller copies inputs B
    over them (a few hundred MB from pinned memory), enqueues an evaluation,
    then copies A back without waiting.  The evaluation must see B: it runs
    after the first copy and before the second."""
    a = gpp_inputs.seeded_inputs(48, 20000, 400, seed=21)
    b = gpp_inputs.seeded_inputs(48, 20000, 400, seed=22)
    for k in ("inv_igp_index", "wx_array", "vcoul"):
        b[k] = a[k]
    p = problem_of(a, 3)
    ctx.upload(p, b)
    want_b = ctx.sigma(form)
    ctx.upload(p, a)
    want_a = ctx.sigma(form)
    dev = {k: hiprt.DeviceArray(v) for k, v in a.items()}
    pin_a = {k: hiprt.PinnedArray(a[k]) for k in BIG}
    pin_b = {k: hiprt.PinnedArray(b[k]) for k in BIG}
    s = hiprt.Stream()
    try:
        ctx.bind_device(p, {k: d.ptr for k, d in dev.items()})
        ctx.set_stream(s.handle)
        for rep in range(2):
            for k in BIG:
                pin_b[k].copy_to_device_async(dev[k].ptr, s)
            ctx.sigma_async(form)
            for k in BIG:
                pin_a[k].copy_to_device_async(dev[k].ptr, s)
            got = ctx.wait()
            for k in ("achtemp", "asxtemp", "achstemp"):
                np.testing.assert_array_equal(got[k], want_b[k], err_msg=f"rep {rep} {k}")
            s.synchronize()
            # the inputs are A again: an evaluation now sees A
            ctx.sigma_async(form)
            got = ctx.wait()
            for k in ("achtemp", "asxtemp", "achstemp"):
                np.testing.assert_array_equal(got[k], want_a[k], err_msg=f"rep {rep} {k}")
        # synchronous calls work on the caller's stream too
        np.testing.assert_array_equal(ctx.sigma(form)["achtemp"], want_a["achtemp"])
    finally:
        ctx.set_stream(None)
        s.close()
        for d in list(pin_a.values()) + list(pin_b.values()):
            d.close()
    np.testing.assert_array_equal(ctx.sigma(form)["achtemp"], want_a["achtemp"])


def test_set_stream_rejects_nothing_else(ctx):
    """NULL restores the context's own stream; setting it twice is a no-op."""
    ctx.set_stream(None)
    ctx.set_stream(None)


@pytest.mark.parametrize("form", ["direct", "direct-strict", "factored-f64", "factored-ozaki-i8"])
def test_slab_equals_bound_sub_arrays(ctx, form):
    """Evaluating columns [j0, j1) of the uploaded inputs is bitwise the same
    as binding only that slab's rows of I_eps / wtilde and entries of
    inv_igp_index (what a rank of the MPI build holds)."""
    a = gpp_inputs.near_pole_inputs(40, 3000, 120, seed=8)
    p = problem_of(a, 4)
    j0, j1 = 45, 90
    ctx.upload(p, a)
    ctx.set_columns(j0, j1)
    got = ctx.sigma(form, acht_n1_loc=True)
    ctx.set_columns(0, 0)
    s = dict(a)
    for k in ("I_eps_array", "wtilde_array", "inv_igp_index"):
        s[k] = np.ascontiguousarray(a[k][j0:j1])
    dev = {k: hiprt.DeviceArray(v) for k, v in s.items()}
    ps = problem_of(s, 4)
    ctx.bind_device(ps, {k: d.ptr for k, d in dev.items()})
    want = ctx.sigma(form, acht_n1_loc=True)
    for k in ("achtemp", "asxtemp", "acht_n1_loc", "achstemp"):
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)


def test_slabs_vs_oracle(ctx, oracle):
    """Three uneven slabs summed in rank order against the oracle, every form."""
    a = gpp_inputs.seeded_inputs(20, 1500, 70, nfreq=5, seed=12)
    ref = oracle.sigma(a, nvband=3, threads=8)
    p = problem_of(a, 3)
    ctx.upload(p, a)
    bounds = [0, 17, 50, 70]
    for form, tol in (("direct", 1e-12), ("factored-f64", 1e-10), ("factored-ozaki", 1e-10)):
        total = None
        for r in range(3):
            ctx.set_columns(bounds[r], bounds[r + 1])
            part = ctx.sigma(form, acht_n1_loc=True)
            total = dict(part) if total is None else {k: total[k] + part[k] for k in part}
        ctx.set_columns(0, 0)
        assert_close(total, ref, tol)


def test_set_columns_validates(ctx):
    import pygpp
    a = gpp_inputs.seeded_inputs(6, 200, 10, seed=2)
    ctx.upload(problem_of(a, 1), a)
    for j0, j1 in ((-1, 3), (3, 3), (4, 2), (0, 11)):
        with pytest.raises(pygpp.GppError) as e:
            ctx.set_columns(j0, j1)
        assert e.value.status == pygpp.GPP_ERR_INVALID_ARGUMENT
    ctx.set_columns(2, 5)
    ctx.upload(problem_of(a, 1), a)  # new inputs reset the slab
    whole = ctx.sigma("direct")
    ctx.set_columns(0, 10)
    np.testing.assert_array_equal(ctx.sigma("direct")["achtemp"], whole["achtemp"])
    ctx.set_columns(0, 0)
