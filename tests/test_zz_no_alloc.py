"""d_noFlagOCCSolver_noalloc, the launcher that allocates no device memory
(its scratch is a library-owned device array), against d_noFlagOCCSolver
(GPPComplex.h:310) through the C ABI: bitwise equal results on two calls in
a row, the second with a smaller input (the scratch keeps the first call's
contents), accumulation into caller-zeroed outputs, the size limit and the
index check."""
import numpy as np
import pytest

import gpp_inputs
import hiprt

pytestmark = pytest.mark.gpu


def _device(a):
    return {k: hiprt.DeviceArray(v) for k, v in a.items()}


def _launch(lib, fn, a, d, nv, outs):
    nb, ncouls = a["aqsmtemp"].shape
    ngpown = a["I_eps_array"].shape[0]
    return getattr(lib, fn)(nb, nv, ngpown, ncouls, d["inv_igp_index"].ptr, d["wx_array"].ptr,
                            d["wtilde_array"].ptr, d["aqsmtemp"].ptr, d["aqsntemp"].ptr, d["I_eps_array"].ptr,
                            d["vcoul"].ptr, outs[0].ptr, outs[1].ptr, outs[2].ptr)


def _outs(nb):
    return (hiprt.DeviceArray(np.zeros(3)), hiprt.DeviceArray(np.zeros(3)),
            hiprt.DeviceArray(np.zeros(nb, np.complex128)))


def _get(outs):
    return outs[0].get() + 1j * outs[1].get(), outs[2].get()


CASES = [  # nb, nv, ncouls, ngpown, seed: the second call is smaller
    (96, 3, 6000, 330, 1),
    (20, 2, 1500, 70, 2),
]


def test_noalloc_matches_allocating_launcher(gpplib):
    arrays = [(gpp_inputs.near_pole_inputs(nb, nc, ng, seed=sd), nv) for nb, nv, nc, ng, sd in CASES]
    devs = [_device(a) for a, _ in arrays]
    want = []
    for (a, nv), d in zip(arrays, devs):
        o = _outs(a["aqsmtemp"].shape[0])
        assert _launch(gpplib, "d_noFlagOCCSolver", a, d, nv, o) == 0
        want.append(_get(o))
    got = []
    for (a, nv), d in zip(arrays, devs):  # two calls in a row, the second smaller
        o = _outs(a["aqsmtemp"].shape[0])
        assert _launch(gpplib, "d_noFlagOCCSolver_noalloc", a, d, nv, o) == 0
        got.append((o, _get(o)))
    for (o, (ach, n1)), (wach, wn1) in zip(got, want):
        np.testing.assert_array_equal(ach, wach)
        np.testing.assert_array_equal(n1, wn1)
    # outputs accumulate, as the reference's launchers do (gppCudaVer.cpp:376-377)
    (a, nv), d = arrays[1], devs[1]
    o = got[1][0]
    assert _launch(gpplib, "d_noFlagOCCSolver_noalloc", a, d, nv, o) == 0
    ach, n1 = _get(o)
    np.testing.assert_allclose(ach, 2 * want[1][0], rtol=1e-14)
    np.testing.assert_allclose(n1, 2 * want[1][1], rtol=1e-14, atol=1e-300)


def test_noalloc_benchmark_columns_and_limit(gpplib):
    """The largest supported shape class (512 bands x 1638 columns, the
    benchmark's) runs; a problem whose scratch exceeds the array is refused
    before any device work."""
    import pygpp
    a = gpp_inputs.seeded_inputs(512, 2048, 1638, seed=4)
    d = _device(a)
    o1, o2 = _outs(512), _outs(512)
    assert _launch(gpplib, "d_noFlagOCCSolver", a, d, 2, o1) == 0
    assert _launch(gpplib, "d_noFlagOCCSolver_noalloc", a, d, 2, o2) == 0
    for x, y in zip(_get(o1), _get(o2)):
        np.testing.assert_array_equal(x, y)
    # 1400 bands x 250 column tiles x 3 frequencies of partial sums alone
    # exceed the 16 MiB array (the inputs are real, in-bounds arrays)
    big = gpp_inputs.seeded_inputs(1400, 1024, 1000, seed=5)
    ob = _outs(1400)
    assert _launch(gpplib, "d_noFlagOCCSolver_noalloc", big, _device(big), 2, ob) == pygpp.GPP_ERR_INVALID_ARGUMENT
    assert b"scratch" in gpplib.gpp_last_error()
    assert not np.any(_get(ob)[0])


def test_noalloc_index_check_and_recovery(gpplib):
    """A negative inv_igp_index entry is refused (the flag in the scratch is
    cleared for the next call, which succeeds)."""
    import pygpp
    a = gpp_inputs.seeded_inputs(12, 800, 40, seed=6)
    bad = dict(a)
    bad["inv_igp_index"] = a["inv_igp_index"].copy()
    bad["inv_igp_index"][7] = -2
    o = _outs(12)
    assert _launch(gpplib, "d_noFlagOCCSolver_noalloc", bad, _device(bad), 2, o) == pygpp.GPP_ERR_BAD_INDEX
    d = _device(a)
    o1, o2 = _outs(12), _outs(12)
    assert _launch(gpplib, "d_noFlagOCCSolver_noalloc", a, d, 2, o1) == 0
    assert _launch(gpplib, "d_noFlagOCCSolver", a, d, 2, o2) == 0
    for x, y in zip(_get(o1), _get(o2)):
        np.testing.assert_array_equal(x, y)
