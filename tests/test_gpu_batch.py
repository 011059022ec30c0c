"""The batched factored forms: many uneven problems through one launch per
stage (the pole reduce: one per occupied band class), replayed from a
captured hipGraph.  Each instance must
match the oracle -- the FP64 Gram product within 1e-10 per real / imaginary
component, the Ozaki INT8 / FP16 products within 1e-10 of each output's
modulus (the digit split's error bound is normwise) -- and, for the same tile
and reduce schedule, be bitwise identical to the single-problem path of the
same form."""
import numpy as np
import pytest

import gpp_inputs
import hiprt
from test_gpu_parity import assert_close, problem_of, separate_reduce

pytestmark = pytest.mark.gpu

TOL = 1e-10
KEYS = ("achtemp", "asxtemp", "achstemp")

SHAPES = [  # nb, nv, ncouls, ngpown, nfreq
    (24, 2, 700, 35, 3),
    (9, 0, 130, 9, 1),
    (40, 4, 1200, 60, 9),
    (17, 3, 513, 25, 5),
    (6, 6, 300, 12, 3),
    (33, 2, 1000, 50, 16),
    (20, 8, 400, 20, 4),   # the reduce kernels are specialised per occupied count:
    (15, 1, 350, 14, 3),   # 8, 5 and 7 here, 1 and 0 take the generic kernel
    (21, 5, 450, 22, 7),
    (19, 7, 380, 19, 2),
]


def assert_close_modulus(out, ref, tol, keys):
    """|out - ref| / |ref| per complex entry: the Ozaki split bounds the error
    relative to the magnitude of each output (a real or imaginary part much
    smaller than the modulus, from cancellation, carries an absolute error of
    order 1e-12 |ref|, not a relative one)."""
    for k in keys:
        g = np.atleast_1d(np.asarray(out[k], np.complex128))
        r = np.atleast_1d(np.asarray(ref[k], np.complex128))
        err = float(np.max(np.abs(g - r) / np.maximum(np.abs(r), 1e-300)))
        assert err <= tol, f"{k}: relative error {err:.3e} > {tol:.1e}"


def make_arrays(seed0=100):
    out = []
    for i, (nb, nv, nc, ng, nf) in enumerate(SHAPES):
        a = (gpp_inputs.near_pole_inputs(nb, nc, ng, seed=seed0 + i) if nf == 3 and i % 2 == 0
             else gpp_inputs.seeded_inputs(nb, nc, ng, nfreq=nf, seed=seed0 + i, dw=2.0 / max(nf - 1, 1)))
        out.append((a, nv))
    return out


def device_batch(ctx, arrays):
    import pygpp
    dev = [{k: hiprt.DeviceArray(v) for k, v in a.items()} for a, _ in arrays]
    probs = [problem_of(a, nv) for a, nv in arrays]
    b = pygpp.Batch(ctx, probs, device_inputs=[{k: d.ptr for k, d in dd.items()} for dd in dev])
    return b, dev


def test_batch_vs_oracle(ctx, oracle):
    arrays = make_arrays()
    b, dev = device_batch(ctx, arrays)
    try:
        outs = b.sigma()
    finally:
        b.close()
    for (a, nv), out in zip(arrays, outs):
        assert_close(out, oracle.sigma(a, nvband=nv, threads=8), TOL, KEYS)


def test_batch_bitwise_equals_single_path(ctx):
    """The batch runs the Gram product and the pole reduce as separate
    stages, as the single path does by default (GPP_OPT_FUSED_REDUCE = 0):
    bitwise equal to it (the fused single path is compared in
    test_gpu_fused)."""
    import pygpp
    probs = [pygpp.Problem(*s) for s in SHAPES]
    seeds = [11 + i for i in range(len(probs))]
    with pygpp.Batch(ctx, probs, seeds=seeds) as b:
        outs = b.sigma()
        cfg = b.last_timing()
    with separate_reduce(ctx):
        for p, s, out in zip(probs, seeds, outs):
            ctx.generate_seeded(p, s)
            ref = ctx.sigma("factored-f64")
            assert cfg["gram_tile"] == ctx.get_option(pygpp.GPP_OPT_GRAM_TILE)
            for k in KEYS:
                np.testing.assert_array_equal(out[k], ref[k], err_msg=k)


@pytest.mark.parametrize("tile", [0, 3])
def test_batch_gram_tiles_bitwise(ctx, tile):
    """The batched FP64 Gram with the 128 x 128 (0) and 64 x 64 (3) tiles,
    with an instance of several hundred tiles: bitwise equal to the
    single-problem path with the same tile."""
    import pygpp
    probs = [pygpp.Problem(*s) for s in SHAPES] + [pygpp.Problem(7, 2, 30000, 300, 3)]
    seeds = [41 + i for i in range(len(probs))]
    ctx.set_option(pygpp.GPP_OPT_GRAM_TILE, tile)
    try:
        with pygpp.Batch(ctx, probs, seeds=seeds) as b:
            outs = b.sigma()
            assert b.last_timing()["gram_tile"] == tile
        with separate_reduce(ctx):
            for p, s, out in zip(probs, seeds, outs):
                ctx.generate_seeded(p, s)
                ref = ctx.sigma("factored-f64")
                for k in KEYS:
                    np.testing.assert_array_equal(out[k], ref[k], err_msg=k)
    finally:
        ctx.set_option(pygpp.GPP_OPT_GRAM_TILE, 0)


@pytest.mark.parametrize("form", ["factored-ozaki-i8", "factored-ozaki"])
def test_batch_ozaki_vs_oracle_and_single_path(ctx, oracle, form):
    """The Ozaki forms in a batch: every instance within 1e-10 of the oracle,
    relative to each output's modulus, and bitwise equal to the
    single-problem Ozaki path (same digit product, same reduce)."""
    import pygpp
    arrays = make_arrays(seed0=300)
    b, dev = device_batch(ctx, arrays)
    try:
        outs = b.sigma(form)
    finally:
        b.close()
    for (a, nv), out in zip(arrays, outs):
        assert_close_modulus(out, oracle.sigma(a, nvband=nv, threads=8), TOL, KEYS)
    probs = [pygpp.Problem(*s) for s in SHAPES]
    seeds = [31 + i for i in range(len(probs))]
    with pygpp.Batch(ctx, probs, seeds=seeds) as bs:
        outs = bs.sigma(form)
    for p, s, out in zip(probs, seeds, outs):
        ctx.generate_seeded(p, s)
        ref = ctx.sigma(form)
        for k in KEYS:
            np.testing.assert_array_equal(out[k], ref[k], err_msg=f"{form} {k}")


@pytest.mark.parametrize("form", ["factored-ozaki-i8", "factored-ozaki"])
def test_batch_ozaki_five_slices(ctx, oracle, form):
    """At 5 slices (base-254 digits; FP16 sums its level-3 products in two
    accumulators) the batch's digit product equals the single-problem one
    bitwise, and every instance is within 1e-10 of the oracle relative to
    each output's modulus."""
    import pygpp
    probs = [pygpp.Problem(*s) for s in SHAPES]
    seeds = [71 + i for i in range(len(probs))]
    try:
        ctx.set_option(pygpp.GPP_OPT_OZAKI_SLICES, 5)
        with pygpp.Batch(ctx, probs, seeds=seeds) as bs:
            outs = bs.sigma(form)
        for p, s, out in zip(probs, seeds, outs):
            ctx.generate_seeded(p, s)
            ref = ctx.sigma(form)
            for k in KEYS:
                np.testing.assert_array_equal(out[k], ref[k], err_msg=f"{form} {k}")
            assert_close_modulus(out, oracle.sigma(pygpp.fill_seeded(p, s), nvband=p.nvband, threads=8), TOL, KEYS)
    finally:
        ctx.set_option(pygpp.GPP_OPT_OZAKI_SLICES, 6)


def test_batch_switches_forms(ctx):
    """One batch evaluated in turn with the FP64, INT8 and FP16 Gram products
    (each re-plans the workspace and recaptures the graph) and back."""
    import pygpp
    probs = [pygpp.Problem(*s) for s in SHAPES[:4]]
    with pygpp.Batch(ctx, probs, seeds=[5, 6, 7, 8]) as b:
        first = b.sigma("factored-f64")
        i8 = b.sigma("factored-ozaki-i8")
        f16 = b.sigma("factored-ozaki")
        again = b.sigma("factored-f64")
    for x, y, z, w in zip(first, i8, f16, again):
        for k in KEYS:
            np.testing.assert_array_equal(x[k], w[k])
        assert_close_modulus(y, x, TOL, KEYS)
        assert_close_modulus(z, x, TOL, KEYS)


def test_graph_replays_match_eager(ctx):
    """Replays of the captured graph, with eager work and synchronisation in
    between, all equal the eager launch bitwise."""
    import pygpp
    probs = [pygpp.Problem(*s) for s in SHAPES]
    with pygpp.Batch(ctx, probs, seeds=list(range(50, 50 + len(probs)))) as b:
        b.use_graph(False)
        eager = b.sigma()
        b.use_graph(True)
        for rep in range(4):
            got = b.sigma()
            # unrelated eager work on the same stream between replays
            ctx.generate_seeded(pygpp.Problem(8, 2, 256, 16), rep)
            ctx.sigma("direct")
            for g, e in zip(got, eager):
                for k in KEYS:
                    np.testing.assert_array_equal(g[k], e[k])


def test_batch_reduce_freqs_bitwise(ctx, oracle):
    """The batched reduce in passes of 4, 8 or 16 frequencies (16: passes of
    16, then 8 and 4; the negative-frequency class at most 8) regroups the
    same per-frequency sums: bitwise equal for every pass width, within 1e-10
    of the oracle."""
    import pygpp
    arrays = make_arrays(seed0=500)
    arrays.append((gpp_inputs.seeded_inputs(26, 900, 30, nfreq=28, seed=512, dw=2.0 / 27), 4))
    arrays.append((gpp_inputs.seeded_inputs(22, 700, 25, nfreq=24, seed=513, dw=2.0 / 23), 2))
    neg = gpp_inputs.seeded_inputs(10, 500, 25, nfreq=20, seed=514, wx_center=0.0, dw=0.5)
    arrays.append((neg, 2))
    base = None
    try:
        for fc in (4, 8, 16):
            ctx.set_option(pygpp.GPP_OPT_REDUCE_FREQS, fc)
            b, dev = device_batch(ctx, arrays)
            try:
                outs = b.sigma()
                assert b.last_timing()["reduce_freqs"] == fc
            finally:
                b.close()
            if base is None:
                base = outs
                for (a, nv), out in zip(arrays, outs):
                    assert_close(out, oracle.sigma(a, nvband=nv, threads=8), TOL, KEYS)
            for i, (out, ref) in enumerate(zip(outs, base)):
                for k in KEYS:
                    np.testing.assert_array_equal(out[k], ref[k], err_msg=f"fc {fc} instance {i} {k}")
    finally:
        ctx.set_option(pygpp.GPP_OPT_REDUCE_FREQS, 0)


def test_batch_autotune_keeps_accuracy(ctx, oracle):
    import pygpp
    arrays = make_arrays(seed0=300)
    b, dev = device_batch(ctx, arrays)
    try:
        b.autotune()
        outs = b.sigma()
        cfg = b.last_timing()
    finally:
        b.close()
    assert 0 <= cfg["gram_tile"] <= 3 and cfg["reduce_freqs"] in (4, 8)
    for (a, nv), out in zip(arrays, outs):
        assert_close(out, oracle.sigma(a, nvband=nv, threads=8), TOL, KEYS)


def test_batch_rejects_unsupported(ctx):
    import pygpp
    with pygpp.Batch(ctx, [pygpp.Problem(8, 2, 200, 10)], seeds=[1]) as b:
        with pytest.raises(pygpp.GppError) as e:
            b.sigma("direct")
        assert e.value.status == pygpp.GPP_ERR_UNSUPPORTED


def test_ctx_autotune(ctx, oracle):
    import pygpp
    arrs = gpp_inputs.seeded_inputs(30, 3000, 200, seed=7)
    ctx.upload(problem_of(arrs, 2), arrs)
    try:
        cfg = ctx.autotune("factored-f64")
        assert cfg["gram_tile"] in range(4)
        assert_close(ctx.sigma("factored-f64"), oracle.sigma(arrs, nvband=2, threads=8), TOL, KEYS)
    finally:
        ctx.set_option(pygpp.GPP_OPT_GRAM_TILE, 0)
        ctx.set_option(pygpp.GPP_OPT_REDUCE_CHUNK, 8192)
        ctx.set_option(pygpp.GPP_OPT_REDUCE_FREQS, 0)


@pytest.mark.parametrize("form", ["factored-f64", "factored-ozaki-i8"])
def test_batch_negative_frequency_instances(ctx, oracle, form):
    """Instances with negative frequencies (the ssx cutoff of
    gppComplex.cpp:58-59 live, cutoff-straddling elements included) go to the
    batch's negative-frequency reduce class beside ordinary instances: each
    matches the oracle and, bitwise, the single-problem path."""
    cut = gpp_inputs.ssx_cutoff_inputs(14, 900, 40, seed=9)
    cut.pop("_kind")
    neg4 = gpp_inputs.seeded_inputs(12, 700, 30, nfreq=4, seed=10)
    neg4["wx_array"] = np.array([-2.0, -0.5, 1.0, 3.0])
    neg9 = gpp_inputs.seeded_inputs(10, 500, 25, nfreq=9, seed=11, wx_center=0.0, dw=0.5)
    arrays = [(cut, 4), (gpp_inputs.seeded_inputs(16, 800, 40, seed=12), 2), (neg4, 3), (neg9, 2)]
    b, dev = device_batch(ctx, arrays)
    try:
        outs = b.sigma(form)
    finally:
        b.close()
    for (a, nv), out in zip(arrays, outs):
        ref = oracle.sigma(a, nvband=nv, threads=8)
        if form == "factored-f64":
            assert_close(out, ref, TOL, KEYS)
        else:
            assert_close_modulus(out, ref, TOL, KEYS)
        ctx.upload(problem_of(a, nv), a)
        with separate_reduce(ctx):
            single = ctx.sigma(form)
        for k in KEYS:
            np.testing.assert_array_equal(out[k], single[k], err_msg=k)
