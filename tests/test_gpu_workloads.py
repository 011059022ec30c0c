"""The benchmark workloads at their real sizes, on the GPU, against the CPU
oracle (the reference's loop restated in C, bitwise equal to the compiled
reference on every fixture):

* the benchmark shape itself (512 bands, nvband 2, ncouls 32768, ngpown 1638,
  nfreq 3, seed 1): every form against the oracle run over all 8.2e10
  updates, all four outputs -- the direct forms within 1e-12 relative per
  component, the factored forms within 1e-10.  A band's acht_n1_loc sums 5e7
  terms left to right in the reference, whose own rounding then moves some
  components by up to a few 1e-12 (the direct-strict form, whose terms are
  the reference's bit for bit, differs from it by that much); there the
  direct forms are held to 1e-12 of the compensated oracle (the exact sum of
  the reference's terms), or to no more than the reference's own distance
  from it (assert_close_conditioned);
* the batched workload (64 instances, nb up to 1024, ncouls up to 16384,
  nfreq up to 32) through one batch with the FP64 and the INT8 Ozaki Gram
  product: every instance against the direct form run on that instance, the
  smallest instances against the oracle;
* single problems with 24 and 32 frequencies (three and four passes of the
  reduce) for every form.

Tolerances: the FP64 forms are compared per real / imaginary component
(a component below 1e-6 of its modulus relative to the modulus).  The Ozaki
forms' error is normwise -- the digit split bounds each Gram element's error
relative to its column scales, so an output's error is bounded relative to
its modulus, not to a real or imaginary part that cancellation made small --
and they are held to 1e-10 of each output's modulus (at the benchmark shape
they also meet 1e-10 per component, checked below).
"""
import os

import numpy as np
import pytest

import gpp_inputs
from test_gpu_batch import assert_close_modulus
from test_gpu_parity import assert_close, assert_close_conditioned, comp_rel_err, problem_of, run_gpu

pytestmark = pytest.mark.gpu

TOL_DIRECT = 1e-12
TOL_FACTORED = 1e-10
ALL = ("achtemp", "asxtemp", "acht_n1_loc", "achstemp")


def host_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


@pytest.fixture(scope="module")
def c1_full(oracle):
    """The benchmark shape's inputs (seed 1) and the oracle's outputs over the
    whole loop, sequential and compensated, computed once on the host's cores
    (about 30 s and 60 s at 16)."""
    import pygpp
    p = pygpp.Problem(512, 2, 32768, 1638)
    print(f"c1_full: oracle over {p.updates:.3e} updates on {host_threads()} threads", flush=True)
    arrs = pygpp.fill_seeded(p, seed=1)
    ref = oracle.sigma(arrs, nvband=2, threads=host_threads())
    print("c1_full: compensated oracle", flush=True)
    dd = oracle.sigma_dd(arrs, nvband=2, threads=host_threads(), all_outputs=True)
    return p, arrs, ref, dd


@pytest.mark.parametrize("form", ["direct", "direct-strict", "factored-f64", "factored-ozaki-i8", "factored-ozaki"])
def test_c1_full_vs_cpu_oracle(ctx, c1_full, form):
    p, arrs, ref, dd = c1_full
    ctx.generate_seeded(p, seed=1)  # the device generator gives the host generator's bits
    out = ctx.sigma(form, acht_n1_loc=True)
    if form.startswith("direct"):
        assert_close(out, ref, TOL_DIRECT, ("achtemp", "asxtemp", "achstemp"))
        assert_close_conditioned(out, ref, dd, TOL_DIRECT, ALL)
    else:
        assert_close(out, ref, TOL_FACTORED, ALL)
    print(f"{form}: worst per-component error vs sequential / compensated oracle: " +
          ", ".join(f"{k} {comp_rel_err(out[k], ref[k]):.2e} / {comp_rel_err(out[k], dd[k]):.2e}" for k in ALL),
          flush=True)


def test_c1_full_slabs_sum_to_whole(ctx, c1_full):
    """The single-GPU part of the reference's G'-column decomposition
    (gppMPIOpenMP3.cpp:172, 255-271, 361-370): eight slabs of the benchmark
    shape (204-205 columns each), evaluated one at a time and summed on the
    host in rank order, equal the whole evaluation and the oracle."""
    p, arrs, ref, dd = c1_full
    ctx.generate_seeded(p, seed=1)
    nr = 8
    bounds = [r * p.ngpown // nr for r in range(nr + 1)]
    for form, tol in (("direct", TOL_DIRECT), ("factored-f64", TOL_FACTORED), ("factored-ozaki-i8", TOL_FACTORED)):
        whole = ctx.sigma(form, acht_n1_loc=True)
        total = None
        for r in range(nr):
            ctx.set_columns(bounds[r], bounds[r + 1])
            part = ctx.sigma(form, acht_n1_loc=True)
            if total is None:
                total = {k: np.array(part[k]) for k in ALL}
            else:
                for k in ALL:
                    total[k] = total[k] + part[k]
        ctx.set_columns(0, 0)
        assert_close(total, whole, tol, ALL)
        assert_close_conditioned(total, ref, dd, tol, ALL)


C4_FORMS = ["factored-f64", "factored-ozaki-i8"]


@pytest.fixture(scope="module")
def c4_batch_outputs(ctx):
    import gpp_workloads
    import pygpp
    inst = gpp_workloads.c4_instances()
    probs = [q[0] for q in inst]
    outs = {}
    with pygpp.Batch(ctx, probs, seeds=[q[1] for q in inst], wx_center=[q[2] for q in inst],
                     dw=[q[3] for q in inst]) as b:
        for form in C4_FORMS:
            outs[form] = b.sigma(form)
            print(f"c4 {form}: {b.last_timing()['ms']:.2f} ms", flush=True)
    return inst, outs


def test_c4_batch_vs_direct_every_instance(ctx, c4_batch_outputs):
    inst, outs = c4_batch_outputs
    keys = ("achtemp", "asxtemp", "achstemp")
    worst = {f: 0.0 for f in C4_FORMS}
    for i, (p, seed, wxc, dw) in enumerate(inst):
        ctx.generate_seeded(p, seed, wxc, dw)
        ref = ctx.sigma("direct")
        assert_close(outs["factored-f64"][i], ref, TOL_FACTORED, keys)
        assert_close_modulus(outs["factored-ozaki-i8"][i], ref, TOL_FACTORED, keys)
        for f in C4_FORMS:
            worst[f] = max(worst[f], max(comp_rel_err(outs[f][i][k], ref[k]) for k in keys))
    print("c4 worst per-component error vs direct: " + ", ".join(f"{f} {w:.2e}" for f, w in worst.items()),
          flush=True)


def test_c4_smallest_instances_vs_oracle(ctx, oracle, c4_batch_outputs):
    import pygpp
    inst, outs = c4_batch_outputs
    order = sorted(range(len(inst)), key=lambda i: inst[i][0].updates)[:4]
    keys = ("achtemp", "asxtemp", "achstemp")
    for i in order:
        p, seed, wxc, dw = inst[i]
        arrs = pygpp.fill_seeded(p, seed, wxc, dw)
        ref = oracle.sigma(arrs, nvband=p.nvband, threads=host_threads())
        dd = oracle.sigma_dd(arrs, nvband=p.nvband, threads=host_threads(), all_outputs=True)
        assert_close(outs["factored-f64"][i], ref, TOL_FACTORED, keys)
        assert_close_modulus(outs["factored-ozaki-i8"][i], ref, TOL_FACTORED, keys)
        ctx.generate_seeded(p, seed, wxc, dw)
        assert_close_conditioned(ctx.sigma("direct"), ref, dd, TOL_DIRECT, ALL)


WIDE = [  # nb, nv, ncouls, ngpown, nfreq
    (12, 3, 2000, 60, 32),
    (10, 2, 1500, 50, 24),
    (9, 8, 900, 30, 32),     # the 8-occupied-band reduce kernel
    (11, 5, 700, 25, 27),    # generic occupied count, ragged passes
]


@pytest.mark.parametrize("shape", WIDE, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("form", ["direct", "direct-strict", "factored-f64", "factored-ozaki", "factored-ozaki-i8"])
def test_many_frequencies_vs_oracle(ctx, oracle, shape, form):
    nb, nv, ncouls, ngpown, nfreq = shape
    arrs = gpp_inputs.seeded_inputs(nb, ncouls, ngpown, nfreq=nfreq, seed=sum(shape), dw=2.0 / (nfreq - 1))
    ref = oracle.sigma(arrs, nvband=nv, threads=8)
    out = run_gpu(ctx, arrs, nv, form)
    if form.startswith("direct"):
        dd = oracle.sigma_dd(arrs, nvband=nv, threads=8, all_outputs=True)
        assert_close_conditioned(out, ref, dd, TOL_DIRECT, ALL)
    else:
        assert_close(out, ref, TOL_FACTORED, ("achtemp", "asxtemp", "achstemp"))
        assert_close(ctx.sigma(form, acht_n1_loc=True), ref, TOL_FACTORED, ALL)


@pytest.mark.parametrize("form", ["factored-f64", "factored-ozaki-i8"])
def test_many_frequencies_batched(ctx, oracle, form):
    """Instances of 24-32 frequencies in one batch, against the oracle."""
    arrays = []
    for i, (nb, nv, ncouls, ngpown, nfreq) in enumerate(WIDE):
        a = gpp_inputs.seeded_inputs(nb, ncouls, ngpown, nfreq=nfreq, seed=100 + i, dw=2.0 / (nfreq - 1))
        arrays.append((a, nv))
    from test_gpu_batch import device_batch
    b, dev = device_batch(ctx, arrays)
    try:
        outs = b.sigma(form)
    finally:
        b.close()
    for (a, nv), out in zip(arrays, outs):
        assert_close(out, oracle.sigma(a, nvband=nv, threads=8), TOL_FACTORED, ("achtemp", "asxtemp", "achstemp"))
