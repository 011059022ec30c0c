#!/usr/bin/env python3
"""GPP sigma throughput on one MI355X.

One step is one evaluation of the GPP sigma band loop (gppComplex.cpp
:335-383) over the benchmark problem resident in HBM: number_bands=512,
nvband=2, ncouls=32768, ngpown=1638, nfreq=3, seeded complex128 inputs
(generated on the device).  The headline step computes achtemp, asxtemp and
achstemp (the outputs the metric names); "factored-f64+acht" in `forms` is
the same step with acht_n1_loc as well, every output of the reference's
loop.  The metric counts one update per (n1, igp, ig, iw) tuple.

Prints ONE JSON line.  `value` is the updates/s of the headline form; the
other forms' numbers, their errors, per-stage GPU times (HIP events on the
library's stream), the roofline of the dominant kernel, one G'-column slab's
time and the reference's OpenMP CPU path timed on this host are attached.
The GPU work runs first; the CPU leg then times the reference and, at the
benchmark shape, runs the CPU oracle over the whole problem as the checker
of every form's outputs (``rel_err_vs_cpu_oracle``).

    python bench.py [--steps K] [--warmup W] [--form direct|...] [--no-cpu] [--no-cpu-parity]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "doe-csfg-hackathon_amd")
ORACLE = os.path.join(REPO, "oracle")
sys.path.insert(0, PKG)

METRIC = "GPP sigma updates/s (bands\u00b7ngpown\u00b7ncouls\u00b7nfreq), direct vs factored @\u22641e-10 rel err"
# Workloads: c1 = the benchmark shape; c3 = the large single instance.
WORKLOADS = {
    "c1": dict(nb=512, nv=2, ncouls=32768, ngpown=1638, nfreq=3, wxc=4.0, dw=1.0),
    "c3": dict(nb=1536, nv=8, ncouls=65536, ngpown=3276, nfreq=16, wxc=4.0, dw=0.25),
    # batched: the CPU sample below uses a mid-size instance shape
    "c4": dict(nb=576, nv=5, ncouls=16384, ngpown=819, nfreq=3, wxc=4.0, dw=1.0),
}
NB, NV, NCOULS, NGPOWN, NFREQ, SEED = 512, 2, 32768, 1638, 3, 1

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 vector and FP64 matrix, dense
# The FP64 MFMA products form each complex product from three real products
# (3M), 6 flop per complex multiply-add executed (8 in the 4-product count).
F64_FLOP_PER_CMAC = 6
PEAK_HBM_GBS = 8000.0
PEAK_FP16_TFLOPS = 2500.0 # dense FP16 MFMA
PEAK_INT8_TOPS = 5000.0   # dense INT8 MFMA
FLOP_UNOCC = 33           # per unoccupied update, counted on the reference's expression
FLOP_OCC = 62             # per occupied update (branch-1 path)


def host_cores():
    """Cores this process may run on (its affinity mask)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


# The CPU legs use at most 16 threads: a GPU box gives one GPU's job a share
# of 16 host cores, whatever the machine's core count and affinity show.
CPU_THREAD_CAP = 16


def cpu_cores():
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return min(CPU_THREAD_CAP, host_cores())


def cpu_core_note(cores):
    return {"host_cores_available": host_cores(), "threads_used": cores,
            "thread_cap": CPU_THREAD_CAP,
            "cap_reason": "one GPU's share of the host is 16 cores (OMP_NUM_THREADS overrides)"}


def cpu_baseline():
    """The reference's OpenMP build (gppOpenMP3.cpp, compiled from source into
    oracle/_ref) on a bounded band sample of the benchmark shape; falls back
    to the oracle's race-free OpenMP restatement when the reference binary is
    absent.  Runs before the GPU is initialised."""
    cores = cpu_cores()
    # about 5e10 updates (10 s on 16 host cores): 305 bands at c1, 77 at c3
    nb_s = min(NB, max(16, int(5e10 / (NGPOWN * NCOULS * 3))))
    nv_s = int(round(NV * nb_s / NB))  # keep the occupied-band fraction of the full problem
    ref_exe = os.path.join(ORACLE, "_ref", "gppOpenMP3.ex")
    # the reference fixes nfreq = 3 (gppComplex.cpp:12-13); its rate per update is what is compared
    updates = nb_s * NGPOWN * NCOULS * 3
    if os.path.exists(ref_exe):
        env = dict(os.environ, OMP_NUM_THREADS=str(cores))
        r = subprocess.run([ref_exe, str(nb_s), str(nv_s), str(NCOULS), "20"], capture_output=True, text=True,
                           env=env, timeout=600)
        m = re.search(r"Time Taken \*+= ([0-9.eE+-]+)", r.stdout)
        if r.returncode == 0 and m:
            secs = float(m.group(1))
            return {"value": updates / secs, "unit": "updates/s", "cores": cores, "kind": "reference",
                    "sample": f"reference gppOpenMP3 (g++ -O3 -fopenmp, {cores} threads) on '{nb_s} {nv_s} {NCOULS} 20': "
                              f"{nb_s} of {NB} bands, ngpown={NCOULS // 20}, nfreq=3, its deterministic init; "
                              f"{updates:.3e} updates in {secs:.3f} s (its own loop timer)",
                    "seconds": secs, "cores_detail": cpu_core_note(cores)}
    sys.path.insert(0, ORACLE)
    import gpp_oracle
    import pygpp
    p = pygpp.Problem(nb_s, nv_s, NCOULS, NGPOWN, 3)
    arrs = pygpp.fill_seeded(p, SEED)
    t0 = time.perf_counter()
    gpp_oracle.sigma(arrs, nvband=nv_s, threads=cores)
    secs = time.perf_counter() - t0
    return {"value": updates / secs, "unit": "updates/s", "cores": cores, "kind": "port",
            "sample": f"oracle restatement of gppComplex.cpp (race-free OpenMP, {cores} threads) on {nb_s} of {NB} "
                      f"bands at ncouls={NCOULS}, ngpown={NGPOWN}, nfreq=3, seeded; {updates:.3e} updates in {secs:.3f} s",
            "seconds": secs, "cores_detail": cpu_core_note(cores)}


def cpu_parity(results):
    """The checker half of the CPU leg: the CPU oracle (the reference's loop
    restated in C, bitwise equal to the compiled reference on its fixtures)
    over the whole benchmark problem on the host's cores, and every GPU
    form's largest per-component relative error against it."""
    sys.path.insert(0, ORACLE)
    import gpp_oracle
    import pygpp
    cores = cpu_cores()
    p = pygpp.Problem(NB, NV, NCOULS, NGPOWN, NFREQ)
    arrs = pygpp.fill_seeded(p, SEED)  # the device generator's bits
    t0 = time.perf_counter()
    ref = gpp_oracle.sigma(arrs, nvband=NV, threads=cores)
    secs = time.perf_counter() - t0
    t0 = time.perf_counter()
    dd = gpp_oracle.sigma_dd(arrs, nvband=NV, threads=cores, all_outputs=True)
    secs_dd = time.perf_counter() - t0
    del arrs
    errs, errs_dd = {}, {}
    for form, out in results.items():
        keys = ("achtemp", "asxtemp", "achstemp") + (("acht_n1_loc",) if out.get("acht_n1_loc") is not None else ())
        errs[form] = {k: rel_err(out[k], ref[k]) for k in keys}
        errs_dd[form] = {k: rel_err(out[k], dd[k]) for k in keys}
    return {"reference": f"CPU oracle (gpp_oracle.c, race-free OpenMP, {cores} threads) over all "
                         f"{p.updates:.3e} updates of the benchmark shape, seed {SEED}: its sequential sums "
                         f"(the reference's rounding, bitwise) and its compensated sums (the exact sum of the "
                         f"reference's terms)",
            "seconds": secs, "seconds_compensated": secs_dd,
            "oracle_own_summation_error": {k: rel_err(ref[k], dd[k]) for k in dd},
            "rel_err_per_component": errs, "rel_err_per_component_compensated": errs_dd}


def run_c4(args):
    import numpy as np
    import pygpp
    from gpp_workloads import c4_instances
    inst = c4_instances()
    probs = [p for p, _, _, _ in inst]
    updates = sum(p.updates for p in probs)
    ctx = pygpp.Context(0)
    batch = pygpp.Batch(ctx, probs, seeds=[s for _, s, _, _ in inst], wx_center=[w for _, _, w, _ in inst],
                        dw=[d for _, _, _, d in inst])
    t0 = time.perf_counter()
    batch.autotune()
    tune_s = time.perf_counter() - t0
    batch.use_graph(False)
    batch.sigma()
    stages = batch.last_timing()["stages_ms"]
    batch.use_graph(True)
    for _ in range(args.warmup):
        batch.sigma()
    gpu_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        outs = batch.sigma()
        gpu_ms.append(batch.last_timing()["ms"])
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    cfg = batch.last_timing()
    # the same batch with the Ozaki INT8 and FP16 Gram products (autotuned reduce)
    oz = {}
    for form in ("factored-ozaki-i8", "factored-ozaki"):
        batch.autotune(form)
        batch.use_graph(False)
        batch.sigma(form)
        st = batch.last_timing()["stages_ms"]
        batch.use_graph(True)
        for _ in range(args.warmup):
            batch.sigma(form)
        t1 = time.perf_counter()
        for _ in range(args.steps):
            o = batch.sigma(form)
        oms = (time.perf_counter() - t1) * 1e3 / args.steps
        oz[form] = {"ms_per_step": oms, "updates_per_s": updates / (oms * 1e-3), "stages_ms_eager": st,
                    "outs": o}
    # per-instance FP64 references (direct kernel) and the one-at-a-time factored path
    errs, single_ms, direct_ms = [], 0.0, 0.0
    oz_errs = {f: [] for f in oz}
    for i, ((p, seed, wxc, dw), out) in enumerate(zip(inst, outs)):
        ctx.generate_seeded(p, seed, wxc, dw)
        ref = ctx.sigma("direct")
        direct_ms += ctx.last_timing()["total"]
        ctx.sigma("factored-f64")
        single_ms += ctx.last_timing()["total"]
        errs.append({k: rel_err(out[k], ref[k]) for k in ("achtemp", "asxtemp", "achstemp")})
        for f in oz:
            oz_errs[f].append({k: rel_err(oz[f]["outs"][i][k], ref[k]) for k in ("achtemp", "asxtemp", "achstemp")})
            oz_errs[f][-1].update({k + "_modulus": rel_err_modulus(oz[f]["outs"][i][k], ref[k])
                                   for k in ("achtemp", "asxtemp", "achstemp")})
    for f in oz:
        del oz[f]["outs"]
        oz[f]["max_rel_err_vs_direct_per_output"] = {k: max(e[k] for e in oz_errs[f]) for k in oz_errs[f][0]}
        oz[f]["speedup_vs_direct"] = direct_ms / oz[f]["ms_per_step"]
    batch.close()
    ctx.close()
    cpu = None if args.no_cpu else cpu_baseline()  # after the GPU work
    # executed work of the 3M complex product: three real products (6 flop)
    # per complex multiply-add; the 4-product (ZGEMM) count is 8
    gram_flop = sum(F64_FLOP_PER_CMAC * p.ngpown * p.ncouls * (p.number_bands - p.nvband) for p in probs)
    achieved = gram_flop / (stages["gram"] * 1e-3) / 1e12
    line = {
        "metric": METRIC, "value": updates / (ms * 1e-3), "unit": "updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: 64 seeded instances (counter-based SplitMix64, seeds 1000..1063) generated in HBM",
        "config": {"workload": "GPP sigma factored-f64, batched: 64 instances, nb~U{128..1024}, nv~U{2..8}, "
                               "ncouls~U{2048..16384}, ngpown=ncouls/20, nfreq~U{3..32}; persistent kernels "
                               "replayed from a hipGraph, autotuned",
                   "form": "factored-f64", "instances": len(probs), "updates_per_step": updates,
                   "autotuned": {k: cfg[k] for k in ("gram_tile", "reduce_chunk", "reduce_freqs")},
                   "autotune_seconds": tune_s},
        "roofline": {"bound": "mfma", "kernel": "k_batch_gram (v_mfma_f64_16x16x4, persistent)",
                     "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None, "algorithmic_flop_per_launch": gram_flop,
                     "flop_count": "3M: 6 flop per complex multiply-add executed on MFMA",
                     "zgemm_equivalent_tflops": achieved * 8 / F64_FLOP_PER_CMAC,
                     "avg_launch_ms": stages["gram"], "stages_ms_eager": stages},
        "cpu_baseline": cpu,
        "batch": {"graph_gpu_ms_avg": sum(gpu_ms) / len(gpu_ms),
                  "max_rel_err_vs_direct_per_output": {k: max(e[k] for e in errs) for k in errs[0]},
                  "median_rel_err_vs_direct": float(np.median([max(e.values()) for e in errs])),
                  "one_at_a_time_factored_gpu_ms": single_ms,
                  "direct_gpu_ms": direct_ms,
                  "speedup_vs_direct": direct_ms / ms,
                  "speedup_vs_one_at_a_time": single_ms / ms,
                  "ozaki_forms": oz},
    }
    if cpu:
        line["gpu_over_cpu"] = line["value"] / cpu["value"]
    print(json.dumps(line), flush=True)


def load_traffic(kernel):
    path = os.path.join(REPO, "profiles", f"traffic_{kernel}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def rel_err_modulus(a, b):
    """max |a - b| / |b| over the complex entries (normwise per output entry)."""
    import numpy as np
    a = np.atleast_1d(np.asarray(a, np.complex128))
    b = np.atleast_1d(np.asarray(b, np.complex128))
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def rel_err(a, b):
    import numpy as np
    a = np.atleast_1d(np.asarray(a, np.complex128))
    b = np.atleast_1d(np.asarray(b, np.complex128))
    w = 0.0
    for x, y in zip(a, b):
        for u, v in ((x.real, y.real), (x.imag, y.imag)):
            w = max(w, abs(u - v) / max(abs(v), 1e-6 * abs(y), 1e-300))
    return w


def time_form(ctx, form, steps, warmup, acht=None):
    """K evaluations enqueued back to back on the library's stream, then one
    wait that copies the last one's outputs to the host; the warmup ends
    with a synchronous evaluation, so the host clock brackets exactly the K
    evaluations.  Stage times are the HIP-event averages over the K."""
    for _ in range(warmup):
        ctx.sigma(form, acht_n1_loc=acht)
    want_acht = bool(acht) or form.startswith("direct")
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.sigma_async(form, acht_n1_loc=bool(acht))
    out = ctx.wait(acht_n1_loc=want_acht)
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / steps
    stages = ctx.timing_summary(steps)
    return out, ms, stages


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps of the headline form (default: 50 at c1, 10 at c3 and c4)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps before them (default: 10 at c1, 3 at c3 and c4)")
    ap.add_argument("--form", default="factored-f64",
                    choices=["direct", "direct-strict", "factored-f64", "factored-ozaki", "factored-ozaki-i8"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU leg (baseline and full-size parity)")
    ap.add_argument("--no-cpu-parity", action="store_true",
                    help="skip the CPU oracle over the whole benchmark problem (about 30 s on 16 cores)")
    ap.add_argument("--workload", default="c1", choices=sorted(WORKLOADS))
    args = ap.parse_args()
    # A few milliseconds of GPU work after an idle host wait run at a lower
    # shader clock than the sustained rate (512/2/32768/1638 shape, FP64
    # factored form: 5 steps after one warmup average 3.54-3.58 ms, 50 after
    # ten 3.28 ms, 200 after twenty 3.27 ms), so that short step gets more
    # warmup and timed steps.
    if args.steps is None:
        args.steps = 50 if args.workload == "c1" else 10
    if args.warmup is None:
        args.warmup = 10 if args.workload == "c1" else 3
    global NB, NV, NCOULS, NGPOWN, NFREQ
    wl = WORKLOADS[args.workload]
    NB, NV, NCOULS, NGPOWN, NFREQ = wl["nb"], wl["nv"], wl["ncouls"], wl["ngpown"], wl["nfreq"]
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE={world}); "
              "multi-GPU sharding of the G' columns is a design note, not run here.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "c4":
        run_c4(args)
        return

    import pygpp
    prob = pygpp.Problem(NB, NV, NCOULS, NGPOWN, NFREQ)
    updates = prob.updates
    ctx = pygpp.Context(0)
    ctx.generate_seeded(prob, SEED, wl["wxc"], wl["dw"])

    others = ("factored-f64", "factored-ozaki-i8", "factored-ozaki", "direct", "direct-strict")
    if args.workload != "c1":
        others = ("factored-f64", "factored-ozaki-i8", "direct")  # the strict form would take ~7 s per step here
    run_forms = [args.form] + [f for f in others if f != args.form]
    forms = {}
    results = {}
    # The factored forms compute acht_n1_loc only on request (a second product
    # of the Gram product's size); "factored-f64+acht" is that run, with every
    # output of the reference's band loop.
    # "factored-f64-fused" is factored-f64 with the pole reduce in the Gram
    # epilogue (GPP_OPT_FUSED_REDUCE = 1, at most three frequencies): no G'xG
    # workspace, slower where that fits.
    # "...@5": the Ozaki form at 5 slices (39-bit digits) instead of the
    # default 6 (47 bits for INT8, 42 for FP16): 15 digit pairs instead of 21.
    # "factored-ozaki-i8@5+acht": the fastest form with every output (its
    # acht_n1_loc product runs on FP64 MFMA, as in factored-f64+acht).
    extra = ["factored-f64+acht"] + (["factored-f64-fused"] if NFREQ <= 3 else [])
    extra += ["factored-ozaki-i8@5"] + (["factored-ozaki@5", "factored-ozaki-i8@5+acht"] if args.workload == "c1" else [])
    for form in run_forms + extra:
        name, acht = (form[:-5], True) if form.endswith("+acht") else (form, None)
        fused = name.endswith("-fused")
        if fused:
            name = name[:-len("-fused")]
        slices = 5 if name.endswith("@5") else 6
        name = name[:-2] if name.endswith("@5") else name
        # the direct forms (50 ms per step at c1, 3 s at c3) run at their
        # sustained clock from the first step: a few steps and one warmup
        if form == args.form:
            steps, warm = args.steps, args.warmup
        elif name.startswith("direct"):
            steps, warm = max(1, min(args.steps // 2, 5 if args.workload == "c1" else 2)), min(args.warmup, 1)
        else:
            steps, warm = max(1, args.steps // 2), args.warmup
        ctx.set_option(pygpp.GPP_OPT_FUSED_REDUCE, 1 if fused else 0)
        ctx.set_option(pygpp.GPP_OPT_OZAKI_SLICES, slices)
        out, ms, stages = time_form(ctx, name, steps, warm, acht)
        ctx.set_option(pygpp.GPP_OPT_OZAKI_SLICES, 6)
        ctx.set_option(pygpp.GPP_OPT_FUSED_REDUCE, 0)
        results[form] = out
        forms[form] = {"ms_per_step": ms, "updates_per_s": updates / (ms * 1e-3), "gpu_stage_ms": stages}
    head = forms[args.form]
    # One G'-column slab at its real shape: the share of one of 8 ranks in the
    # reference's MPI decomposition (gppMPIOpenMP3.cpp:172, 255-271).
    slab_cols = NGPOWN // 8
    ctx.set_columns(0, slab_cols)
    _, slab_ms, slab_st = time_form(ctx, args.form, max(1, args.steps // 2), args.warmup)
    ctx.set_columns(0, 0)
    slab = {"columns": [0, slab_cols], "of": NGPOWN, "form": args.form, "ms_per_step": slab_ms,
            "updates_per_s": NB * slab_cols * NCOULS * NFREQ / (slab_ms * 1e-3), "gpu_stage_ms": slab_st}
    # Agreement with the strict direct form (reference operation order, FP64),
    # per output: largest per-component relative error.  The CPU leg below
    # adds every form's error against the CPU oracle over the whole problem.
    bname = "direct-strict" if "direct-strict" in results else "direct"
    base = results[bname]
    for form, out in results.items():
        keys = ("achtemp", "asxtemp", "achstemp") + (("acht_n1_loc",) if out.get("acht_n1_loc") is not None else ())
        forms[form]["rel_err_vs_" + bname.replace("-", "_") + "_gpu"] = {k: rel_err(out[k], base[k]) for k in keys}
        if form.startswith("factored-ozaki"):
            forms[form]["rel_err_modulus_vs_" + bname.replace("-", "_") + "_gpu"] = {
                k: rel_err_modulus(out[k], base[k]) for k in keys}
    for f in forms:
        if f.startswith("factored"):
            forms[f]["speedup_vs_direct"] = forms["direct"]["ms_per_step"] / forms[f]["ms_per_step"]

    def direct_roofline(form):
        kname = "direct_unoccupied_strict" if form == "direct-strict" else "direct_unoccupied"
        kms = forms[form]["gpu_stage_ms"][kname]
        flop = (NB - NV) * NGPOWN * NCOULS * NFREQ * FLOP_UNOCC
        achieved = flop / (kms * 1e-3) / 1e12
        tr = load_traffic("k_unocc") if (form == "direct" and args.workload == "c1") else None
        return {"bound": "valu", "pipe": "FP64 VALU (FP64 vector peak = FP64 matrix peak on MI355X)",
                "kernel": kname, "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64_TFLOPS,
                "traffic": (tr["fetch_bytes"] + tr["write_bytes"]) if tr else None,
                "algorithmic_flop_per_launch": flop, "avg_launch_ms": kms}

    def factored_roofline(form):
        st = forms[form]["gpu_stage_ms"]
        gkey = [k for k in st if k.startswith("factored_gram")][0]
        gms = st[gkey]
        K = NB - NV
        fused = gkey == "factored_gram_pole"
        if form.startswith("factored-f64"):
            flop = F64_FLOP_PER_CMAC * NGPOWN * NCOULS * K
            peak, pname = PEAK_FP64_TFLOPS, "v_mfma_f64_16x16x4, 3M: 6 flop per complex multiply-add"
            if fused:
                pname += "; the pole reduce of its elements in the epilogue, timed with it"
            tr = load_traffic("k_gram_pole" if fused else "k_gram_f64") if args.workload == "c1" else None
        else:
            # executed Ozaki work: S (S + 1) / 2 digit pairs x 4 real products, split included in the
            # stage time
            S = 5 if form.endswith("@5") else 6
            flop = S * (S + 1) // 2 * 4 * 2 * NGPOWN * NCOULS * K
            i8 = "-i8" in form
            peak = PEAK_INT8_TOPS if i8 else PEAK_FP16_TFLOPS
            # the default block layout (4, persistent layout 3) runs 32x32 MFMAs
            pname = ("v_mfma_i32_32x32x32_i8" if i8 else "v_mfma_f32_32x32x16_f16") + ", digit splits included"
            tr = None
        achieved = flop / (gms * 1e-3) / 1e12
        extra = {"zgemm_equivalent_tflops": achieved * 8 / F64_FLOP_PER_CMAC} if form.startswith("factored-f64") else {}
        line = {**extra, "bound": "mfma", "kernel": f"{gkey} ({pname})", "achieved": achieved,
                "peak": peak, "unit": "TOP/s" if "-i8" in form else "TFLOP/s", "frac": achieved / peak,
                "traffic": (tr["fetch_bytes"] + tr["write_bytes"]) if tr else None,
                "algorithmic_flop_per_launch": flop, "avg_launch_ms": gms}
        if fused:
            # the Gram kernel also streams wtilde and I_eps for the pole terms
            cms = st["factored_pole_cols"]
            line["fused_epilogue"] = {"algorithmic_bytes": 2 * 16 * NGPOWN * NCOULS,
                                      "note": "wtilde and I_eps read in the Gram epilogue; Mu stays in registers",
                                      "column_sums_kernel": "factored_pole_cols", "column_sums_ms": cms}
            return line
        red_ms = st["factored_pole_reduce"]
        red_bytes = 3 * 16 * NGPOWN * NCOULS  # wtilde, I_eps, Mu
        tr2 = load_traffic("k_pole_reduce") if args.workload == "c1" else None
        line["second_kernel"] = {"bound": "hbm", "kernel": "factored_pole_reduce", "avg_launch_ms": red_ms,
                                 "algorithmic_bytes": red_bytes,
                                 "achieved": red_bytes / (red_ms * 1e-3) / 1e9, "peak": PEAK_HBM_GBS,
                                 "unit": "GB/s", "frac": red_bytes / (red_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                                 "traffic": (tr2["fetch_bytes"] + tr2["write_bytes"]) if tr2 else None}
        return line

    roofline = direct_roofline(args.form) if args.form.startswith("direct") else factored_roofline(args.form)
    for form in forms:
        if form != args.form and not form.endswith("+acht"):
            forms[form]["roofline"] = direct_roofline(form) if form.startswith("direct") else factored_roofline(form)
    st = forms["factored-f64+acht"]["gpu_stage_ms"]
    acht_flop = F64_FLOP_PER_CMAC * NB * NGPOWN * NCOULS
    forms["factored-f64+acht"]["acht_kernel"] = {
        "kernel": "factored_acht (k_acht_f64, v_mfma_f64_16x16x4, 3M)", "avg_launch_ms": st["factored_acht"],
        "algorithmic_flop_per_launch": acht_flop, "achieved": acht_flop / (st["factored_acht"] * 1e-3) / 1e12,
        "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
        "frac": acht_flop / (st["factored_acht"] * 1e-3) / 1e12 / PEAK_FP64_TFLOPS}

    line = {
        "metric": METRIC,
        "value": head["updates_per_s"],
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic: seeded complex128 inputs (seed {SEED}, counter-based SplitMix64) generated in HBM",
        "config": {"workload": f"GPP sigma {args.form} ({args.workload}): number_bands={NB} nvband={NV} "
                               f"ncouls={NCOULS} ngpown={NGPOWN} nfreq={NFREQ} "
                               f"(wx = {wl['wxc']} + {wl['dw']}*(iw - {(NFREQ - 1) / 2}))",
                   "form": args.form, "number_bands": NB, "nvband": NV, "ncouls": NCOULS, "ngpown": NGPOWN,
                   "nfreq": NFREQ, "updates_per_step": updates},
        "roofline": roofline,
        "cpu_baseline": None,
        "forms": forms,
        "slab": slab,
        "error_definition": ("largest relative error per real / imaginary component of each output (a component "
                             "below 1e-6 of its complex modulus relative to the modulus); the Ozaki forms also "
                             "report it relative to each output's modulus (rel_err_modulus), the bound their "
                             "digit split guarantees"),
    }
    ctx.close()
    # The CPU leg, after the GPU work: the reference timed on the host, then
    # the oracle over the whole benchmark problem as the checker.
    if not args.no_cpu:
        cpu = cpu_baseline()
        line["cpu_baseline"] = cpu
        line["gpu_over_cpu"] = head["updates_per_s"] / cpu["value"]
        if args.workload == "c1" and not args.no_cpu_parity:
            par = cpu_parity(results)
            line["cpu_parity"] = {k: v for k, v in par.items() if not k.startswith("rel_err_per_component")}
            for form, e in par["rel_err_per_component"].items():
                forms[form]["rel_err_vs_cpu_oracle"] = e
                forms[form]["rel_err_vs_cpu_oracle_compensated"] = par["rel_err_per_component_compensated"][form]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
