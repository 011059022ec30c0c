#!/usr/bin/env python3
"""Profiling driver: one GPP sigma workload through the C ABI, sized for
rocprofv3 runs, or a device-to-device copy of known size used to calibrate
the FETCH_SIZE / WRITE_SIZE counters.

    rocprofv3 --kernel-trace --stats -d OUT -o run --output-format csv -- \\
        python3 doe-csfg-hackathon_amd/gpp_profile.py --form direct
    rocprofv3 --pmc FETCH_SIZE -d OUT -o run --output-format csv -- \\
        python3 doe-csfg-hackathon_amd/gpp_profile.py --form direct --steps 1
    rocprofv3 --pmc FETCH_SIZE -d OUT -o run --output-format csv -- \\
        python3 doe-csfg-hackathon_amd/gpp_profile.py --copy-bytes 1073741824

Workloads (gpp_workloads.py): c1 (the benchmark shape), c3 (the large
single instance), c4 (64 uneven instances through the batched persistent
kernels, replayed from a hipGraph).  --acht also requests acht_n1_loc from the
single-instance forms.
"""
import argparse
import ctypes
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import pygpp  # noqa: E402
from gpp_workloads import SINGLE, c4_instances  # noqa: E402


def copy_calibration(nbytes: int, reps: int) -> None:
    pygpp.load()
    rt = ctypes.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rt.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert rt.hipMalloc(ctypes.byref(a), nbytes) == 0 and rt.hipMalloc(ctypes.byref(b), nbytes) == 0
    assert rt.hipMemset(a, 1, nbytes) == 0
    for _ in range(reps):
        assert rt.hipMemcpy(b, a, nbytes, 3) == 0  # hipMemcpyDeviceToDevice
    assert rt.hipDeviceSynchronize() == 0
    print(f"copied {nbytes} bytes x {reps}", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--form", default="direct", choices=sorted(pygpp.FORMS))
    ap.add_argument("--workload", default="c1", choices=sorted(SINGLE) + ["c4"])
    ap.add_argument("--acht", action="store_true", help="also request acht_n1_loc")
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--copy-bytes", type=int, default=0)
    args = ap.parse_args()
    if args.copy_bytes:
        copy_calibration(args.copy_bytes, args.steps)
        return
    if args.workload == "c4":
        inst = c4_instances()
        with pygpp.Context(0) as ctx:
            with pygpp.Batch(ctx, [p for p, _, _, _ in inst], seeds=[s for _, s, _, _ in inst],
                             wx_center=[w for _, _, w, _ in inst], dw=[d for _, _, _, d in inst]) as batch:
                batch.autotune(args.form)
                batch.use_graph(True)
                for i in range(args.steps):
                    outs = batch.sigma(args.form)
                    print(f"step {i}: batch {args.form} c4 {batch.last_timing()['ms']:.3f} ms "
                          f"achtemp[0] of instance 0 = {outs[0]['achtemp'][0]:.10e}", flush=True)
        return
    prob, wxc, dw = SINGLE[args.workload]
    with pygpp.Context(0) as ctx:
        ctx.generate_seeded(prob, args.seed, wxc, dw)
        for i in range(args.steps):
            out = ctx.sigma(args.form, acht_n1_loc=True if args.acht else None)
            t = ctx.last_timing()
            print(f"step {i}: {args.form} {args.workload} total {t['total']:.3f} ms "
                  f"achtemp[0]={out['achtemp'][0]:.10e}", flush=True)


if __name__ == "__main__":
    main()
