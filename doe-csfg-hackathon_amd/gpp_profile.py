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

Workloads: c1 = number_bands 512, nvband 2, ncouls 32768, ngpown 1638,
nfreq 3 (the benchmark shape); c3 = 1536, 8, 65536, 3276, nfreq 16 with the
frequency grid 4 + 0.25 (iw - 7.5).
"""
import argparse
import ctypes
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import pygpp  # noqa: E402

WORKLOADS = {
    "c1": (pygpp.Problem(512, 2, 32768, 1638, 3), 4.0, 1.0),
    "c3": (pygpp.Problem(1536, 8, 65536, 3276, 16), 4.0, 0.25),
}


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
    ap.add_argument("--workload", default="c1", choices=sorted(WORKLOADS))
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--copy-bytes", type=int, default=0)
    args = ap.parse_args()
    if args.copy_bytes:
        copy_calibration(args.copy_bytes, args.steps)
        return
    prob, wxc, dw = WORKLOADS[args.workload]
    with pygpp.Context(0) as ctx:
        ctx.generate_seeded(prob, args.seed, wxc, dw)
        for i in range(args.steps):
            out = ctx.sigma(args.form)
            t = ctx.last_timing()
            print(f"step {i}: {args.form} {args.workload} total {t['total']:.3f} ms "
                  f"achtemp[0]={out['achtemp'][0]:.10e}", flush=True)


if __name__ == "__main__":
    main()
