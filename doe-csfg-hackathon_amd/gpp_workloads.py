"""The benchmark workloads, shared by bench.py and gpp_profile.py.

c1  the benchmark shape: number_bands 512, nvband 2, ncouls 32768,
    ngpown 1638, nfreq 3 (wx = 3, 4, 5), seed 1
c3  the large single instance: 1536, 8, 65536, 3276, nfreq 16 on the grid
    wx = 4 + 0.25 (iw - 7.5), seed 1
c4  64 independent instances with nb ~ U{128..1024}, nvband ~ U{2..8},
    ncouls ~ U{2048..16384}, ngpown = ncouls // 20, nfreq ~ U{3..32}, each on
    the frequency window [3, 5] (wx = 4 + dw (iw - (nfreq-1)/2),
    dw = 2/(nfreq-1)), seeds 1000..1063

Inputs have the seeded benchmark distributions of gpp_fill_seeded (aqs re/im
~U(-1,1), I_eps ~U(-0.5,0.5), wtilde re ~U(0.5,5) im ~U(-0.05,0.05), vcoul
~U(0.1,10), sorted distinct random inv_igp_index).
"""
import numpy as np

import pygpp

SEED = 1
SINGLE = {
    "c1": (pygpp.Problem(512, 2, 32768, 1638, 3), 4.0, 1.0),
    "c3": (pygpp.Problem(1536, 8, 65536, 3276, 16), 4.0, 0.25),
}


def c4_instances(seed=4, count=64):
    """[(Problem, seed, wx_center, dw)] of the batched workload."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        nb = int(rng.integers(128, 1025))
        nv = int(rng.integers(2, 9))
        ncouls = int(rng.integers(2048, 16385))
        nfreq = int(rng.integers(3, 33))
        out.append((pygpp.Problem(nb, nv, ncouls, ncouls // 20, nfreq), 1000 + i, 4.0, 2.0 / (nfreq - 1)))
    return out
