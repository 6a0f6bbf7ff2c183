"""Writes tests/golden/flops_<workload>.json: algorithmic operation counts of the plain
FP64 method on a workload's initial data, measured by the oracle's counting scalar
(oracle/counted.hpp) on the oracle's own path (same roundings, same data-dependent Newton
iterations), twice: along the reference's algorithm, and along the device's, whose HLLD
starts the nested c2p of the HLL state from the average of the two states (fewer Newton
iterations). bench.py reads the device-path counts to turn the flux kernel's measured
duration into FLOP/s.

c5's grid (512^3) is counted on its generator at 256^3 (the oracle's arrays at 512^3 do
not fit this host's memory); the count per interface depends on the data's resolution,
so this over-states the Newton work per interface of the smoother 512^3 data.

Usage:  python tests/golden/make_flop_counts.py [c2 ...]
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ctf_amd import native, workloads  # noqa: E402
from oracle.oracle import Lib, Setup  # noqa: E402


def count(name, nsample=6000):
    w = workloads.ALL[name]
    if w.N > 256:
        w = w.resized(256)
    P = native.init_turbulence(**w.ic_kwargs())
    s = Setup(N=(w.N,) * 3, extrap=w.extrap, riemann=w.riemann, advance=w.advance, gamma=w.gamma)
    solver = Lib("oracle").solver(s)
    out = solver.count_flops(P, nsample=nsample)
    dev = solver.count_flops(P, nsample=nsample, device_path=True)
    rec = {
        "workload": name,
        "grid": w.N,
        "scheme": f"{w.extrap}/{w.riemann}/{w.advance}",
        "state": "initial data (t = 0), ghosts synchronized",
        "nsample_per_axis": nsample,
        "flops_per_interface": {"x": out[0], "y": out[1], "z": out[2]},
        "flops_per_c2p_zone": out[3],
        "transcendentals_per_interface": {"x": out[4], "y": out[5], "z": out[6]},
        "transcendentals_per_c2p_zone": out[7],
        "device_path": {
            "what": "the device's HLLD: nested c2p by anton2dzw from the average state first",
            "flops_per_interface": {"x": dev[0], "y": dev[1], "z": dev[2]},
            "transcendentals_per_interface": {"x": dev[4], "y": dev[5], "z": dev[6]},
        },
        "convention": "+ - * / sqrt count 1 each; pow acos cos sin counted separately",
    }
    path = os.path.join(HERE, f"flops_{name}.json")
    with open(path, "w") as f:
        json.dump(rec, f, indent=1)
    print(path, rec["flops_per_interface"], rec["flops_per_c2p_zone"])


if __name__ == "__main__":
    for n in (sys.argv[1:] or ["c1", "c2", "c5"]):
        count(n)
