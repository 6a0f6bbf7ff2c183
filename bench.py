"""Benchmark of the device-resident RMHD update on one MI355X.

Workload (default): c2 — 256^3 periodic RMHD turbulence, WENO5 / HLLD / RK3, CFL 0.4,
plain FP64 (see ctf_amd/workloads.py). A step is one full time step (all RK stages) of
the whole grid with the state resident in HBM. The first step of a run has dt = 0 (the
reference's priming step) and is untimed; W warm-up steps follow, then K timed steps.

The same initial data is then run through the mixed-precision form (and the learned-c2p
form when a network file is present) with the plain run's dt sequence, and each form
reports zone-updates/s, per-field L1 difference from the plain FP64 state and max |div B|.

value = interior zones x timed steps / wall seconds of the plain FP64 form.

Printed: one JSON line (rank 0). It measures one GPU; --gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from ctf_amd import native, workloads  # noqa: E402

METRIC = "zone-updates/s at 256^3 RMHD (plain vs learned-c2p); L1 err vs FP64"
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector (and matrix) peak
FP32_PEAK_TFLOPS = 157.3     # MI355X FP32 vector peak (counts packed FP32, two per lane)
FP32_UNPACKED_TFLOPS = 78.6  # the same ALUs without packed FP32, which this build does not use
BF16_PEAK_TFLOPS = 2500.0    # MI355X dense BF16 matrix-core peak
HBM_PEAK_GBS = 8000.0
MLP_FLOPS_PER_ZONE = 20480   # matrix-core flops the network issues per zone (padded tiles)
UPDATE_FLOPS_PER_ZONE = 150  # CT EMF averages, flux divergence and the RK combination
FIELDS = ["rho", "pre", "vx", "vy", "vz", "Bx", "By", "Bz"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def flop_counts(w):
    path = os.path.join(ROOT, "tests", "golden", f"flops_{w.name}.json")
    if not os.path.exists(path):
        path = os.path.join(ROOT, "tests", "golden", "flops_c2.json")
    with open(path) as f:
        return json.load(f)


def interfaces_per_sweep(w):
    n = w.N + 2 * w.Ng
    T = n**3
    strides = (n * n, n, 1)
    return [T - 5 * s for s in strides]


def sweep_flops(w, fc=None):
    """Algorithmic flops of one sweep launch (mean over the three axes): the oracle's count
    per interface along the device's path (nested c2p from the average state) times the
    interfaces of the sweep."""
    fc = fc or flop_counts(w)
    fpi = fc.get("device_path", fc)["flops_per_interface"]
    nif = interfaces_per_sweep(w)
    return (nif[0] * fpi["x"] + nif[1] * fpi["y"] + nif[2] * fpi["z"]) / 3.0


def sweep_hw_ops(w):
    """Hardware FP32 operations of one FP32 sweep launch (mean over the axes): the count per
    interface measured at 256^3 times the interfaces of the sweep; None without the file."""
    path = os.path.join(ROOT, "profiles", "sweep_ops_c2_256.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        per = json.load(f)["fp32_ops_per_interface"]
    return per * sum(interfaces_per_sweep(w)) / 3.0


def stage_count(w):
    return {"single": 1, "rk2": 2, "rk3": 3, "rk4": 4}[w.advance]


def update_bytes_per_zone(w, flux_bytes, chained=False):
    """Algorithmic HBM bytes of one fused update + c2p pass per interior zone, averaged
    over the stages: the three axes' 8 fluxes and U^n read; U read by the stages that
    combine it and written by the stage before each of those (rk2, rk3); the RK4
    accumulator read and written by the stages that use it; the primitives written. The
    previous primitives are read only where the first attempt fails (none at these
    workloads). The non-plain forms read the zone's stored FP32 start (Z, W) and write the
    new one (8 B each)."""
    stages = stage_count(w)
    if w.advance in ("rk2", "rk3"):
        u_traffic = 2 * 64 * (stages - 1) / stages
    else:
        u_traffic = 0.0
    acc = (64 + 128 * (stages - 2) + 64) / stages if w.advance == "rk4" else 0.0
    return 3 * 8 * flux_bytes + 64 + u_traffic + acc + 64 + (16 if chained else 0)


def p2c_bytes_per_zone(chained=False):
    """PrimToCons + the first c2p pass: P read, U^n and P written (and, in the non-plain
    forms, the FP32 start for the next pass)."""
    return 192 + (8 if chained else 0)


def form_roofline(w, r, form):
    """The roofline view of one form's kernels from the HIP-event durations the library
    records around them (averages over the timed steps)."""
    fc = flop_counts(w)
    nint = w.N**3
    T = (w.N + 2 * w.Ng) ** 3
    nif = sum(interfaces_per_sweep(w)) / 3.0
    fp32 = form != "plain"
    f_sweep = sweep_flops(w, fc)
    t_sweep = r["flux_ms_per_launch"] * 1e-3
    peak = FP32_PEAK_TFLOPS if fp32 else FP64_PEAK_TFLOPS
    # The FP64 sweep is priced by the oracle's count along its path. The FP32 sweep accepts
    # the HLLD start without a Newton step at most interfaces, so that count would credit it
    # with work it skips: it is priced by the hardware FP32 operations it issues per
    # interface (rocprofv3 SQ counters at 256^3, profiles/sweep_ops_c2_256.json).
    hw = sweep_hw_ops(w) if fp32 else None
    f_priced = hw if hw is not None else f_sweep
    ach = f_priced / t_sweep / 1e12
    sweep = {"kernel": f"k_flux_hlld_fast<{'float' if fp32 else 'double'}> + k_flux_redo (one axis sweep)",
             "bound": "valu", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
             "flops_per_launch": f_priced, "avg_launch_ms": r["flux_ms_per_launch"],
             "flops_source": ("hardware FP32 operations per interface (FMA = 2) x interfaces, "
                              "profiles/sweep_ops_c2_256.json" if hw is not None else
                              "oracle count along the device's path (tests/golden/flops_*.json device_path)"),
             "bytes_per_launch": nif * (64 + (32 if fp32 else 64)),
             "hbm_GBs": nif * (64 + (32 if fp32 else 64)) / t_sweep / 1e9}
    if fp32:
        sweep["oracle_flops_per_launch"] = f_sweep
        sweep["peak_without_packed_fp32"] = FP32_UNPACKED_TFLOPS
        sweep["frac_without_packed_fp32"] = ach / FP32_UNPACKED_TFLOPS
    chained = form != "plain"
    ub = update_bytes_per_zone(w, 4 if fp32 else 8, chained) * nint
    t_upd = r["update_ms_per_launch"] * 1e-3
    upd = {"kernel": f"k_update_c2p<{'float' if fp32 else 'double'}, {'C2P_STORED' if chained else 'C2P_CASCADE'}>",
           "bound": "hbm", "achieved": ub / t_upd / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": ub / t_upd / 1e9 / HBM_PEAK_GBS, "bytes_per_launch": ub, "avg_launch_ms": r["update_ms_per_launch"]}
    pb = p2c_bytes_per_zone(chained) * nint
    t_p2c = r["c2p_ms_per_launch"] * 1e-3
    p2c = {"kernel": f"k_p2c_c2p<{'C2P_NET' if form == 'learned' else 'C2P_CASCADE'}>", "bound": "hbm",
           "achieved": pb / t_p2c / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": pb / t_p2c / 1e9 / HBM_PEAK_GBS, "bytes_per_launch": pb, "avg_launch_ms": r["c2p_ms_per_launch"]}
    stages = stage_count(w)
    c2p_zone = fc["flops_per_c2p_zone"]
    step_flops = (3 * stages * f_priced + (stages + 1) * nint * c2p_zone + stages * nint * UPDATE_FLOPS_PER_ZONE)
    step_s = r["ms_per_step"] * 1e-3
    step = {"flops_per_step": step_flops, "achieved": step_flops / step_s / 1e12, "peak": peak,
            "unit": "TFLOP/s", "frac": step_flops / step_s / 1e12 / peak,
            "what": "the step's flops (sweeps as priced above, c2p passes at the reference cascade's "
                    "count, updates) over its wall time"}
    out = {"sweep": sweep, "update_c2p": upd, "p2c_c2p": p2c, "step": step}
    if form == "learned":
        # the network runs once per step, in the step-start pass
        mf = MLP_FLOPS_PER_ZONE * nint
        out["network"] = {"bound": "mfma", "achieved": mf / t_p2c / 1e12, "peak": BF16_PEAK_TFLOPS,
                          "unit": "TFLOP/s", "frac": mf / t_p2c / 1e12 / BF16_PEAK_TFLOPS,
                          "flops_per_step": mf,
                          "what": "the network's matrix-core flops over the time of the step-start kernel "
                                  "it runs in (k_p2c_c2p<C2P_NET>, which also does PrimToCons and the "
                                  "Newton polish)"}
    return out


def run_form(w, P0, steps, warmup, dts=None, learned_model=None, total_steps=0):
    """Runs one form; returns timings, stats, the state after total_steps steps (the
    priming step not counted; untimed steps follow the timed ones up to that count) and
    the dt sequence."""
    cfg = native.make_config(**w.config_kwargs())
    ctx = native.Context(cfg)
    if learned_model is not None:
        native.mlp_load(ctx, learned_model)
    ctx.upload(P0)
    t0 = time.perf_counter()
    e, st = ctx.advance(0.0)                      # priming step, dt = 0
    prime_ms = 1e3 * (time.perf_counter() - t0)
    if e:
        raise RuntimeError(f"{w.name}: priming step failed on {e} zones")
    used = [] if dts is None else list(dts)
    k = 0

    def next_dt():
        nonlocal k
        if dts is not None:
            d = dts[k]
        else:
            d = ctx.get_timestep(w.cfl)
            used.append(d)
        k += 1
        return d

    for _ in range(warmup):
        e, st = ctx.advance(next_dt())
        if e:
            raise RuntimeError(f"{w.name}: warm-up step failed on {e} zones")
    flux_ms = c2p_ms = upd_ms = 0.0
    flux_n = c2p_n = upd_n = 0
    agg = {}
    t0 = time.perf_counter()
    for _ in range(steps):
        e, st = ctx.advance(next_dt())
        if e:
            raise RuntimeError(f"{w.name}: step failed on {e} zones")
        kt = ctx.kernel_times()
        flux_ms += kt["flux_ms"]; flux_n += kt["flux_launches"]
        c2p_ms += kt["c2p_ms"]; c2p_n += kt["c2p_launches"]
        upd_ms += kt["update_ms"]; upd_n += kt["update_launches"]
        for key, val in st.as_dict().items():
            agg[key] = agg.get(key, 0) + val
    wall = time.perf_counter() - t0
    done = warmup + steps
    while done < total_steps:
        e, _ = ctx.advance(next_dt())
        if e:
            raise RuntimeError(f"{w.name}: step {done + 1} failed on {e} zones")
        done += 1
    divB = ctx.divB()
    P = ctx.download()
    ctx.close()
    nint = w.N**3
    return {
        "wall_s": wall,
        "zone_updates_per_s": nint * steps / wall,
        "ms_per_step": 1e3 * wall / steps,
        "prime_ms": prime_ms,
        "flux_ms_per_launch": flux_ms / max(flux_n, 1),
        "c2p_ms_per_launch": c2p_ms / max(c2p_n, 1),
        "update_ms_per_launch": upd_ms / max(upd_n, 1),
        "flux_ms_per_step": flux_ms / steps,
        "c2p_ms_per_step": c2p_ms / steps,
        "update_ms_per_step": upd_ms / steps,
        "stats": agg,
        "max_divB": divB[2],
        "P": P,
        "dts": used,
        "steps_run": done,
    }


def l1_vs(P, Pref, Ng):
    it = (slice(Ng, -Ng),) * 3
    a, b = P[it], Pref[it]
    out = {}
    for q, name in enumerate(FIELDS):
        den = np.abs(b[..., q]).sum()
        out[name] = float(np.abs(a[..., q] - b[..., q]).sum() / den) if den > 0 else 0.0
    return out


def cpu_baseline(w, seconds_target=15.0):
    """The reference's own CPU path (oracle/_ref, single thread) on a bounded sample of
    the workload: the same generator and scheme on a smaller grid, a priming step plus
    timed steps. Falls back to the oracle's restatement when _ref is absent."""
    sys.path.insert(0, ROOT)
    from oracle.oracle import Lib, Setup, have_ref
    kind = "reference" if have_ref() else "port"
    lib = Lib("ref" if kind == "reference" else "oracle")
    if kind == "port":
        lib.lib.oracle_set_threads(1)
    n = 32
    ws = w.resized(n)
    P = native.init_turbulence(**ws.ic_kwargs())
    s = Setup(N=(n,) * 3, extrap=w.extrap, riemann=w.riemann, advance=w.advance, gamma=w.gamma)
    solver = lib.solver(s)
    e, _ = solver.advance(P, 0.0)
    steps = 0
    t0 = time.perf_counter()
    while True:
        dt = solver.get_timestep(w.cfl)
        e, _ = solver.advance(P, dt)
        steps += 1
        el = time.perf_counter() - t0
        if e or el > seconds_target or steps >= 20:
            break
    rate = n**3 * steps / el
    return {"value": rate, "unit": "zone-updates/s", "cores": 1, "kind": kind,
            "sample": f"{w.name} scheme ({w.extrap}/{w.riemann}/{w.advance}, FP64) on the same "
                      f"generator at {n}^3, {steps} timed step(s) after the dt=0 priming step, "
                      f"{el:.1f} s single-threaded on the GPU box's host"}


def driving_record(w, steps, k1=3):
    """Driving and cooling at the workload's grid: the OU field of ctf-core/drvtrb.lua
    (k1 = 3, 343 modes) sampled at every zone on the device (ResampleField), and the cost of
    the drive and T4 cooling source steps per advance (plain form, with against without).
    Beside it, the reference's own ResampleField (oracle/_ref, one thread) on a bounded
    32^3 sample, scaled to zones per second."""
    f = native.OuField.create(0.8, 0.2, k1, 12345)
    f.advance(0.01)
    P0 = native.init_turbulence(**w.ic_kwargs())
    ctx = native.Context(native.make_config(**w.config_kwargs()))
    ctx.upload(P0)
    ctx.set_driving(f)                                   # first call allocates the planes
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        ctx.set_driving(f)
    sample_s = (time.perf_counter() - t0) / reps
    T = (w.N + 2 * w.Ng) ** 3
    e, _ = ctx.advance(0.0)
    dt = ctx.get_timestep(w.cfl)

    def step_ms(n):
        t = time.perf_counter()
        for _ in range(n):
            err, _ = ctx.advance(dt)
            if err:
                raise RuntimeError("driving record: step failed")
        return 1e3 * (time.perf_counter() - t) / n

    with_src = None
    ctx.set_cooling("T4", 1e-18, 100.0)
    ctx.advance(dt)
    with_src = step_ms(steps)
    ctx.set_driving(None)
    ctx.set_cooling(None)
    without = step_ms(steps)
    ctx.close()
    rec = {"grid": w.N, "modes": (2 * k1 + 1) ** 3, "zones_sampled": T,
           "resample_ms": 1e3 * sample_s, "resample_zones_per_s": T / sample_s,
           "step_ms_with_drive_and_cool": with_src, "step_ms_without": without,
           "source_steps_ms": with_src - without}
    sys.path.insert(0, ROOT)
    from oracle.oracle import Lib, Setup, have_ref
    if have_ref():
        n = 32
        lib = Lib("ref")
        so = lib.solver(Setup(N=(n,) * 3))
        P = native.init_turbulence(**w.resized(n).ic_kwargs())
        t = time.perf_counter()
        lib.drive(so, f.bytes(), P, 0.0)                 # DrivingProcedure samples every zone
        el = time.perf_counter() - t
        rec["cpu_reference_resample_zones_per_s"] = (n + 2 * w.Ng) ** 3 / el
        rec["cpu_sample"] = f"reference DrivingProcedure at {n}^3 ({(n + 6) ** 3} zones), {el:.1f} s, one thread"
    return rec


def weno_char_roofline(N, prec, launch_ms):
    """The characteristic sweep against the FP64 vector peak: its hardware FP64 operations
    per launch (SQ counters of the same launch at 256^3, profiles/weno_split_char_256.json)
    over the live duration of one axis sweep (k_zone_flux + k_flux_weno_split_char). The
    oracle does not count this path, so the count is the hardware's."""
    path = os.path.join(ROOT, "profiles", "weno_split_char_256.json")
    if N != 256 or not os.path.exists(path):
        return None
    with open(path) as f:
        prof = json.load(f)
    ops = prof["fp64" if prec == "fp64" else "mixed"]["hw_fp64_ops_per_launch"]
    ach = ops / (launch_ms * 1e-3) / 1e12
    return {"kernel": "k_zone_flux + k_flux_weno_split_char (one axis sweep)", "bound": "valu",
            "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFLOPS,
            "ops_per_launch": ops, "avg_launch_ms": launch_ms, "ops_source": prof["what"] + ", " + prof["source"]}


def weno_split_record(w, N, steps):
    """The weno-split operator (local Lax-Friedrichs, finite-difference WENO5, RK3) on the
    same turbulence at N^3: component-wise and RMHD-characteristic projections, FP64 and
    mixed, timed over `steps` steps after the priming step and one warm-up step."""
    base = workloads.replace(w.resized(N), godunov="weno-split", c2p="plain")
    P0 = native.init_turbulence(**base.ic_kwargs())
    rec = {"grid": N, "steps": steps, "fsplit": "llf", "extrap": "weno5 (finite difference)",
           "advance": base.advance}
    ref = None
    for eigen in ("component", "rmhd"):
        for prec in ("fp64", "mixed"):
            log(f"[bench] weno-split {eigen} {prec} {N}^3")
            wf = workloads.replace(base, eigen=eigen, precision=prec)
            r = run_form(wf, P0, steps, 1, dts=None if ref is None else ref["dts"],
                         total_steps=1 + steps)
            if ref is None:
                ref = r
            st = r["stats"]
            item = {"zone_updates_per_s": r["zone_updates_per_s"], "ms_per_step": r["ms_per_step"],
                    "flux_ms_per_step": r["flux_ms_per_step"], "max_divB": r["max_divB"]}
            if eigen == "rmhd":
                stages = {"single": 1, "rk2": 2, "rk3": 3, "rk4": 4}[base.advance]
                solves = 3 * stages * steps * (N + 1) * (N + 2 * base.Ng) ** 2
                item["identity_projection_fraction"] = st.get("eig_identity", 0) / solves
                item["roofline"] = weno_char_roofline(N, prec, r["flux_ms_per_launch"])
            if r is not ref:
                item["L1_vs_component_fp64"] = l1_vs(r["P"], ref["P"], w.Ng)
            rec[f"{eigen}_{prec}"] = item
    return rec


def plm_muscl_record(w, N, steps, cpu=True):
    """The plm-muscl operator (CTU-Hancock: predictor face fluxes, a Hancock half step per
    axis, corrector fluxes) against plm-split with the same PLM reconstruction, HLLD and
    Runge-Kutta scheme, on the same turbulence at N^3, FP64: `steps` timed steps after the
    priming step and one warm-up step, with the same dt sequence for both."""
    base = workloads.replace(w.resized(N), extrap="plm", c2p="plain", precision="fp64")
    P0 = native.init_turbulence(**base.ic_kwargs())
    rec = {"grid": N, "steps": steps, "riemann": base.riemann, "advance": base.advance,
           "theta": 2.0}
    log(f"[bench] plm-split (plm) {N}^3")
    split = run_form(workloads.replace(base, godunov="plm-split"), P0, steps, 1, total_steps=1 + steps)
    log(f"[bench] plm-muscl {N}^3")
    ctu = run_form(workloads.replace(base, godunov="plm-muscl"), P0, steps, 1, dts=split["dts"],
                   total_steps=1 + steps)
    for name, r in (("plm_split_plm", split), ("plm_muscl", ctu)):
        st = r["stats"]
        rec[name] = {"zone_updates_per_s": r["zone_updates_per_s"], "ms_per_step": r["ms_per_step"],
                     "flux_ms_per_step": r["flux_ms_per_step"], "c2p_ms_per_step": r["c2p_ms_per_step"],
                     "update_ms_per_step": r["update_ms_per_step"], "max_divB": r["max_divB"],
                     "riemann_solves": st.get("hlld_success", 0) + st.get("hlld_failure", 0)}
    rec["plm_muscl"]["L1_vs_plm_split"] = l1_vs(ctu["P"], split["P"], w.Ng)
    # the reference's own PlmCtuHancockOperator on the host, one thread, beside it
    sys.path.insert(0, ROOT)
    from oracle.oracle import Lib, Setup, have_ref
    if cpu and have_ref():
        n = 32
        lib = Lib("ref")
        so = lib.solver(Setup(N=(n,) * 3, godunov="plm-muscl", riemann=base.riemann, advance=base.advance,
                              gamma=w.gamma))
        P = native.init_turbulence(**w.resized(n).ic_kwargs())
        so.advance(P, 0.0)
        steps_cpu = 0
        t = time.perf_counter()
        while steps_cpu < 2:
            e, _ = so.advance(P, so.get_timestep(w.cfl))
            steps_cpu += 1
            if e:
                break
        el = time.perf_counter() - t
        rec["cpu_reference_zone_updates_per_s"] = n ** 3 * steps_cpu / el
        rec["cpu_sample"] = (f"reference PlmCtuHancockOperator ({base.riemann}/{base.advance}) at {n}^3, "
                             f"{steps_cpu} step(s) after the priming step, {el:.1f} s, one thread")
    return rec


def traffic_record(w):
    path = os.path.join(ROOT, "profiles", f"traffic_{w.name}.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def forms_record(w, forms, steps, warmup, acc_steps, model):
    """Runs the plain FP64 form, then the mixed and mixed + learned forms on its dt
    sequence; returns the per-form records and the plain run."""
    log(f"[bench] initial data {w.name} {w.N}^3")
    P0 = native.init_turbulence(**w.ic_kwargs())
    results = {}
    log(f"[bench] {w.name} plain FP64")
    plain = run_form(w, P0, steps, warmup, total_steps=acc_steps)
    results["plain"] = plain
    dts = plain["dts"]
    if "mixed" in forms:
        log(f"[bench] {w.name} mixed precision")
        wm = workloads.replace(w, precision="mixed")
        results["mixed"] = run_form(wm, P0, steps, warmup, dts=dts, total_steps=acc_steps)
    if "learned" in forms:
        if not os.path.exists(model):
            raise SystemExit(f"learned form requested but no network file at {model}")
        log(f"[bench] {w.name} mixed + learned c2p")
        wl = workloads.replace(w, precision="mixed", c2p="learned")
        results["learned"] = run_form(wl, P0, steps, warmup, dts=dts, learned_model=model, total_steps=acc_steps)
    del P0
    forms_out = {}
    for name, r in results.items():
        rec = {"zone_updates_per_s": r["zone_updates_per_s"], "ms_per_step": r["ms_per_step"],
               "speedup_vs_plain": r["zone_updates_per_s"] / plain["zone_updates_per_s"],
               "flux_ms_per_step": r["flux_ms_per_step"], "c2p_ms_per_step": r["c2p_ms_per_step"],
               "update_ms_per_step": r["update_ms_per_step"], "max_divB": r["max_divB"],
               "priming_step_ms": r["prime_ms"]}
        rec["accuracy_after_steps"] = r["steps_run"]
        if name != "plain":
            rec["L1_vs_fp64"] = l1_vs(r["P"], plain["P"], w.Ng)
        st = r["stats"]
        zones = max(st.get("c2p_zones", 0), 1)
        rec["newton_iterations_per_zone"] = st.get("c2p_iterations", 0) / zones
        rec["c2p_fallback_fraction"] = st.get("c2p_fallbacks", 0) / zones
        rec["hlld_to_hllc_fallbacks"] = st.get("hlld_failure", 0)
        rec["hlld_solves"] = st.get("hlld_success", 0) + st.get("hlld_failure", 0)
        if rec["hlld_solves"]:
            rec["hlld_newton_iterations_per_solve"] = st.get("hlld_newton_iterations", 0) / rec["hlld_solves"]
        if st.get("fp64_redo", 0) or st.get("fp64_redo_overflow", 0):
            rec["fp64_redo_interfaces"] = st.get("fp64_redo", 0)
            rec["fp64_redo_overflow"] = st.get("fp64_redo_overflow", 0)
        if name == "learned":
            g = max(st.get("mlp_guesses", 0), 1)
            rec["mlp_fallback_fraction"] = st.get("mlp_fallbacks", 0) / g
        rec["roofline"] = form_roofline(w, r, name)
        forms_out[name] = rec
    for r in results.values():
        r.pop("P", None)
    return forms_out, plain


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--accuracy-steps", type=int, default=None,
                    help="steps after which L1 and div B are measured (default: the workload's)")
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--grid", type=int, default=None, help="override the workload's grid size")
    ap.add_argument("--forms", default="plain,mixed,learned")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-c5", dest="c5", action="store_false",
                    help="skip the c5 record (512^3, every form)")
    ap.add_argument("--c5-steps", type=int, default=3)
    ap.add_argument("--weno-split-grid", type=int, default=256,
                    help="grid of the weno-split measurement (0: skip it)")
    ap.add_argument("--weno-split-steps", type=int, default=3)
    ap.add_argument("--no-driving", action="store_true", help="skip the driving/cooling record")
    ap.add_argument("--plm-muscl-grid", type=int, default=256,
                    help="grid of the plm-muscl (CTU-Hancock) measurement (0: skip it)")
    ap.add_argument("--plm-muscl-steps", type=int, default=3)
    ap.add_argument("--model", default=os.path.join(ROOT, "ctf_amd", "models", "c2p_mlp.bin"))
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (calibrates the "
                         "HBM byte counters read by rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU only: the update runs on a single "
                                   "MI355X, and --gpus above 1 is not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    w = workloads.ALL[args.workload]
    if args.grid:
        w = w.resized(args.grid)
    w = workloads.replace(w, precision="fp64", c2p="plain")
    forms = [f for f in args.forms.split(",") if f]
    acc_steps = w.steps if args.accuracy_steps is None else args.accuracy_steps
    forms_out, plain = forms_record(w, forms, args.steps, args.warmup, acc_steps, args.model)

    fc = flop_counts(w)
    flops_per_launch = sweep_flops(w, fc)
    avg_flux_s = plain["flux_ms_per_launch"] * 1e-3
    achieved = flops_per_launch / avg_flux_s / 1e12
    roofline = {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                "kernel": "k_flux_hlld_fast<double> + k_flux_redo (one axis sweep, WENO5/HLLD)",
                "flops_per_launch": flops_per_launch,
                "flops_source": "oracle count along the device's path (nested c2p from the average state), "
                                "tests/golden/flops_%s.json" % w.name,
                "avg_launch_ms": plain["flux_ms_per_launch"],
                "traffic": None}
    tr = traffic_record(w)
    if tr:
        roofline["traffic"] = tr.get("bytes_per_launch")
        roofline["traffic_source"] = tr.get("source")
        # the same launch against HBM: measured bytes (the profile's counters) over the
        # live launch time; the kernel is not HBM-bound, this says how far it is from it
        if roofline["traffic"]:
            gbs = roofline["traffic"] / avg_flux_s / 1e9
            roofline["hbm"] = {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}

    c5 = None
    if args.c5 and w.name == "c2":
        log("[bench] c5 record")
        w5 = workloads.replace(workloads.C5, precision="fp64", c2p="plain")
        c5_forms, _ = forms_record(w5, forms, args.c5_steps, 1, w5.steps, args.model)
        c5 = {"workload": f"c5: {w5.N}^3 periodic RMHD turbulence, 4-velocity rms {w5.v_rms}, Bz = {w5.Bz}, "
                          f"{w5.extrap.upper()}/{w5.riemann.upper()}/{w5.advance.upper()}",
              "grid": w5.N, "timed_steps": args.c5_steps, "warmup": 1, "forms": c5_forms}

    ws = None
    if args.weno_split_grid:
        ws = weno_split_record(w, args.weno_split_grid, args.weno_split_steps)
    pm = None
    if args.plm_muscl_grid:
        pm = plm_muscl_record(w, args.plm_muscl_grid, args.plm_muscl_steps, cpu=not args.no_cpu_baseline)
    dr = None
    if not args.no_driving:
        log("[bench] driving and cooling")
        dr = driving_record(w, 3)

    out = {
        "metric": METRIC,
        "value": plain["zone_updates_per_s"],
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": plain["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{w.name}: {w.N}^3 periodic RMHD turbulence, "
                               f"{w.extrap.upper()}/{w.riemann.upper()}/{w.advance.upper()}, "
                               f"CFL {w.cfl}, plain FP64 (Mara plm-split + extrap=weno5)",
                   "grid": w.N, "ghosts": w.Ng, "gamma": w.gamma,
                   "interior_zones": w.N**3, "priming_step_excluded": True},
        "forms": forms_out,
        "roofline": roofline,
    }
    if c5:
        out["c5"] = c5
    if ws:
        out["weno_split"] = ws
    if pm:
        out["plm_muscl"] = pm
    if dr:
        out["driving"] = dr
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline")
        cb = cpu_baseline(w)
        out["cpu_baseline"] = cb
    print(json.dumps(out))


if __name__ == "__main__":
    main()
