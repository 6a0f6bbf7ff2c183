"""Trains the learned c2p initial guess on the GPU and saves it.

Training pairs (conserved -> primitive) come from two sources, both generated on the
device: snapshots of the 256^3 WENO5/HLLD/RK3 turbulence run (the committed primitive
state after a number of steps, sampled zone by zone) and uniform samples over Lorentz
factor 1..W_max, magnetisation sigma, p/rho and random field-velocity angles. The
network is trained by the library's own BF16 matrix-core forward/backward kernels with
Adam, then checked on held-out states: the learned-guess recovery must return the same
primitives as the plain c2p cascade, and the script reports Newton iterations per zone
for both and the fraction of zones that fell back.

    python -m ctf_amd.train_c2p --out ctf_amd/models/c2p_mlp.bin
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from ctf_amd import native, workloads  # noqa: E402

DEFAULT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models", "c2p_mlp.bin")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def snapshot_pairs(data, w, N, steps_at, per_snapshot, seed):
    """Runs workload w at N^3 (mixed form) and adds per_snapshot pairs from the committed
    state at each step count in steps_at."""
    ctx = native.Context(native.make_config(**dict(w.config_kwargs(N), precision="mixed",
                                                   c2p="plain")))
    ctx.upload(native.init_turbulence(**w.ic_kwargs(N)))
    done, dt = 0, 0.0
    for k, target in enumerate(sorted(steps_at)):
        while done < target:
            e, _ = ctx.advance(dt)
            if e:
                raise RuntimeError(f"snapshot run failed at step {done} ({e} zones)")
            dt = ctx.get_timestep(w.cfl)
            done += 1
        data.add_state(ctx, per_snapshot, seed=seed + k, gamma=w.gamma)
        log(f"[train] snapshot at step {done}: {len(data)} pairs")
    ctx.close()


def held_out_states(n, seed, gamma, W_max, sigma, ptr):
    """Primitive states over the sampling ranges (host, numpy), and their conserved form."""
    rng = np.random.default_rng(seed)
    W = np.exp(rng.uniform(0.0, np.log(W_max), n))
    vd = rng.normal(size=(n, 3))
    vd /= np.linalg.norm(vd, axis=1, keepdims=True)
    v = vd * np.sqrt(1.0 - 1.0 / W**2)[:, None]
    rho = np.exp(rng.uniform(np.log(0.1), np.log(10.0), n))
    pre = rho * np.exp(rng.uniform(np.log(ptr[0]), np.log(ptr[1]), n))
    h = 1.0 + gamma / (gamma - 1.0) * pre / rho
    sig = np.exp(rng.uniform(np.log(sigma[0]), np.log(sigma[1]), n))
    bd = rng.normal(size=(n, 3))
    bd /= np.linalg.norm(bd, axis=1, keepdims=True)
    P = np.zeros((n, 8))
    P[:, 0], P[:, 1], P[:, 2:5] = rho, pre, v
    P[:, 5:8] = bd * np.sqrt(sig * rho * h)[:, None]
    U, err = native.prim_to_cons(P, gamma)
    return P[err == 0], U[err == 0]


def evaluate(net, U, gamma):
    """Plain cascade vs learned guess on the same conserved states (device)."""
    Pp, ep, ip = native.cons_to_prim(U, gamma)
    Pl, el, il, used, _ = net.cons_to_prim(U, gamma)
    ok = (ep == 0) & (el == 0)
    rel = np.abs(Pl[ok] - Pp[ok]).max(1) / np.abs(Pp[ok]).max(1)
    return {
        "states": int(len(U)),
        "plain_failed": int((ep != 0).sum()),
        "learned_failed": int((el != 0).sum()),
        "max_rel_diff": float(rel.max()) if ok.any() else None,
        "plain_iters_per_zone": float(ip[ok].mean()),
        "learned_iters_per_zone": float(il[ok].mean()),
        "fallback_fraction": float(1.0 - used[el == 0].mean()),
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--pairs", type=float, default=1e7, help="total training pairs")
    ap.add_argument("--snapshot-fraction", type=float, default=0.3)
    ap.add_argument("--snapshot-N", type=int, default=256)
    ap.add_argument("--snapshot-steps", default="0,10,25,50")
    ap.add_argument("--W-max", type=float, default=50.0)
    ap.add_argument("--sigma", default="0.01,100")
    ap.add_argument("--ptr", default="1e-4,10")
    ap.add_argument("--epochs", type=int, default=24)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--lr-final", type=float, default=1e-5)
    ap.add_argument("--seed", type=int, default=20240601)
    ap.add_argument("--report", default=None, help="write the training report (JSON) here")
    a = ap.parse_args(argv)

    w = workloads.C2
    sigma = tuple(float(x) for x in a.sigma.split(","))
    ptr = tuple(float(x) for x in a.ptr.split(","))
    n_total = int(a.pairs)
    steps_at = [int(s) for s in a.snapshot_steps.split(",") if s != ""]
    n_snap = int(n_total * a.snapshot_fraction) if steps_at else 0
    per_snapshot = n_snap // max(1, len(steps_at))
    n_syn = n_total - per_snapshot * len(steps_at)

    t0 = time.time()
    data = native.MlpData(n_total)
    data.add_synthetic(n_syn, seed=a.seed, gamma=w.gamma, W_max=a.W_max, sigma=sigma, ptr=ptr)
    log(f"[train] synthetic: {len(data)} pairs")
    if n_snap:
        snapshot_pairs(data, w, a.snapshot_N, steps_at, per_snapshot, a.seed + 100)
    t_data = time.time() - t0

    net = native.Mlp.create(seed=a.seed)
    rep = net.train(data, epochs=a.epochs, batch=a.batch, lr=a.lr, lr_final=a.lr_final, seed=a.seed + 1)
    log(f"[train] {rep['steps']} steps over {rep['samples']} pairs in {rep['seconds']:.1f} s: "
        f"loss {rep['loss_first']:.4g} -> {rep['loss_last']:.4g}")
    net.save(a.out)

    _, U = held_out_states(200_000, a.seed + 7, w.gamma, a.W_max, sigma, ptr)
    ev = evaluate(net, U, w.gamma)
    log(f"[train] held-out synthetic: {ev}")
    ctx = native.Context(native.make_config(**dict(w.config_kwargs(64), precision="fp64")))
    ctx.upload(native.init_turbulence(**w.ic_kwargs(64)))
    for _ in range(3):
        ctx.advance(0.0 if _ == 0 else ctx.get_timestep(w.cfl))
    P = ctx.download()[3:-3, 3:-3, 3:-3].reshape(-1, 8)
    Ut, _ = native.prim_to_cons(P, w.gamma)
    ev_run = evaluate(net, Ut, w.gamma)
    log(f"[train] turbulence states: {ev_run}")
    out = {"model": os.path.relpath(a.out), "pairs": len(data), "data_seconds": t_data,
           "train": rep, "held_out_synthetic": ev, "turbulence_64": ev_run,
           "config": {k: v for k, v in vars(a).items() if k not in ("out", "report")}}
    if a.report:
        with open(a.report, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
