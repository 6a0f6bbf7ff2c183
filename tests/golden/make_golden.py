"""Writes the golden vectors under tests/golden/ from the compiled reference.

Run in a container that has the reference tree:  make -C oracle ref && python tests/golden/make_golden.py

Every expected output in the fixtures is computed by the reference's own Mara sources
(oracle/_ref/libmararef.so, built by oracle/Makefile); the inputs are seeded samples
drawn here. The fixtures are data only: inputs and the reference's outputs.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle.oracle import Lib, Setup, synchronize  # noqa: E402

GAMMA = 4.0 / 3.0


def sample_states(rng, n, Wmax=50.0, sig=(0.01, 100.0), ptr=(1e-4, 10.0)):
    """Primitive states over Lorentz factor 1..Wmax, magnetization sig, p/rho ptr and
    random field-velocity angles (the learned-c2p training ranges)."""
    W = np.exp(rng.uniform(0.0, np.log(Wmax), n))
    vdir = rng.normal(size=(n, 3))
    vdir /= np.linalg.norm(vdir, axis=1, keepdims=True)
    v = vdir * np.sqrt(1.0 - 1.0 / W**2)[:, None]
    rho = np.exp(rng.uniform(np.log(0.1), np.log(10.0), n))
    pre = rho * np.exp(rng.uniform(np.log(ptr[0]), np.log(ptr[1]), n))
    h = 1.0 + GAMMA / (GAMMA - 1.0) * pre / rho
    sigma = np.exp(rng.uniform(np.log(sig[0]), np.log(sig[1]), n))
    b = np.sqrt(sigma * rho * h)
    bdir = rng.normal(size=(n, 3))
    bdir /= np.linalg.norm(bdir, axis=1, keepdims=True)
    P = np.zeros((n, 8))
    P[:, 0] = rho
    P[:, 1] = pre
    P[:, 2:5] = v
    P[:, 5:8] = bdir * b[:, None]
    return P


def edge_states():
    """Hand-picked edge cases: zero field, aligned field, cold gas, error paths."""
    base = np.array([1.0, 1.0, 0.3, -0.2, 0.1, 0.5, 0.7, -0.4])
    out = [base.copy()]
    s = base.copy(); s[5:8] = 0.0; out.append(s)                     # no field
    s = base.copy(); s[5:8] = s[2:5] * 3.0; out.append(s)            # B parallel to v
    s = base.copy(); s[1] = 1e-6; out.append(s)                      # cold
    s = base.copy(); s[2:5] = 0.0; out.append(s)                     # at rest
    s = base.copy(); s[2:5] = [0.99, 0.0, 0.0]; out.append(s)        # W ~ 7
    s = base.copy(); s[2:5] = [0.6, 0.6, 0.6]; out.append(s)         # superluminal
    s = base.copy(); s[0] = -1.0; out.append(s)                      # negative density
    s = base.copy(); s[1] = np.nan; out.append(s)                    # NaN
    return np.array(out)


def riemann_pairs(rng, n):
    PL = sample_states(rng, n, Wmax=10.0, sig=(0.01, 10.0), ptr=(1e-2, 10.0))
    PR = PL * (1.0 + 0.2 * rng.normal(size=PL.shape))
    PR[:, 0:2] = np.abs(PR[:, 0:2])
    PR[:, 2:5] = PL[:, 2:5] * rng.uniform(0.5, 1.0, size=(n, 1))
    # a quarter of them: large jumps (shock-tube like)
    m = n // 4
    PR[:m, 0] *= 10.0
    PR[:m, 1] *= 100.0
    return PL, PR


def grid_state(N, kind, rng):
    n = N + 6
    x = (np.arange(n) - 3 + 0.5) / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    P = np.zeros((n, n, n, 8))
    tp = 2 * np.pi
    P[..., 0] = 1.0 + 0.2 * np.sin(tp * X) * np.cos(tp * Z)
    P[..., 1] = 1.0 + 0.3 * np.cos(tp * Y)
    P[..., 2] = 0.3 * np.sin(tp * Y)
    P[..., 3] = 0.3 * np.sin(tp * Z + 0.4)
    P[..., 4] = 0.3 * np.sin(tp * X + 1.1)
    # a divergence-free perturbation: B = B0 + curl of A = (0, 0, a sin(2 pi x) sin(2 pi y))
    P[..., 5] = 0.2 * np.cos(tp * Y) * np.sin(tp * X) * 1.0
    P[..., 6] = -0.2 * np.cos(tp * X) * np.sin(tp * Y)
    P[..., 7] = 1.0
    P[..., :5] *= 1.0 + 0.01 * rng.normal(size=P[..., :5].shape)
    if kind == "jump":
        P[: n // 2, ..., 0] *= 5.0
        P[: n // 2, ..., 1] *= 20.0
    if kind == "superluminal":
        P[6, 6, 6, 2:5] = [0.7, 0.7, 0.2]
    if kind == "cold":
        P[..., 1] *= 1e-4
    return synchronize(P, 3)


GRID_CASES = [
    # name, kind, setup kwargs, steps
    ("plm_hlld_rk2", "smooth", dict(extrap="plm", riemann="hlld", advance="rk2"), 3),
    ("weno5_hlld_rk3", "smooth", dict(extrap="weno5", riemann="hlld", advance="rk3"), 3),
    ("plm_hllc_rk2", "smooth", dict(extrap="plm", riemann="hllc", advance="rk2"), 3),
    ("pcm_hll_single", "smooth", dict(extrap="pcm", riemann="hll", advance="single"), 3),
    ("plm_hlld_rk4", "smooth", dict(extrap="plm", riemann="hlld", advance="rk4"), 2),
    ("weno5_b08_hlld_rk3", "smooth", dict(extrap="weno5", IS="b08", riemann="hlld", advance="rk3"), 2),
    ("weno5_sz10_hlld_rk3", "smooth", dict(extrap="weno5", IS="sz10", sz10A=100.0, riemann="hlld",
                                           advance="rk3"), 2),
    ("plm_hlld_rk2_jump", "jump", dict(extrap="plm", riemann="hlld", advance="rk2"), 3),
    ("plm_theta1_hlld_rk2", "smooth", dict(extrap="plm", theta=1.0, riemann="hlld", advance="rk2"), 2),
    ("plm_hlld_rk2_superluminal", "superluminal", dict(extrap="plm", riemann="hlld", advance="rk2"), 1),
    ("plm_hlld_rk2_superluminal_ereset", "superluminal",
     dict(extrap="plm", riemann="hlld", advance="rk2", ereset=True, preset=0.5, dreset=2.0), 2),
    ("plm_hlld_rk2_cold_pfloor", "cold", dict(extrap="plm", riemann="hlld", advance="rk2",
                                             pfloor=1e-3), 2),
    # weno-split (local Lax-Friedrichs splitting), run by the reference's own WenoSplit with the
    # identity eigensystem supplied by oracle/ref_harness.cpp (component-wise projection)
    ("weno_split_rk3", "smooth", dict(godunov="weno-split", advance="rk3"), 3),
    ("weno_split_rk2_jump", "jump", dict(godunov="weno-split", advance="rk2"), 3),
    ("weno_split_sz10_rk3", "smooth", dict(godunov="weno-split", IS="sz10", sz10A=100.0,
                                           advance="rk3"), 2),
    ("weno_split_rk2_superluminal_ereset", "superluminal",
     dict(godunov="weno-split", advance="rk2", ereset=True, preset=0.5, dreset=2.0), 2),
    # ... with the characteristic projection: the reference's WenoSplit driven by the
    # oracle's RMHD eigensystem
    ("weno_split_char_rk3", "smooth", dict(godunov="weno-split", eigen="rmhd", advance="rk3"), 3),
    ("weno_split_char_rk2_jump", "jump", dict(godunov="weno-split", eigen="rmhd", advance="rk2"), 3),
    # plm-muscl: the reference's PlmCtuHancockOperator (its own theta)
    ("plm_muscl_hlld_single", "smooth", dict(godunov="plm-muscl", riemann="hlld", advance="single"), 3),
    ("plm_muscl_hll_rk2_theta125", "smooth", dict(godunov="plm-muscl", riemann="hll", advance="rk2",
                                                  theta=1.25), 3),
    ("plm_muscl_hlld_single_jump", "jump", dict(godunov="plm-muscl", riemann="hlld", advance="single"), 3),
    ("plm_muscl_hlld_superluminal", "superluminal", dict(godunov="plm-muscl", riemann="hlld",
                                                          advance="single"), 1),
]


def save(write, path, **arrays):
    if write:
        np.savez_compressed(path, **arrays)


def main(only=None):
    """Writes every fixture, or with `only` (grid-case names) just those grid cases; the
    random stream, hence each case's initial data, is the same either way."""
    ref = Lib("ref")
    rng = np.random.default_rng(20240601)
    write_kats(ref, rng, write=not only)
    write_grids(ref, rng, only)
    print("golden vectors written to", HERE)


def write_kats(ref, rng, write=True):
    out = {}

    # --- conserved -> primitive -------------------------------------------------
    P = np.concatenate([sample_states(rng, 600), edge_states()])
    U, perr = ref.prim_to_cons(P, GAMMA)
    guess = P * (1.0 + 1e-3 * rng.normal(size=P.shape))
    Pc, ec, ic = ref.cons_to_prim(U, GAMMA, guess=guess)
    Pa, ea, ia = ref.c2p_solver(U, 0, GAMMA)
    Pn, en, inn = ref.c2p_solver(U, 1, GAMMA)
    Pz, ez, iz = ref.cons_to_prim(U, GAMMA, guess=np.zeros_like(P))
    Pf, ef, itf = ref.cons_to_prim(U, GAMMA, pfloor=0.05, guess=np.zeros_like(P))
    save(write, os.path.join(HERE, "kat_c2p.npz"), P=P, U=U, p2c_err=perr, guess=guess,
                        cascade_P=Pc, cascade_err=ec, anton_P=Pa, anton_err=ea, anton_iters=ia,
                        noble_P=Pn, noble_err=en, noble_iters=inn, zero_guess_P=Pz,
                        zero_guess_err=ez, pfloor=0.05, pfloor_P=Pf, pfloor_err=ef, gamma=GAMMA)

    # --- fluxes and fast speeds ---------------------------------------------------
    P = sample_states(rng, 400, Wmax=20.0)
    d = {"P": P, "gamma": GAMMA}
    for dim in (1, 2, 3):
        F, ap, am = ref.flux_eigen(P, dim, GAMMA)
        d[f"F{dim}"], d[f"ap{dim}"], d[f"am{dim}"] = F, ap, am
    save(write, os.path.join(HERE, "kat_flux.npz"), **d)

    # --- quartic ------------------------------------------------------------------
    C = rng.normal(size=(300, 5))
    C[:100, 0] = np.abs(C[:100, 0]) + 0.1
    roots = np.full((300, 4), np.nan)
    nr = np.zeros(300, dtype=np.int32)
    for i, c in enumerate(C):
        r = ref.quartic(*c)
        nr[i] = len(r)
        roots[i, :len(r)] = r
    save(write, os.path.join(HERE, "kat_quartic.npz"), coeffs=C, roots=roots, nroots=nr)

    # --- reconstruction -------------------------------------------------------------
    V = np.concatenate([rng.normal(size=(300, 6)),
                        np.cumsum(rng.uniform(0, 1, size=(100, 6)), axis=1),   # monotone
                        np.repeat(rng.normal(size=(50, 1)), 6, axis=1)])       # constant
    d = {"V": V}
    for op in ("PLM_C2L", "PLM_C2R"):
        for th in (1.0, 1.5, 2.0):
            d[f"{op}_theta{th}"] = ref.reconstruct(V, op, theta=th)
    for op in ("WENO5_FV_C2R", "WENO5_FV_C2L", "WENO5_FD_C2R", "WENO5_FD_C2L"):
        for IS in ("js96", "b08", "sz10"):
            d[f"{op}_{IS}"] = ref.reconstruct(V, op, IS=IS, sz10A=100.0)
    save(write, os.path.join(HERE, "kat_reconstruct.npz"), **d)

    # --- Riemann solvers --------------------------------------------------------------
    PL, PR = riemann_pairs(rng, 400)
    d = {"PL": PL, "PR": PR, "gamma": GAMMA}
    for s in ("hll", "hllc", "hlld"):
        for dim in (1, 2, 3):
            F, err, lam, ok = ref.riemann(s, PL, PR, dim, GAMMA)
            d[f"{s}_F{dim}"], d[f"{s}_err{dim}"], d[f"{s}_lam{dim}"] = F, err, lam
            if s == "hlld":
                d[f"hlld_ok{dim}"] = ok
    save(write, os.path.join(HERE, "kat_riemann.npz"), **d)



def write_grids(ref, rng, only):
    # --- full steps on small grids ------------------------------------------------------
    N = 8
    for name, kind, kw, steps in GRID_CASES:
        st = Setup(N=(N, N, N), **kw)
        P0 = grid_state(N, kind, rng)
        if only and name not in only:
            continue
        solver = ref.solver(st)
        P = P0.copy()
        dts, errs, lams, succ, fail = [], [], [], [], []
        dt = 0.0
        for _ in range(steps):
            e, stats = solver.advance(P, dt)
            dts.append(dt)
            errs.append(e)
            lams.append(stats.max_lambda)
            succ.append(stats.hlld_success)
            fail.append(stats.hlld_failure)
            dt = solver.get_timestep(0.4)
            if not np.isfinite(dt):
                dt = 0.0
        setup = {k: v for k, v in st.__dict__.items()}
        np.savez_compressed(os.path.join(HERE, f"grid_{name}.npz"), P0=P0[3:-3, 3:-3, 3:-3],
                            P=P[3:-3, 3:-3, 3:-3], dts=np.array(dts), errors=np.array(errs),
                            max_lambda=np.array(lams), hlld_success=np.array(succ),
                            hlld_failure=np.array(fail), next_dt=dt,
                            setup=np.array(repr(setup)))


if __name__ == "__main__":
    import sys
    main(sys.argv[1].split(",") if len(sys.argv) > 1 else None)
