"""The mixed-precision form against FP64.

The mixed form runs reconstruction and the Riemann fluxes in FP32 and keeps the c2p
Newton solves, the RK accumulation and the CT update in FP64. It is a different method
from the reference, so it is held to accuracy bounds, not to round-off parity:
  * Riemann fluxes (FP32 solvers on the reference's KAT states, Lorentz factors up to 10):
    the same error codes on >= 99 % and finite fluxes everywhere; where both states have
    W <= 3, within 1e-4 relative of the reference's FP64 flux on >= 97 % of interfaces
    (HLL, HLLC) or >= 93 % (HLLD, whose attempt ladder can change branch) and within
    1e-2 on >= 98 %. Above W ~ 3 the four roots of the characteristic quartic crowd
    within ~1/W^2 of the flow speed and FP32 Ferrari roots lose their digits, so the
    mixed form's wave speeds are not held to FP64 there;
  * full steps of turbulence against the oracle's FP64 run on the same seeded data and
    dt sequence: per-field L1 <= 1e-5 and max |div B| at its FP64 round-off (the CT
    update is FP64 on widened FP32 fluxes). The 24^3 grid resolves the shortest
    wavelength of the data with 4 zones, so its interface jumps are an order of
    magnitude larger than at 256^3 (where bench.py measures L1 ~ 2e-7 after 50 steps);
  * no more HLLD-to-HLLC fallbacks than the FP64 run in any step: interfaces whose FP32
    HLLD fails are recomputed in FP64 (the FP32 solve can succeed where FP64 does not).
"""
import numpy as np
import pytest

from conftest import load_golden
from oracle.oracle import Setup

pytestmark = pytest.mark.gpu

GAMMA = 4.0 / 3.0


def lorentz(P):
    return 1.0 / np.sqrt(1.0 - (P[:, 2:5] ** 2).sum(axis=1))


def rel_rows(a, b):
    return np.abs(a - b).max(axis=1) / np.maximum(np.abs(b).max(axis=1), 1e-300)


@pytest.mark.parametrize("solver", ["hll", "hllc", "hlld"])
def test_mixed_riemann_close_to_fp64(gpu, solver):
    g = load_golden("kat_riemann.npz")
    for dim in (1, 2, 3):
        F, err, lam = gpu.riemann(solver, g["PL"], g["PR"], dim, GAMMA, precision="mixed")
        Fr, er = g[f"{solver}_F{dim}"], g[f"{solver}_err{dim}"]
        assert np.mean(err == er) >= 0.99
        ok = (err == 0) & (er == 0)
        assert np.isfinite(F[ok]).all()
        mild = ok & (lorentz(g["PL"]) <= 3.0) & (lorentz(g["PR"]) <= 3.0)
        assert mild.sum() > 150
        rel = rel_rows(F[mild], Fr[mild])
        need = 0.93 if solver == "hlld" else 0.97
        assert np.mean(rel < 1e-4) >= need, (dim, np.quantile(rel, [0.5, 0.9, 0.99]))
        assert np.mean(rel < 1e-2) >= 0.98, (dim, np.sort(rel)[-5:])


@pytest.mark.parametrize("extrap,advance,riemann", [("weno5", "rk3", "hlld"), ("plm", "rk2", "hlld"),
                                                    ("plm", "rk2", "hll"), ("plm", "single", "hlld"),
                                                    ("plm", "rk4", "hlld")])
def test_mixed_steps_close_to_the_oracle(gpu, oracle_lib, extrap, advance, riemann):
    N, steps = 24, 6
    P0 = gpu.init_turbulence((N, N, N), v_rms=0.5, v_kmax=6.0, four_velocity=True, dB_rms=0.5,
                             A_kmax=4.0, seed=31)
    so = oracle_lib.solver(Setup(N=(N, N, N), extrap=extrap, riemann=riemann, advance=advance))
    ctx = gpu.Context(gpu.make_config(N=(N, N, N), extrap=extrap, riemann=riemann, advance=advance,
                                      precision="mixed"))
    ctx.upload(P0)
    Po = P0.copy()
    ref = gpu.Context(gpu.make_config(N=(N, N, N), extrap=extrap, riemann=riemann, advance=advance))
    ref.upload(P0)
    dt = 0.0
    for _ in range(steps):
        eo, _ = so.advance(Po, dt)
        eg, st = ctx.advance(dt)
        er, sr = ref.advance(dt)
        assert eo == 0 and eg == 0 and er == 0
        assert st.hlld_failure <= sr.hlld_failure
        assert st.fp64_redo_overflow == 0
        dt = so.get_timestep(0.4)
        # the FP32 wave speeds give the same CFL step to 1e-4 (the extreme quartic root
        # carries a few FP32 ulps times its conditioning)
        assert abs(ctx.get_timestep(0.4) - dt) <= 1e-4 * dt
    Pg = ctx.download()
    it = (slice(3, -3),) * 3
    a, b = Pg[it], Po[it]
    L1 = np.abs(a - b).sum(axis=(0, 1, 2)) / np.abs(b).sum(axis=(0, 1, 2))
    assert L1.max() <= 1e-5, L1
    _, _, absmax = ctx.divB()
    assert absmax < 1e-11
