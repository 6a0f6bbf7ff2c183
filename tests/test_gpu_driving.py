"""Driving and cooling on the device against the oracle (pinned bitwise to the reference's
DrivingProcedure and CoolingModuleT4/E4).

The device evaluates the field at every zone as a product of per-axis factors
exp(2 pi i n x) instead of the exponential of the summed phase, so the sampled field agrees
with the reference's to round-off (1e-13 of the summed mode amplitudes); the driven and
cooled states then follow the plain-form tolerances."""
import numpy as np
import pytest

from oracle.oracle import Setup

pytestmark = pytest.mark.gpu

G = 4.0 / 3.0


def compare(a, b, l1=1e-10, mx=1e-7):
    it = (slice(3, -3),) * 3
    a, b = a[it], b[it]
    L1 = np.abs(a - b).sum(axis=(0, 1, 2)) / np.maximum(np.abs(b).sum(axis=(0, 1, 2)), 1e-300)
    assert L1.max() < l1, L1
    scale = np.maximum(np.abs(b).max(axis=(0, 1, 2)), 1e-300)
    assert (np.abs(a - b).max(axis=(0, 1, 2)) / scale).max() < mx


def test_sampled_field_vs_oracle(gpu, oracle_lib):
    f = gpu.OuField.create(0.8, 0.2, 3, 4)
    for _ in range(20):
        f.advance(0.01)
    data = f.bytes()
    xyz = np.random.default_rng(0).uniform(-0.2, 1.2, size=(4000, 3))
    Fd, Fo = f.sample(xyz), oracle_lib.ou_sample(data, xyz)
    # the scale of the sum: the summed magnitudes of the mode amplitudes
    numk = 7 ** 3
    raw = np.frombuffer(data[20:20 + 6 * numk * 16], dtype=np.complex128)
    amp = np.abs(raw[3 * numk:6 * numk]).reshape(3, numk).sum(axis=1)
    assert (np.abs(Fd - Fo) <= 1e-13 * amp[None, :]).all(), np.abs(Fd - Fo).max(axis=0) / amp


@pytest.mark.parametrize("N", [(16, 16, 16), (12, 20, 8)], ids=["16^3", "12x20x8"])
@pytest.mark.parametrize("cool", [None, "T4", "E4"])
def test_driven_cooled_steps_vs_oracle(gpu, oracle_lib, cool, N):
    s = Setup(N=N, extrap="plm", riemann="hlld", advance="rk2")
    so = oracle_lib.solver(s)
    P0 = gpu.init_turbulence(N, v_rms=0.3, four_velocity=True, dB_rms=0.3, seed=21)
    f = gpu.OuField.create(2.0, 0.3, 2, 77)
    ctx = gpu.Context(gpu.make_config(N=N, extrap="plm", riemann="hlld", advance="rk2"))
    ctx.upload(P0)
    ref = 1e-18 if cool == "T4" else 0.4
    ctx.set_cooling(cool, ref, 50.0)
    Po = P0.copy()
    dt = 0.0
    removed_o = 0.0
    for step in range(4):
        if step % 2 == 0:
            ctx.set_driving(f)                      # resample
            sampled = f.bytes()
        eg, _ = ctx.advance(dt)
        eo, _ = so.advance(Po, dt)
        assert eg == eo == 0
        oracle_lib.drive(so, sampled, Po, dt)
        if cool:
            removed_o += oracle_lib.cool(so, cool, ref, 50.0, Po, dt)
        f.advance(0.05)
        dt = so.get_timestep(0.4)
        assert abs(ctx.get_timestep(0.4) - dt) <= 1e-12 * dt
    compare(ctx.download(), Po)
    removed = ctx.cooling_energy_removed()
    if cool:
        assert removed > 0 and abs(removed - removed_o) <= 1e-11 * abs(removed_o)
        assert ctx.cooling_energy_removed() == 0.0              # read resets
    else:
        assert removed == 0.0
    # the power of the sampled field
    pw_o = oracle_lib.drive(so, sampled, Po.copy(), 0.0)
    assert abs(ctx.driving_power() - pw_o) <= 1e-12 * pw_o
    ctx.close()


def test_driving_verbs_like_drvtrb(gpu, oracle_lib):
    """The drvtrb loop (ctf-core/drvtrb.lua:480-488, 728-733): new_ou_field, set_driving,
    advance, driving.Advance every step and driving.Resample every other one, cooling."""
    from ctf_amd.mara import Mara
    N = 12
    Mara.start()
    Mara.set_fluid("rmhd")
    Mara.set_eos("gamma-law", G)
    Mara.set_godunov("plm-split")
    Mara.set_riemann("hlld")
    Mara.set_advance("rk2")
    Mara.set_boundary("periodic")
    Mara.set_domain((0, 0, 0), (1, 1, 1), (N, N, N), 8, 3)
    field = Mara.new_ou_field(3, 1.5, 0.2, 2, 31)
    Mara.set_driving(field)
    Mara.set_cooling("E4", 0.5, 80.0)
    so = oracle_lib.solver(Setup(N=(N, N, N), extrap="plm", riemann="hlld", advance="rk2"))
    P = gpu.init_turbulence((N, N, N), v_rms=0.2, four_velocity=True, dB_rms=0.2, seed=4)
    Po = P.copy()
    sampled = field
    dt, removed_o = 0.0, 0.0
    for it in range(4):
        kz, ne = Mara.advance(P, dt)
        eo, _ = so.advance(Po, dt)
        assert ne == eo == 0
        oracle_lib.drive(so, sampled, Po, dt)
        removed_o += oracle_lib.cool(so, "E4", 0.5, 80.0, Po, dt)
        Mara.driving.Advance(dt)
        field = oracle_lib.ou_advance(field, dt, 1)
        if it % 2 == 1:
            Mara.driving.Resample()
            sampled = field
        assert Mara.driving.Serialize() == field
        dt = Mara.get_timestep(0.4)
    compare(P, Po)
    rate = Mara.cooling_rate(0.5)
    assert abs(rate - removed_o / 0.5) <= 1e-11 * abs(removed_o / 0.5)
    Mara.close()
