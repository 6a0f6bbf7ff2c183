"""The oracle's driving field and cooling against the compiled reference.

The reference's StochasticVectorField3d (Mara/ou-field.cpp) with its OU processes and random
streams (Mara/ornuhl.hpp, Mara/random.cpp), DrivingProcedure (Mara/driving.cpp) and
CoolingModuleT4/E4 (Mara/cooling.cpp) run in the reference harness; the oracle restates them
and must agree bitwise: the serialized field after construction and after advances, the
sampled field, the driven and cooled primitives and the energy removed. Skipped where the
reference tree is absent."""
import numpy as np
import pytest

from oracle.oracle import Setup

G = 4.0 / 3.0


def field(lib, k1=2, seed=12345, zeta=0.2, P0=0.8, advances=0, dt=0.01):
    f = lib.ou_new(P0, zeta, k1, seed)
    if advances:
        f = lib.ou_advance(f, dt, advances)
    return f


def test_field_bytes_match_the_reference(oracle_lib, ref_lib):
    for k1, seed in ((1, 7), (2, 12345), (3, 99)):
        a, b = field(oracle_lib, k1, seed), field(ref_lib, k1, seed)
        assert len(a) == 4 + 16 + 6 * (2 * k1 + 1) ** 3 * 16 + (2 * k1 + 1) ** 3 * 40
        assert a == b
        a, b = oracle_lib.ou_advance(a, 0.013, 5), ref_lib.ou_advance(b, 0.013, 5)
        assert a == b
        # the field moved
        assert a != field(oracle_lib, k1, seed)


def test_sampled_field_matches_the_reference(oracle_lib, ref_lib):
    f = field(oracle_lib, 3, 4, advances=20)
    xyz = np.random.default_rng(0).uniform(-0.2, 1.2, size=(500, 3))
    Fa, Fb = oracle_lib.ou_sample(f, xyz), ref_lib.ou_sample(f, xyz)
    assert np.array_equal(Fa, Fb)
    # a real, periodic, divergence-free-leaning field: periodic images agree to round-off
    Fc = oracle_lib.ou_sample(f, xyz + np.array([1.0, -1.0, 2.0]))
    assert np.abs(Fc - Fa).max() < 1e-12 * np.abs(Fa).max()
    assert np.abs(Fa).max() > 0


@pytest.mark.parametrize("kind", ["T4", "E4"])
def test_drive_and_cool_match_the_reference(oracle_lib, ref_lib, kind):
    from ctf_amd import native
    N = 6
    s = Setup(N=(N, N, N))
    so, sr = oracle_lib.solver(s), ref_lib.solver(s)
    P = native.init_turbulence((N, N, N), v_rms=0.4, dB_rms=0.2, seed=3)
    f = field(oracle_lib, 2, 5, advances=10, dt=0.05)
    Pa, Pb = P.copy(), P.copy()
    pa = oracle_lib.drive(so, f, Pa, 0.02)
    pb = ref_lib.drive(sr, f, Pb, 0.02)
    assert np.array_equal(Pa, Pb) and pa == pb
    assert not np.array_equal(Pa, P)
    ref = 1e-18 if kind == "T4" else 0.5
    ea = oracle_lib.cool(so, kind, ref, 100.0, Pa, 0.02)
    eb = ref_lib.cool(sr, kind, ref, 100.0, Pb, 0.02)
    assert np.array_equal(Pa, Pb) and ea == eb
    assert ea > 0


def test_drive_keeps_the_rest_mass_per_zone(oracle_lib):
    # Drive_rmhd scales rho by W0 / W1 so that D = rho W is unchanged
    from ctf_amd import native
    N = 5
    s = Setup(N=(N, N, N))
    so = oracle_lib.solver(s)
    P = native.init_turbulence((N, N, N), v_rms=0.5, four_velocity=True, seed=8)
    f = field(oracle_lib, 2, 1, advances=5, dt=0.1)
    Q = P.copy()
    oracle_lib.drive(so, f, Q, 0.05)
    W = lambda A: 1 / np.sqrt(1 - (A[..., 2:5] ** 2).sum(-1))
    assert np.allclose(P[..., 0] * W(P), Q[..., 0] * W(Q), rtol=1e-14)
