"""What the mathematics fixes, checked on the oracle (independently of the reference's
own outputs): closed forms, reductions to textbook cases and invariants of the scheme.
Each check is chosen so that a plausible slip (a dropped term, a sign, a swapped index
or axis) fails it.
"""
import numpy as np
import pytest

from oracle.oracle import Setup, synchronize

GAMMA = 4.0 / 3.0


def rand_states(rng, n, Wmax=10.0, bmax=2.0):
    u = rng.normal(size=(n, 3))
    u *= rng.uniform(0, np.sqrt(Wmax**2 - 1), size=(n, 1)) / np.linalg.norm(u, axis=1, keepdims=True)
    W = np.sqrt(1 + (u**2).sum(1))
    P = np.zeros((n, 8))
    P[:, 0] = 10 ** rng.uniform(-1, 1, n)
    P[:, 1] = P[:, 0] * 10 ** rng.uniform(-3, 1, n)
    P[:, 2:5] = u / W[:, None]
    P[:, 5:8] = rng.normal(size=(n, 3)) * bmax
    return P


def conserved_textbook(P, g):
    """U from the covariant definitions: T^{0mu} of ideal RMHD with b^mu = (W v.B,
    B/W + b0 v), h = 1 + g/(g-1) p/rho (the closed form is written independently of
    the oracle's intermediate variables)."""
    rho, p, v, B = P[:, 0], P[:, 1], P[:, 2:5], P[:, 5:8]
    W = 1 / np.sqrt(1 - (v**2).sum(1))
    b0 = W * (v * B).sum(1)
    bvec = B / W[:, None] + b0[:, None] * v
    b2 = (bvec**2).sum(1) - b0**2
    h = 1 + g / (g - 1) * p / rho
    wtot = rho * h + b2
    ptot = p + 0.5 * b2
    D = rho * W
    tau = wtot * W**2 - ptot - b0**2 - D
    S = (wtot * W**2)[:, None] * v - b0[:, None] * bvec
    return np.column_stack([D, tau, S, B])


def test_prim_to_cons_closed_form(oracle_lib):
    P = rand_states(np.random.default_rng(1), 500)
    U, err = oracle_lib.prim_to_cons(P, GAMMA)
    assert (err == 0).all()
    ref = conserved_textbook(P, GAMMA)
    assert np.allclose(U, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_c2p_round_trip(oracle_lib):
    """ConsToPrim(PrimToCons(P)) = P over W 1..50, sigma 0.01..100, p/rho 1e-4..10
    (what luaC_test_rmhd_c2p checks, Mara/mara.cpp:1819-1879)."""
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    from make_golden import sample_states
    P = sample_states(np.random.default_rng(2), 2000)
    U, _ = oracle_lib.prim_to_cons(P, GAMMA)
    Q, err, it = oracle_lib.cons_to_prim(U, GAMMA)
    ok = err == 0
    assert ok.mean() > 0.99
    rel = np.abs(Q[ok] - P[ok]).max(1) / np.abs(P[ok]).max(1)
    assert np.median(rel) < 1e-13 and rel.max() < 1e-8


def test_flux_closed_form(oracle_lib):
    """F^i = T^{i0}: D v^i, S v^i - b^i B/W... written out for all three axes."""
    P = rand_states(np.random.default_rng(3), 300)
    U = conserved_textbook(P, GAMMA)
    rho, p, v, B = P[:, 0], P[:, 1], P[:, 2:5], P[:, 5:8]
    W = 1 / np.sqrt(1 - (v**2).sum(1))
    b0 = W * (v * B).sum(1)
    bvec = B / W[:, None] + b0[:, None] * v
    b2 = (bvec**2).sum(1) - b0**2
    ptot = p + 0.5 * b2
    for d in range(3):
        F, ap, am = oracle_lib.flux_eigen(P, d + 1, GAMMA)
        ref = np.zeros_like(F)
        ref[:, 0] = U[:, 0] * v[:, d]
        ref[:, 1] = U[:, 1] * v[:, d] - b0 * B[:, d] / W + ptot * v[:, d]
        for k in range(3):
            ref[:, 2 + k] = U[:, 2 + k] * v[:, d] - bvec[:, k] * B[:, d] / W + ptot * (k == d)
            ref[:, 5 + k] = v[:, d] * B[:, k] - v[:, k] * B[:, d]
        assert np.allclose(F, ref, rtol=1e-10, atol=1e-10 * np.abs(ref).max())
        assert (am <= ap).all() and (np.abs(ap) <= 1).all() and (np.abs(am) <= 1).all()


def test_fast_speeds_reduce_to_hydro(oracle_lib):
    """With B = 0 the fast speeds are the relativistic sound speeds (v +- cs)/(1 +- v cs)
    along the axis (for motion along that axis)."""
    rng = np.random.default_rng(4)
    n = 200
    P = np.zeros((n, 8))
    P[:, 0] = 10 ** rng.uniform(-1, 1, n)
    P[:, 1] = P[:, 0] * 10 ** rng.uniform(-2, 1, n)
    v = rng.uniform(-0.9, 0.9, n)
    for d in range(3):
        Q = P.copy()
        Q[:, 2 + d] = v
        F, ap, am = oracle_lib.flux_eigen(Q, d + 1, GAMMA)
        h = 1 + GAMMA / (GAMMA - 1) * Q[:, 1] / Q[:, 0]
        cs = np.sqrt(GAMMA * Q[:, 1] / (Q[:, 0] * h))
        assert np.allclose(ap, (v + cs) / (1 + v * cs), atol=1e-7)
        assert np.allclose(am, (v - cs) / (1 - v * cs), atol=1e-7)


def test_quartic_roots_satisfy_polynomial(oracle_lib):
    rng = np.random.default_rng(5)
    for _ in range(300):
        r = np.sort(rng.uniform(-1, 1, 4))
        c = np.poly(r) * rng.uniform(0.5, 2)
        got = np.sort(oracle_lib.quartic(*c))
        assert len(got) == 4
        assert np.allclose(got, r, atol=1e-5)
        assert np.abs(np.polyval(c, got)).max() < 1e-8 * np.abs(c).max()


def test_reconstruction_exactness(oracle_lib):
    """PLM reproduces linear data for theta >= 1; WENO5 reproduces polynomials up to
    degree 2 at the face (every sub-stencil does, whatever the weights); a constant is
    reproduced by both."""
    x = np.arange(-2, 4, dtype=float)  # cell centres of v[-2..3], face at x = 0.5 / -0.5
    rng = np.random.default_rng(6)
    for _ in range(50):
        a, b, c = rng.normal(size=3)
        lin = a + b * x
        for th in (1.0, 1.5, 2.0):
            assert np.isclose(oracle_lib.reconstruct(lin, "PLM_C2R", theta=th)[0], a + b * 0.5)
            assert np.isclose(oracle_lib.reconstruct(lin, "PLM_C2L", theta=th)[0], a - b * 0.5)
        quad = a + b * x + c * x**2
        for IS in ("js96", "b08", "sz10"):
            r = oracle_lib.reconstruct(quad, "WENO5_FV_C2R", IS=IS)[0]
            l = oracle_lib.reconstruct(quad, "WENO5_FV_C2L", IS=IS)[0]
            assert np.isclose(r, a + b * 0.5 + c * 0.25, rtol=1e-12, atol=1e-12)
            assert np.isclose(l, a - b * 0.5 + c * 0.25, rtol=1e-12, atol=1e-12)
    # PLM limits to zero slope at an extremum
    bump = np.array([0.0, 1.0, 2.0, 1.0, 0.0, 0.0])
    assert oracle_lib.reconstruct(bump, "PLM_C2R")[0] == 2.0


@pytest.mark.parametrize("solver", ["hll", "hllc", "hlld"])
def test_riemann_consistency(oracle_lib, solver):
    """Equal left and right states give the physical flux."""
    P = rand_states(np.random.default_rng(7), 200, Wmax=5.0, bmax=1.0)
    for d in (1, 2, 3):
        F, err, lam, ok = oracle_lib.riemann(solver, P, P, d, GAMMA)
        Fp, ap, am = oracle_lib.flux_eigen(P, d, GAMMA)
        assert (err == 0).all()
        scale = np.abs(Fp).max(1, keepdims=True)
        # B_n flux is zero, and every other component matches the physical flux
        assert np.allclose(F, Fp, rtol=1e-8, atol=1e-8 * scale)
        assert np.allclose(lam, np.maximum(np.abs(ap), np.abs(am)))


def test_hlld_reduces_to_hllc_without_normal_field(oracle_lib):
    """With B_n = 0 HLLD skips the p* Newton (estimate (55) directly) and, for purely
    hydrodynamic states, agrees with the HLLC flux of Mignone & Bodo (2006)."""
    rng = np.random.default_rng(8)
    n = 100
    PL = np.zeros((n, 8))
    PL[:, 0] = rng.uniform(0.5, 2, n)
    PL[:, 1] = rng.uniform(0.5, 2, n)
    PL[:, 2] = rng.uniform(-0.3, 0.3, n)
    PR = PL.copy()
    PR[:, 0] *= rng.uniform(0.5, 2, n)
    PR[:, 1] *= rng.uniform(0.5, 2, n)
    Fd, ed, _, okd = oracle_lib.riemann("hlld", PL, PR, 1, GAMMA)
    Fc, ec, _, _ = oracle_lib.riemann("hllc", PL, PR, 1, GAMMA)
    assert (ed == 0).all() and (ec == 0).all()
    assert np.allclose(Fd, Fc, rtol=1e-6, atol=1e-8)


def smooth_grid(N, rng):
    n = N + 6
    x = (np.arange(n) - 3 + 0.5) / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    P = np.zeros((n, n, n, 8))
    P[..., 0] = 1 + 0.2 * np.sin(2 * np.pi * (X + Y))
    P[..., 1] = 1 + 0.2 * np.cos(2 * np.pi * Z)
    P[..., 2] = 0.2 * np.sin(2 * np.pi * Y)
    P[..., 3] = 0.2 * np.sin(2 * np.pi * Z)
    P[..., 4] = 0.2 * np.sin(2 * np.pi * X)
    P[..., 5] = 0.3 * np.sin(2 * np.pi * Y)
    P[..., 6] = 0.3 * np.sin(2 * np.pi * Z)
    P[..., 7] = 1.0 + 0.3 * np.sin(2 * np.pi * X)
    return synchronize(P, 3)


def test_uniform_state_is_stationary(oracle_lib):
    s = Setup(N=(6, 6, 6), extrap="weno5", advance="rk3")
    so = oracle_lib.solver(s)
    P = np.zeros((12, 12, 12, 8))
    P[...] = [1.0, 0.5, 0.2, -0.1, 0.3, 0.4, 0.1, 1.0]
    P0 = P.copy()
    e, _ = so.advance(P, 0.0)
    e, _ = so.advance(P, so.get_timestep(0.4))
    assert e == 0
    assert np.allclose(P, P0, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("extrap,riemann", [("plm", "hlld"), ("weno5", "hllc"), ("pcm", "hll")])
def test_conservation_and_divB(oracle_lib, extrap, riemann):
    """Periodic steps conserve sum(D), sum(tau), sum(S), sum(B) to round-off, and the
    flux-CT update keeps the corner-centred div B (Mara/mara.cpp:2298-2314) unchanged."""
    N = 10
    P = smooth_grid(N, np.random.default_rng(9))
    s = Setup(N=(N, N, N), extrap=extrap, riemann=riemann, advance="rk2")
    so = oracle_lib.solver(s)
    it = (slice(3, -3),) * 3
    U0, _ = oracle_lib.prim_to_cons(P[it].reshape(-1, 8), GAMMA)
    div0 = so.mean_max_divB(P)[2]
    dt = 0.0
    for _ in range(3):
        e, _ = so.advance(P, dt)
        assert e == 0
        dt = so.get_timestep(0.4)
    U1, _ = oracle_lib.prim_to_cons(P[it].reshape(-1, 8), GAMMA)
    tot0, tot1 = U0.sum(0), U1.sum(0)
    scale = np.abs(U0).sum(0)
    assert (np.abs(tot1 - tot0) / scale).max() < 1e-9   # c2p tolerance dominates
    assert abs(so.mean_max_divB(P)[2] - div0) < 1e-11


def test_ghost_and_dt_protocol(oracle_lib):
    """First step has MaxLambda = 0 so dt = inf (the Lua drivers use 0); afterwards
    dt = CFL min dx / max|lambda| and the ghosts hold periodic images."""
    N = 8
    P = smooth_grid(N, np.random.default_rng(10))
    so = oracle_lib.solver(Setup(N=(N, N, N)))
    assert np.isinf(so.get_timestep(0.4))
    e, st = so.advance(P, 0.0)
    assert e == 0 and st.max_lambda > 0
    assert np.isclose(so.get_timestep(0.4), 0.4 * (1.0 / N) / st.max_lambda, rtol=1e-15)
    assert np.array_equal(P, synchronize(P, 3))
