"""The learned-c2p network's CPU oracle, pinned by mathematics: the back-propagated
gradient of the exact (FP64) network equals a centred finite difference of its loss, and
the bfloat16 rounding is round-to-nearest-even on known bit patterns."""
import numpy as np

from oracle import mlp_oracle as mo


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-8, 1.0 + 2**-9, -2.5, 0.0, 3.140625],
                 dtype=np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1); 1 + 3*2^-8 ties up to 1 + 2^-6
    want = np.array([1.0, 1.0, 1.0 + 2**-6, 1.0, -2.5, 0.0, 3.140625], dtype=np.float32)
    assert np.array_equal(mo.bf16(x), want)
    r = np.random.default_rng(0).normal(size=1000).astype(np.float32)
    b = mo.bf16(r)
    assert np.all((b.view(np.uint32) & 0xFFFF) == 0)
    assert np.max(np.abs(b - r) / np.abs(r)) <= 2.0**-8


def test_exact_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    theta = (rng.normal(size=mo.NPARAM) * 0.3).astype(np.float64)
    X = rng.normal(size=(32, 6))
    T = rng.normal(size=(32, 2))
    loss, g = mo.loss_and_grad(theta, X, T, exact=True)
    idx = rng.choice(mo.NPARAM, size=60, replace=False)
    for i in idx:
        h = 1e-6
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (mo.loss_and_grad(tp, X, T, exact=True)[0] - mo.loss_and_grad(tm, X, T, exact=True)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_gradient_close_to_exact():
    rng = np.random.default_rng(2)
    theta = (rng.normal(size=mo.NPARAM) * 0.2).astype(np.float32)
    X = rng.normal(size=(256, 6)).astype(np.float32)
    T = rng.normal(size=(256, 2)).astype(np.float32)
    _, g = mo.loss_and_grad(theta, X, T)
    _, ge = mo.loss_and_grad(theta.astype(np.float64), X, T, exact=True)
    assert np.linalg.norm(g - ge) / np.linalg.norm(ge) < 0.05


def test_split_input_carries_sixteen_bits():
    x = np.random.default_rng(3).normal(size=4000).astype(np.float32) * 3.0
    s = mo.split_input(x)
    assert np.max(np.abs(s - x) / np.abs(x)) <= 2.0**-16
    hi = mo.bf16(x)
    # the hi part alone is the plain bf16 input (error up to 2^-8)
    assert np.max(np.abs(hi - x) / np.abs(x)) > 2.0**-12


def test_c2p_from_the_exact_start_converges_at_once(oracle_lib):
    """anton2dzw started at the solution's own (Z, W) stops within two iterations and
    returns the state; started from the cons-based estimate it agrees."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    from make_golden import sample_states
    G = 4.0 / 3.0
    P = sample_states(np.random.default_rng(5), 500, Wmax=10.0, sig=(0.01, 10.0), ptr=(1e-2, 10.0))
    U, _ = oracle_lib.prim_to_cons(P, G)
    W = 1.0 / np.sqrt(1.0 - (P[:, 2:5] ** 2).sum(1))
    h = 1.0 + G / (G - 1.0) * P[:, 1] / P[:, 0]
    start = np.stack([P[:, 0] * h * W * W, W], axis=1)
    Ps, es, its = oracle_lib.c2p_from_start(U, start, G)
    assert np.all(es == 0) and its.max() <= 2
    assert np.max(np.abs(Ps - P) / (1.0 + np.abs(P))) < 1e-12
    Pe, ee, ite = oracle_lib.c2p_solver(U, 0, G)
    ok = ee == 0
    assert ok.mean() > 0.9 and ite[ok].mean() > 3 * its.mean()
    assert np.max(np.abs(Pe[ok] - P[ok]) / (1.0 + np.abs(P[ok]))) < 1e-10
