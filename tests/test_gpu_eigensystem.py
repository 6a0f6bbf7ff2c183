"""The device RMHD characteristic eigensystem (weno-split, eigen = 'rmhd') against the
oracle's, which the CPU suite pins by invariants (L R = I, A R = R diag(lam) for an
independently built conserved Jacobian, the known eigenvalues).

The eigenvalues come from closed forms and Newton iterations that both sides evaluate in
the same order, so they agree to round-off. The eigenvectors come from forward-mode
Jacobians and eliminations whose round-off is amplified by 1/gap between neighbouring
eigenvalues (the device runs one state on 8 cooperating lanes, with its own pivot ties and
sum orders); eigenvectors of a cluster are a basis of its eigenspace, fixed only up to that
choice. So the comparison is basis-free: the conserved Jacobian R diag(lam) L, and column
by column (up to sign) only where the spectrum is well separated.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
from make_golden import sample_states  # noqa: E402

G = 4.0 / 3.0


def states(n, seed, Wmax=10.0):
    P = sample_states(np.random.default_rng(seed), n, Wmax=Wmax, sig=(0.01, 10.0), ptr=(1e-2, 10.0))
    base = np.array([1.0, 1.0, 0.3, -0.2, 0.1, 0.5, 0.7, -0.4])
    extra = [base.copy() for _ in range(4)]
    extra[0][5:8] = 0.0                      # no field
    extra[1][5:8] = [0.9, 0.0, 0.0]          # field along x
    extra[2][2:5] = 0.0                      # at rest
    extra[3][1] = 1e-3                       # cold
    return np.vstack([P, np.array(extra)])


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_device_eigensystem_matches_the_oracle(gpu, oracle_lib, dim):
    P = states(3000, seed=11 + dim)
    Lg, Rg, lg, okg = gpu.eigensystem(P, dim, G)
    Lo, Ro, lo, oko = oracle_lib.eigensystem(P, dim, G)
    assert np.mean(okg == oko) >= 0.999
    assert okg.mean() >= 0.99
    both = okg & oko
    # eigenvalues
    assert np.abs(lg[both] - lo[both]).max() < 1e-9
    # L R = I where usable, the identity where not
    I = np.einsum("nij,njk->nik", Lg[okg], Rg[okg])
    assert np.abs(I - np.eye(8)).max() < 1e-8
    assert np.array_equal(Lg[~okg], np.broadcast_to(np.eye(8), Lg[~okg].shape))
    assert np.array_equal(Rg[~okg], np.broadcast_to(np.eye(8), Rg[~okg].shape))
    # the conserved Jacobian R diag(lam) L (basis-free)
    Ag = np.einsum("nij,nj,njk->nik", Rg[both], lg[both], Lg[both])
    Ao = np.einsum("nij,nj,njk->nik", Ro[both], lo[both], Lo[both])
    scale = np.abs(Ao).max(axis=(1, 2))
    rel = np.abs(Ag - Ao).max(axis=(1, 2)) / scale
    assert np.quantile(rel, 0.99) < 1e-7, np.quantile(rel, [0.5, 0.9, 0.99])
    assert rel.max() < 1e-4, np.sort(rel)[-5:]
    # column by column (up to sign) where neighbouring eigenvalues are at least 1e-3 apart
    gap = np.diff(lo[both], axis=1).min(axis=1)
    sep = gap > 1e-3
    assert sep.sum() > 1000
    Rs, Ros = Rg[both][sep], Ro[both][sep]
    sgn = np.sign(np.einsum("nij,nij->nj", Rs, Ros))
    err = np.abs(Rs * sgn[:, None, :] - Ros).max(axis=(1, 2))
    assert np.quantile(err, 0.99) < 1e-6 and err.max() < 1e-4, np.sort(err)[-5:]
