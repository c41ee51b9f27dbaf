"""Proposal tuning from the resampled population (main.cpp:213-289) on the GPU:
the variances of log shape and log branch length, the mean pInvar, and the rate
move's Cholesky factor, against numpy (np.var, np.linalg.cholesky) on the same
population."""
import sys

import numpy as np
import pytest

from smccat import Engine
from smccat import simulate as sim

pytestmark = pytest.mark.gpu


def test_proposal_tuning_matches_numpy(gpu):
    print("case proposal tuning: 300 GTR particles, 11 taxa", flush=True)
    rng = np.random.default_rng(17)
    T, S, R, N = 11, 30, 4, 300
    batch = sim.prior_particles(rng, N, T, S, num_rates=R, invar=True)
    with Engine(T, S, R, N, 1) as eng:
        eng.set_alignment(rng.integers(0, 20, size=(T, S)).astype(np.int8))
        eng.set_particles(batch)
        vs, vb, pm, diag, chol = eng.proposal_tuning(N)
        roots = [_root(batch.children[p], T) for p in range(N)]
    ls = np.log(batch.shape)
    bl = np.concatenate([np.log(np.delete(batch.branch_len[p], roots[p])) for p in range(N)])
    assert vs == pytest.approx(np.var(ls), rel=1e-10)
    assert vb == pytest.approx(np.var(bl), rel=1e-10)
    assert pm == pytest.approx(np.mean(batch.pinv), rel=1e-12)
    M = batch.base.T @ batch.base
    f, shrink = 2.38 ** 2 / 190 * 0.21, 0.2
    want_diag = np.trace(M) * f / 190 * shrink
    assert diag == pytest.approx(want_diag, rel=1e-12)
    A = M * f * (1 - shrink) + np.eye(190) * (want_diag + 1e-6)
    L = np.linalg.cholesky(A)
    np.testing.assert_allclose(chol, L, rtol=1e-9, atol=1e-12)
    assert np.all(np.triu(chol, 1) == 0.0)


def _root(children, T):
    kids = {int(c) for row in children for c in row}
    return [n for n in range(T, 2 * T - 1) if n not in kids][0]
