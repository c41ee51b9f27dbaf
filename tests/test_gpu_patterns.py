"""Per-class site-pattern compression: identical columns of one CAT class share
one slot whose site log-likelihood counts once per column (the reference evaluates
every column, Alignment.cpp:83, Particle.cpp:22).  On an alignment built from
repeated columns (gaps, unknowns and invariant columns among them) the compressed
engine prunes the number of distinct (class, column) patterns the data hold, and
its likelihoods equal the uncompressed engine's (1e-12) and the oracle's (1e-10),
in full sweeps, in FP32, and after incremental rejuvenation."""
import sys

import numpy as np
import pytest

import oracle_binding as orc
from smccat import Engine
from smccat import simulate as sim

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _repeated_alignment(rng, T=20, unique=60, S=240):
    base = rng.integers(0, 20, size=(T, unique)).astype(np.int8)
    base[:, 0] = 7                     # an invariant column
    base[:, 1] = 20                    # an all-gap column
    base[: T // 2, 2] = 21             # unknowns
    base[:, 3] = 5
    base[4, 3] = 20                    # not invariant: one gap
    cols = rng.integers(0, unique, size=S)
    cols[:unique] = np.arange(unique)  # every column at least once
    rng.shuffle(cols)
    return np.ascontiguousarray(base[:, cols])


def _expected_slots(codes, batch):
    _, pid = np.unique(codes.T, axis=0, return_inverse=True)
    pid = pid.reshape(-1)
    return sum(len(set(zip(batch.assignment[i].tolist(), pid.tolist()))) for i in range(batch.n))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_compressed_patterns_equal_uncompressed_and_oracle(gpu, precision):
    _say("case site patterns %s: 20 taxa x 240 columns of 60 distinct, 10 particles K 1..5" % precision)
    rng = np.random.default_rng(71)
    codes = _repeated_alignment(rng)
    T, S, R = codes.shape[0], codes.shape[1], 4
    batch = sim.prior_particles(rng, 10, T, S, num_rates=R, invar=True, k_range=(1, 5))
    out = {}
    for compress in (True, False):
        with Engine(T, S, R, batch.n, 5, precision) as eng:
            eng.set_patterns(compress)
            eng.set_alignment(codes)
            eng.set_particles(batch)
            out[compress] = (eng.refresh_likelihood()[0], eng.pattern_count())
    (lc, nc), (lu, nu) = out[True], out[False]
    _say("  slots compressed %d, uncompressed %d" % (nc, nu))
    assert nu == batch.n * S
    assert nc == _expected_slots(codes, batch) and nc < 0.75 * nu
    np.testing.assert_allclose(lc, lu, rtol=1e-12)
    if precision == "fp64":
        want = np.array([orc.particle_lnl(codes, batch.particle(i), R) for i in range(batch.n)])
        np.testing.assert_allclose(lc, want, rtol=1e-10)


def test_compressed_patterns_after_rejuvenation(gpu):
    _say("case site patterns: rejuvenation of 8 particles on repeated columns")
    rng = np.random.default_rng(72)
    codes = _repeated_alignment(rng, T=14, unique=40, S=160)
    T, S, R = codes.shape[0], codes.shape[1], 4
    batch = sim.prior_particles(rng, 8, T, S, num_rates=R, invar=True, k_range=(1, 4))
    with Engine(T, S, R, batch.n, 4, workspaces=2) as eng:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        assert eng.pattern_count() < batch.n * S
        lnl, stats = eng.rejuvenate(0.8, 3, seed=9, invariant=True, dpp_alpha=1.0)
        final = eng.snapshot(range(batch.n), batch)
        again, _ = eng.refresh_likelihood()
    assert stats[:, 1].sum() > 0
    np.testing.assert_allclose(lnl, again, rtol=1e-12)
    want = np.array([orc.particle_lnl(codes, final.particle(i), R) for i in range(batch.n)])
    np.testing.assert_allclose(lnl, want, rtol=1e-10)
