"""CAT allocation Gibbs move (Particle::gibbsPartitionMove, Algorithm 8, Q15
corrected) through the C-ABI: the state it leaves is valid and its likelihood
equals the oracle's; one Gibbs update follows the exact conditional
P(k) ~ n_k L_s(k)^tau (checked against oracle site scores over 3,000 draws)."""
import math
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


def test_gibbs_move_state_and_likelihood(gpu):
    _say("case gibbs: 5 particles, 24 taxa x 120 sites, K in 2..6, 60 iterations")
    rng = np.random.default_rng(8)
    T, S, R, Kmax = 24, 120, 4, 12
    codes = rng.integers(0, 20, size=(T, S)).astype(np.int8)
    batch = sim.prior_particles(rng, 5, T, S, num_rates=R, invar=True, k_range=(2, 6),
                                max_classes=Kmax)
    with Engine(T, S, R, batch.n, Kmax) as eng:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        moved, created, erased = eng.gibbs_partition_move(np.arange(5), 0.8, seed=123, num_iter=60,
                                                          dpp_alpha=1.0)
        lnl, nf = eng.refresh_likelihood()
        states = [eng.particle_state(p) for p in range(5)]
    assert moved > 0 and not nf.any()
    for p, (K, logits, asg) in enumerate(states):
        assert 1 <= K <= Kmax and asg.min() >= 0 and asg.max() < K
        assert np.all(np.bincount(asg, minlength=K) > 0)         # no empty class survives
        q = batch.particle(p)
        q.update(num_classes=K, logits=logits, assignment=asg)
        assert lnl[p] == pytest.approx(orc.particle_lnl(codes, q, R), rel=1e-10)


def test_gibbs_conditional_is_exact(gpu):
    _say("case gibbs conditional: 3000 one-step draws")
    rng = np.random.default_rng(21)
    T, S, R = 8, 10, 2
    col = rng.integers(0, 20, size=T).astype(np.int8)
    codes = np.repeat(col[:, None], S, axis=1)               # every site identical
    base = sim.prior_particles(rng, 1, T, S, num_rates=R, invar=True, k_range=(2, 2))
    a = 6                                                    # class 0 holds 6 sites, class 1 holds 4
    base.assignment[0] = np.array([0] * a + [1] * (S - a), dtype=np.int32)
    n = 3000
    batch = base.take(np.zeros(n, dtype=int))
    tau = 0.7
    with Engine(T, S, R, n, 4) as eng:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        eng.gibbs_partition_move(np.arange(n), tau, seed=99, num_iter=1, dpp_alpha=0.0)
        n0 = np.array([np.sum(eng.particle_state(p)[2] == 0) for p in range(n)])
    l = orc.site_class_lnl(codes, base.particle(0), R, 0, base.logits[0, :2], tau)
    L0, L1 = math.exp(l[0]), math.exp(l[1])
    b = S - a
    p_down = (a / S) * b * L1 / ((a - 1) * L0 + b * L1)
    p_up = (b / S) * a * L0 / (a * L0 + (b - 1) * L1)
    for obs, p in [(np.mean(n0 == a - 1), p_down), (np.mean(n0 == a + 1), p_up)]:
        sd = math.sqrt(p * (1 - p) / n)
        assert abs(obs - p) < 4.5 * sd + 1e-12, (obs, p)
    assert np.all((n0 >= a - 1) & (n0 <= a + 1))
