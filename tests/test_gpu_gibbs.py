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


def _crp_outcomes(groups, kmax, alpha, m=5):
    """Exact distribution of one Algorithm-8 update on uninformative data.

    groups: site labels of the starting partition.  A uniformly drawn site leaves
    its class; with K classes left it joins class k with weight n_k or one of
    m' = min(m, kmax - K) auxiliaries with weight alpha / m' each.  Auxiliary 0 is
    the site's own profile when it was alone in its class (label kept), the others
    are fresh prior profiles (label 'F').  Returns {final labels per site: prob}."""
    S = len(groups)
    out = {}
    for s in range(S):
        rest = [g for i, g in enumerate(groups) if i != s]
        labels = sorted(set(rest))
        sizes = {g: rest.count(g) for g in labels}
        K = len(labels)
        naux = min(m, kmax - K)
        singleton = groups[s] not in sizes
        opts = [(g, float(sizes[g])) for g in labels]
        for a in range(naux):
            opts.append((groups[s] if (singleton and a == 0) else "F", alpha / naux))
        tot = sum(w for _, w in opts)
        for g, w in opts:
            key = tuple(groups[:s]) + (g,) + tuple(groups[s + 1:])
            out[key] = out.get(key, 0.0) + w / (S * tot)
    return out


@pytest.mark.parametrize("kmax", [4, 6])
def test_gibbs_update_at_class_capacity(gpu, kmax):
    """Near max_classes the update is Algorithm 8 with m' = min(5, max_classes - K)
    auxiliaries of weight alpha / m' (new-class mass alpha whatever m'), and none at
    K = max_classes; a class emptied by the removal offers its own profile as the
    first auxiliary.  All-gap columns with pInvar 0 make every site likelihood equal
    (L = K_r under any profile), so the draw follows the CRP weights alone: the
    final partitions of 6,000 one-step moves are checked by chi-square against the
    exact enumeration."""
    from scipy import stats

    _say("case gibbs capacity: max_classes %d, K 4, 6000 one-step draws" % kmax)
    rng = np.random.default_rng(31 + kmax)
    T, S, R, alpha = 8, 6, 2, 1.5
    codes = np.full((T, S), 20, dtype=np.int8)               # every column all '-'
    base = sim.prior_particles(rng, 1, T, S, num_rates=R, invar=True, k_range=(4, 4), max_classes=kmax)
    groups = ["A", "A", "B", "B", "C", "D"]
    base.assignment[0] = np.array([0, 0, 1, 1, 2, 3], dtype=np.int32)
    base.pinv[:] = 0.0
    n = 6000
    batch = base.take(np.zeros(n, dtype=int))
    orig = {tuple(base.logits[0, k]): "ABCD"[k] for k in range(4)}
    with Engine(T, S, R, n, kmax) as eng:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        eng.gibbs_partition_move(np.arange(n), 1.0, seed=77, num_iter=1, dpp_alpha=alpha)
        seen = {}
        for p in range(n):
            K, logits, asg = eng.particle_state(p)
            lab = [orig.get(tuple(logits[k]), "F") for k in range(K)]
            key = tuple(lab[a] for a in asg)
            seen[key] = seen.get(key, 0) + 1
    want = _crp_outcomes(groups, kmax, alpha)
    assert set(seen) <= set(want), set(seen) - set(want)
    if kmax == 4:                                            # no room for a fresh profile
        assert all("F" not in k for k in seen)
    keys = sorted(want, key=lambda k: -want[k])
    big = [k for k in keys if want[k] * n >= 5]
    obs = [seen.get(k, 0) for k in big]
    exp = [want[k] * n for k in big]
    rest_o, rest_e = n - sum(obs), n - sum(exp)
    if rest_e >= 5:
        obs.append(rest_o)
        exp.append(rest_e)
    else:
        exp[-1] += rest_e
        obs[-1] += rest_o
    chi2, pval = stats.chisquare(obs, exp)
    _say("  %d cells, chi2 %.1f, p %.3g" % (len(obs), chi2, pval))
    assert pval > 1e-3, (chi2, pval)
