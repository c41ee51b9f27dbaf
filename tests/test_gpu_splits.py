"""Split posteriors and the greedy consensus on the GPU (computeSplitPosteriors,
Miscellaneous.hpp:115-150; Tree::getSplitSet / getSplitBranchMap, Tree.cpp:485-535;
main.cpp:338-383) against the CPU split oracle: the same distinct splits, their
posteriors and weighted mean lengths, and the same consensus split set."""
import sys

import numpy as np
import pytest

import splits_oracle as so
from smccat import Engine
from smccat import simulate as sim

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _to_int(words):
    v = 0
    for k, w in enumerate(words):
        v |= int(w) << (64 * k)
    return v


def _population(rng, N, T, copies=0):
    batch = sim.prior_particles(rng, N, T, 4, num_rates=1)
    for i in range(1, copies + 1):  # some particles share particle 0's tree
        batch.children[i] = batch.children[0]
        batch.branch_len[i] = batch.branch_len[0]
    return batch


def _engine(batch, T):
    eng = Engine(T, 4, 1, batch.n, 1)
    eng.set_alignment(np.zeros((T, 4), dtype=np.int8))
    eng.set_particles(batch)
    return eng


@pytest.mark.parametrize("N,T,copies", [(48, 13, 6), (1024, 200, 20)])
def test_split_posteriors_and_consensus_match_oracle(gpu, N, T, copies):
    _say("case split posteriors: %d particles, %d taxa" % (N, T))
    rng = np.random.default_rng(N + T)
    batch = _population(rng, N, T, copies)
    w = rng.random(N) + 0.05
    w /= w.sum()
    with _engine(batch, T) as eng:
        words, post, mlen = eng.split_posteriors(w)
        cw, cpost, clen = eng.consensus(w)
    trees = [(batch.children[i], batch.branch_len[i]) for i in range(N)]
    want = so.posteriors(trees, w, T)
    got = {_to_int(words[u]): (post[u], mlen[u]) for u in range(len(post))}
    _say("  %d distinct splits" % len(got))
    assert set(got) == set(want)
    for s, (p, ln) in want.items():
        assert got[s][0] == pytest.approx(p, rel=1e-12, abs=1e-15)
        assert got[s][1] == pytest.approx(ln, rel=1e-12)
    order = sorted(want, key=lambda s: -want[s][0])
    kept = set(so.greedy_consensus(order, T))
    assert {_to_int(x) for x in cw} == kept
    assert len(kept) == 2 * T - 3 or len(kept) == len(order)
    assert np.all(np.diff(cpost) <= 1e-12)  # kept in decreasing posterior


def test_identical_trees_give_posterior_one(gpu):
    _say("case split posteriors: 32 copies of one tree")
    rng = np.random.default_rng(5)
    T, N = 17, 32
    batch = _population(rng, N, T, copies=N - 1)
    with _engine(batch, T) as eng:
        words, post, mlen = eng.split_posteriors(np.full(N, 1.0 / N))
        cw, cpost, clen = eng.consensus(np.full(N, 1.0 / N))
    own = so.splits(batch.children[0], batch.branch_len[0], T)
    assert len(post) == 2 * T - 3
    np.testing.assert_allclose(post, 1.0, rtol=1e-13)
    assert {_to_int(x) for x in cw} == set(own)
    for x, ln in zip(cw, clen):
        assert ln == pytest.approx(own[_to_int(x)], rel=1e-12)
