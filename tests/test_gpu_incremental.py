"""Incremental refresh with accept/reject (Particle::refreshLikelihood(false),
accept, reject: Particle.cpp:343-478, 530-675) through the C-ABI.  After every
kind of move the incremental lnL equals the oracle's full recomputation of the
new state (1e-10); after a reject the committed lnL and partials come back."""
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


def _parents(children, T):
    par = np.full(2 * T - 1, -1)
    for k, (a, b) in enumerate(children):
        par[a] = par[b] = T + k
    return par


def _path_flags(flags, par, n):
    while n >= 0:
        flags[n] |= 1
        n = par[n]


def propose(rng, batch, i, T, kind):
    """Applies one move to particle i of batch in place; returns (node_flags, eig_flag)."""
    nn = 2 * T - 1
    flags = np.zeros(nn, dtype=np.uint8)
    ch = batch.children[i]
    par = _parents(ch, T)
    root = int(np.nonzero(par[T:] < 0)[0][0]) + T
    nonroot = [n for n in range(nn) if n != root]
    if kind == "branch":                        # Tree::scaleBranchMove
        n = int(rng.choice(nonroot))
        batch.branch_len[i, n] *= np.exp(rng.random() - 0.5)
        flags[n] |= 2
        _path_flags(flags, par, par[n])
    elif kind == "subtree":                     # Tree::scaleSubtreeMove
        n = int(rng.choice(nonroot))
        sc = np.exp(rng.random() - 0.5)
        stack = [n]
        while stack:
            q = stack.pop()
            batch.branch_len[i, q] *= sc
            flags[q] |= 3
            if q >= T:
                stack.extend(int(c) for c in ch[q - T])
        _path_flags(flags, par, par[n])
    elif kind == "nni":                         # Tree::NNIMove
        cands = [n for n in range(T, nn) if n != root]
        p = int(rng.choice(cands))
        a = par[p]
        n1 = int(ch[p - T][rng.integers(2)])
        n2 = int([c for c in ch[a - T] if c != p][0])
        ch[p - T] = [n2 if c == n1 else c for c in ch[p - T]]
        ch[a - T] = [n1 if c == n2 else c for c in ch[a - T]]
        _path_flags(flags, _parents(ch, T), p)
    elif kind == "spr":                         # Tree::SPRMove, both affected paths marked
        cands = [n for n in range(T, nn) if n != root]
        p = int(rng.choice(cands))
        sub = set()
        stack = [p]
        while stack:
            q = stack.pop()
            sub.add(q)
            if q >= T:
                stack.extend(int(c) for c in ch[q - T])
        targets = [n for n in range(nn) if n not in sub and n != root and n != par[p]]
        if not targets:
            return flags, 0
        s_child = int(ch[p - T][rng.integers(2)])
        sib = int([c for c in ch[p - T] if c != s_child][0])
        old = par[p]
        tgt = int(rng.choice(targets))
        tpar = par[tgt]
        ch[old - T] = [sib if c == p else c for c in ch[old - T]]      # prune p
        ch[p - T] = [s_child, tgt]                                     # regraft above tgt
        if tpar == old:
            tpar = old
        ch[tpar - T] = [p if c == tgt else c for c in ch[tpar - T]]
        np2 = _parents(ch, T)
        _path_flags(flags, np2, p)
        _path_flags(flags, np2, old)
    elif kind == "shape":                       # Particle::shapeMove
        batch.shape[i] *= np.exp(rng.random() - 0.5)
        flags[:] = 3
    elif kind == "rates":                       # Particle::rateMatrixMove (single entry)
        batch.base[i, rng.integers(190)] += 0.25 * rng.normal()
        return np.full(nn, 3, dtype=np.uint8), 1
    elif kind == "invar":                       # Particle::invarMove: root mixture only
        batch.pinv[i] = min(0.95, max(0.01, batch.pinv[i] + 0.1 * (rng.random() - 0.5)))
    return flags, 0


def _oracle(codes, batch, R):
    return np.array([orc.particle_lnl(codes, batch.particle(i), R) for i in range(batch.n)])


@pytest.mark.parametrize("variant", ["mfma", "mfma_levels", "mfma_walk", "mfma-u1", "mfma-u4", "valu"])
def test_incremental_moves_and_reject(gpu, variant):
    """Incremental refreshes against the oracle; the mfma-uN cases run every launch
    (per-kind, merged small buckets, roots) with N column blocks per step.  With 30
    taxa every dirty set is small enough to run as chains ("mfma"); "mfma_levels"
    launches the same nodes level by level."""
    _say("case incremental: 7 move kinds, accept/reject, %s" % variant)
    rng = np.random.default_rng(5)
    T, S, R = 30, 150, 4
    codes = rng.integers(0, 20, size=(T, S)).astype(np.int8)
    codes[rng.random(codes.shape) < 0.04] = 21
    kinds = ["branch", "subtree", "nni", "spr", "shape", "rates", "invar"]
    n = len(kinds) * 2
    batch = sim.prior_particles(rng, n, T, S, num_rates=R, invar=True, k_range=(1, 4))
    with Engine(T, S, R, n, 4, workspaces=2) as eng:
        eng.set_prune_variant(variant.split("-")[0])
        if "-u" in variant:
            u = int(variant.split("-u")[1])
            eng.set_tuning(32, (u, u, u), 8, unroll_runtime=u)
        eng.set_alignment(codes)
        eng.set_particles(batch)
        l0, _ = eng.refresh_likelihood()
        np.testing.assert_allclose(l0, _oracle(codes, batch, R), rtol=1e-10)
        committed = batch.take(np.arange(n))
        for round_ in range(3):
            prop = committed.take(np.arange(n))
            flags = np.zeros((n, 2 * T - 1), dtype=np.uint8)
            eig = np.zeros(n, dtype=np.uint8)
            for i in range(n):
                flags[i], eig[i] = propose(rng, prop, i, T, kinds[(i + round_) % len(kinds)])
            eng.update_params(prop)
            l1, nf = eng.refresh_incremental(flags, eig)
            assert not nf.any()
            np.testing.assert_allclose(l1, _oracle(codes, prop, R), rtol=1e-10)
            accept = (rng.random(n) < 0.5).astype(np.uint8)
            eng.accept_reject(accept)
            for i in range(n):
                if accept[i]:
                    committed = _replace(committed, prop, i)
            # nothing dirty: the roots recombine the committed partials
            l2, _ = eng.refresh_incremental(np.zeros((n, 2 * T - 1), dtype=np.uint8))
            np.testing.assert_allclose(l2, _oracle(codes, committed, R), rtol=1e-10)
            eng.accept_reject(np.ones(n, dtype=np.uint8))


def _replace(dst, src, i):
    out = dst.take(np.arange(dst.n))
    for f in ("children", "branch_len", "base", "pinv", "shape"):
        getattr(out, f)[i] = getattr(src, f)[i]
    return out


@pytest.mark.parametrize("variant", ["mfma", "mfma_walk"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_chains_equal_level_launches(gpu, precision, variant):
    """Dirty paths run as chains (one launch, the partial carried in registers from
    node to node), or every changed particle walked in post-order in one launch,
    give the per-level launches' lnL bit for bit, round after round of accepted and
    rejected moves; 120 taxa, so tree moves run as chains (one or two of them per
    particle, SPR joining two paths) and whole-tree moves keep the bucket launches
    (walks: whole trees walked too, two stored children read back); CAT classes
    split the tiles."""
    _say("case %s vs level launches %s: 120 taxa x 160 sites, 16 particles" % (variant, precision))
    rng = np.random.default_rng(11)
    T, S, R = 120, 160, 4
    codes = rng.integers(0, 20, size=(T, S)).astype(np.int8)
    codes[rng.random(codes.shape) < 0.04] = 20
    kinds = ["branch", "subtree", "nni", "spr", "spr", "shape", "rates", "invar"]
    n = len(kinds) * 2
    batch = sim.prior_particles(rng, n, T, S, num_rates=R, invar=True, k_range=(1, 6))
    engs = [Engine(T, S, R, n, 6, precision, workspaces=2) for _ in range(2)]
    try:
        engs[0].set_prune_variant(variant)
        engs[1].set_prune_variant("mfma_levels")
        for eng in engs:
            eng.set_alignment(codes)
            eng.set_particles(batch)
        l0 = [eng.refresh_likelihood()[0] for eng in engs]
        assert np.array_equal(l0[0], l0[1])
        committed = batch.take(np.arange(n))
        for round_ in range(4):
            prop = committed.take(np.arange(n))
            flags = np.zeros((n, 2 * T - 1), dtype=np.uint8)
            eig = np.zeros(n, dtype=np.uint8)
            for i in range(n):
                flags[i], eig[i] = propose(rng, prop, i, T, kinds[(i + round_) % len(kinds)])
            got = []
            for eng in engs:
                eng.update_params(prop)
                l1, nf = eng.refresh_incremental(flags, eig)
                assert not nf.any()
                got.append(l1)
            assert np.array_equal(got[0], got[1]), round_
            if precision == "fp64" and round_ == 0:
                np.testing.assert_allclose(got[0], _oracle(codes, prop, R), rtol=1e-10)
            accept = (rng.random(n) < 0.5).astype(np.uint8)
            for eng in engs:
                eng.accept_reject(accept)
            for i in range(n):
                if accept[i]:
                    committed = _replace(committed, prop, i)
        # nothing dirty: the roots recombine each engine's committed partials
        l3 = [eng.refresh_incremental(np.zeros((n, 2 * T - 1), dtype=np.uint8))[0] for eng in engs]
        assert np.array_equal(l3[0], l3[1])
        if precision == "fp64":
            np.testing.assert_allclose(l3[0], _oracle(codes, committed, R), rtol=1e-10)
    finally:
        for eng in engs:
            eng.close()
