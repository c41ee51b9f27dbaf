"""Batched MCMC rejuvenation (main.cpp:293-313, Mcmc.cpp:25-73) through the C-ABI:
the final lnL equals the oracle's on the final state; at tempering 0 the chain
keeps the gamma-shape prior Gamma(3, scale 2) (a missing log-scale Hastings
term would drift it towards Gamma(2, scale 2)); at tempering 1 it climbs."""
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


def test_rejuvenation_consistent_with_oracle(gpu):
    _say("case rejuvenation: 12 particles, 20 taxa x 100 sites, CAT K<=4, 6 iterations")
    rng = np.random.default_rng(12)
    T, S, R = 20, 100, 4
    codes = rng.integers(0, 20, size=(T, S)).astype(np.int8)
    batch = sim.prior_particles(rng, 12, T, S, num_rates=R, invar=True, k_range=(1, 4))
    with Engine(T, S, R, batch.n, 6, workspaces=2) as eng:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        lnl, stats = eng.rejuvenate(0.5, 6, seed=7, gibbs_probability=0.5, dpp_alpha=0.5)
        final = eng.snapshot(range(batch.n), batch)
        again, _ = eng.refresh_likelihood()
    assert stats[:, 0].sum() > 0 and 0 < stats[:, 1].sum() < stats[:, 0].sum()
    np.testing.assert_allclose(again, lnl, rtol=1e-12)
    want = np.array([orc.particle_lnl(codes, final.particle(i), R) for i in range(batch.n)])
    np.testing.assert_allclose(lnl, want, rtol=1e-10)


def test_rejuvenation_keeps_prior_at_zero_tempering(gpu):
    _say("case rejuvenation prior: 512 particles at T=0")
    rng = np.random.default_rng(13)
    T, S, R, N = 10, 40, 4, 512
    codes = rng.integers(0, 20, size=(T, S)).astype(np.int8)
    batch = sim.prior_particles(rng, N, T, S, num_rates=R, invar=True)
    with Engine(T, S, R, N, 1, workspaces=2) as eng:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        eng.rejuvenate(0.0, 40, seed=3, shape_delta=2.0)
        shapes = np.array([eng.particle_params(i)[4] for i in range(N)])
    se = np.sqrt(12.0 / N)
    assert abs(shapes.mean() - 6.0) < 4.5 * se, shapes.mean()


def test_rejuvenation_climbs_at_full_tempering(gpu):
    _say("case rejuvenation climb: 64 particles at T=1")
    codes, cfg = sim.simulate_config("c1")
    rng = np.random.default_rng(14)
    from smccat import lg_base_matrix
    batch = sim.prior_particles(rng, 64, 16, 200, num_rates=4, lg=True, lg_table=lg_base_matrix())
    with Engine(16, 200, 4, 64, 1, workspaces=2) as eng:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        l0, _ = eng.refresh_likelihood()
        l1, stats = eng.rejuvenate(1.0, 20, seed=4, using_lg=True, invariant=False)
    assert np.median(l1) > np.median(l0) + 10.0


def test_stream_keys_fix_the_chains_and_their_step_counts(gpu):
    """A chain depends on its particle and its stream key, not on its slot: four
    particles moved with keys 10..13 end as the same particles in reversed slots with
    the keys reversed.  smc_rejuvenation_steps predicts every chain's proposals."""
    _say("case stream keys: 14 taxa x 90 sites GTR+G4+I, 4 particles")
    rng = np.random.default_rng(41)
    T, S, R, n = 14, 90, 4, 4
    codes = rng.integers(0, 20, size=(T, S)).astype(np.int8)
    batch = sim.prior_particles(rng, n, T, S, num_rates=R, invar=True)
    rev = batch.take(np.arange(n)[::-1])
    keys = np.arange(10, 10 + n, dtype=np.int32)
    with Engine(T, S, R, n, 1, workspaces=2) as eng:
        eng.set_alignment(codes)
        steps = eng.rejuvenation_steps(6, 8, keys)
        eng.set_particles(batch)
        l_a, st_a = eng.rejuvenate(0.6, 6, seed=8, stream_keys=keys)
        a = eng.snapshot(range(n), batch)
        eng.set_particles(rev)
        l_b, st_b = eng.rejuvenate(0.6, 6, seed=8, stream_keys=keys[::-1])
        b = eng.snapshot(range(n), batch)
    np.testing.assert_array_equal(l_a, l_b[::-1])
    for f in ("children", "branch_len", "base", "pinv", "shape"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f)[::-1])
    assert st_a[:, 0].sum() == steps.sum() == st_b[:, 0].sum()


def test_one_slot_engine_with_repeated_whole_tree_moves(gpu):
    """An engine of one slot runs chains whose rejected whole-tree proposals (shape,
    rate matrix: every P of the particle) are followed by more of them: the restores
    deferred to the next refresh merge with the new proposal's recomputes without
    listing a node twice (a one-slot engine has room for each node once)."""
    _say("case rejuvenation: one-slot engine, 16 taxa x 80 sites, 12 iterations, 4 seeds")
    rng = np.random.default_rng(21)
    T, S, R = 16, 80, 4
    codes = rng.integers(0, 20, size=(T, S)).astype(np.int8)
    batch = sim.prior_particles(rng, 1, T, S, num_rates=R, invar=True, k_range=(1, 3))
    for seed in range(4):
        with Engine(T, S, R, 1, 4, workspaces=2) as eng:
            eng.set_alignment(codes)
            eng.set_particles(batch)
            lnl, stats = eng.rejuvenate(1.0, 12, seed=seed, gibbs_probability=0.0)
            final = eng.snapshot([0], batch)
        assert stats[1:, 0].sum() > 0  # whole-tree moves were proposed
        want = orc.particle_lnl(codes, final.particle(0), R)
        np.testing.assert_allclose(lnl[0], want, rtol=1e-10)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_lane_gives_the_same_chains(gpu, precision):
    """Whole-tree steps refreshed on the lane (decided when done, while the other
    particles go on) give the same chains as rounds that wait for every particle:
    the same final lnL, parameters and move statistics, bit for bit."""
    _say("case lane: 24 particles, 18 taxa x 120 sites GTR+G4+I CAT K<=4, 8 iterations, %s" % precision)
    rng = np.random.default_rng(61)
    T, S, R, N = 18, 120, 4, 24
    codes = rng.integers(0, 20, size=(T, S)).astype(np.int8)
    batch = sim.prior_particles(rng, N, T, S, num_rates=R, invar=True, k_range=(1, 4))
    out = []
    for lane in (True, False):
        with Engine(T, S, R, N, 6, precision, workspaces=2) as eng:
            eng.set_alignment(codes)
            eng.set_particles(batch)
            eng.set_lane(lane)
            lnl, stats = eng.rejuvenate(0.7, 8, seed=9, dpp_alpha=0.5)
            params = [eng.particle_params(i) for i in range(N)]
            final = eng.snapshot(range(N), batch)
        out.append((lnl, stats, params, final))
    (l1, s1, p1, f1), (l0, s0, p0, _) = out
    assert s1[3, 0] > 0 and s1[1, 0] > 0  # rate-matrix and shape steps were made
    assert np.array_equal(l1, l0)
    assert np.array_equal(s1, s0)
    for a, b in zip(p1, p0):
        for x, y in zip(a, b):
            assert np.array_equal(np.asarray(x), np.asarray(y))
    if precision == "fp64":
        want = np.array([orc.particle_lnl(codes, f1.particle(i), R) for i in range(N)])
        np.testing.assert_allclose(l1, want, rtol=1e-10)
