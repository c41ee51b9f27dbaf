"""Captured sweeps and launch tuning: a sweep replayed from its hipGraph gives
exactly the eager result, on every replay, with eager work between replays
and after the trees change; every launch shape the autotuner may pick gives the
same per-particle lnL bit for bit, and the tuned engine matches the oracle."""
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


def _setup(precision="fp64", N=16, T=24, S=300, R=4, kmax=1, seed=3, workspaces=1):
    codes, _ = sim.simulate_config("c2", taxa=T, sites=S)
    batch = sim.prior_particles(np.random.default_rng(seed), N, T, S, num_rates=R, invar=True,
                                k_range=(1, kmax))
    eng = Engine(T, S, R, N, kmax, precision, workspaces=workspaces)
    eng.set_alignment(codes)
    eng.set_particles(batch)
    return eng, codes, batch


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replays_equal_eager(gpu, precision):
    _say("case graph replay %s: 16 particles, 24 taxa x 300 sites" % precision)
    eng, codes, batch = _setup(precision)
    with eng:
        eng.set_graphs(False)
        want, _ = eng.refresh_likelihood()
        eng.set_graphs(True)
        for k in range(4):
            got, _ = eng.refresh_likelihood()
            assert np.array_equal(got, want), k
            eng.set_graphs(False)
            part, _ = eng.refresh_likelihood(2, 5)  # eager work between replays
            assert np.array_equal(part, want[2:7])
            eng.set_graphs(True)
        st = eng.graph_stats()
        assert st["launches"] >= 4 and st["captures"] <= 2
        # new trees of the same sizes: the captured launches read the new schedule
        other = sim.prior_particles(np.random.default_rng(99), batch.n, 24, 300, num_rates=4, invar=True)
        eng.set_particles(other)
        got, _ = eng.refresh_likelihood()
        eng.set_graphs(False)
        again, _ = eng.refresh_likelihood()
        assert np.array_equal(got, again)
    if precision == "fp64":
        ref = np.array([orc.particle_lnl(codes, other.particle(i), 4) for i in range(4)])
        np.testing.assert_allclose(got[:4], ref, rtol=1e-10)


def test_graph_cache_one_slot_at_a_time(gpu):
    """One-particle refreshes of alternating slots (the plain path's call shape)
    replay one cached graph per slot and equal the eager sweep on every call."""
    _say("case graph cache: 6 particles refreshed one at a time, 24 taxa x 300 sites")
    eng, codes, batch = _setup(N=6)
    with eng:
        eng.set_graphs(False)
        want, _ = eng.refresh_likelihood()
        eng.set_graphs(True)
        c0 = eng.graph_stats()["captures"]
        for rep in range(3):
            for p in range(6):
                got, _ = eng.refresh_likelihood(p, 1)
                assert np.array_equal(got, want[p:p + 1]), (rep, p)
        st = eng.graph_stats()
        assert st["captures"] - c0 == 6 and st["launches"] >= 18


def test_every_launch_shape_gives_identical_lnl(gpu):
    _say("case launch shapes: CAT K<=6, 12 particles")
    eng, codes, batch = _setup(N=12, kmax=6)
    with eng:
        eng.set_tuning(3, (2, 4, 2), 2, unroll_runtime=4, chain_max_nodes=7)
        assert eng.tuning() == {"blocks_per_cu": 3, "unroll": [2, 4, 2], "xcd_groups": 2, "unroll_runtime": 4,
                                "chain_max_nodes": 7}
        eng.set_graphs(False)
        want, _ = eng.refresh_likelihood()
        for g, x in ((1, 1), (3, 8), (8, 2)):
            for u in ((1, 1, 1), (2, 4, 2), (4, 2, 4)):
                eng.set_tuning(g, u, x)
                for graphs in (False, True):
                    eng.set_graphs(graphs)
                    got, _ = eng.refresh_likelihood()
                    assert np.array_equal(got, want), (g, u, x, graphs)
        # the roots' unroll changes which sites a lane sums (and so the last bits of
        # the per-particle sum), never the per-site values
        for rt in (1, 2, 4):
            eng.set_tuning(8, (1, 1, 1), 8, unroll_runtime=rt)
            eng.set_graphs(False)
            eager, _ = eng.refresh_likelihood()
            np.testing.assert_allclose(eager, want, rtol=1e-13, err_msg="unroll_runtime %d" % rt)
            # a graph captured under another unroll_runtime is never replayed for this one
            eng.set_graphs(True)
            for rep in range(2):
                got, _ = eng.refresh_likelihood()
                assert np.array_equal(got, eager), ("unroll_runtime", rt, rep)


def test_autotune_then_oracle(gpu):
    _say("case autotune: 32 particles")
    eng, codes, batch = _setup(N=32, workspaces=2)
    with eng:
        res = eng.autotune(sweeps=2)
        _say("  %s" % res)
        assert res["configs_tried"] > 5 and res["ms_per_sweep"] > 0
        assert eng.tuning()["blocks_per_cu"] == res["blocks_per_cu"]
        got, _ = eng.refresh_likelihood()
    ref = np.array([orc.particle_lnl(codes, batch.particle(i), 4) for i in range(6)])
    np.testing.assert_allclose(got[:6], ref, rtol=1e-10)
