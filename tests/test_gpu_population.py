"""Population management on the device: resampled copies (main.cpp:300,309,313;
Particle::copyFromSerialized, Particle.cpp:218-276) as gathers of device rows.

The copied population's likelihoods equal its parents', bit for bit, and its
states are the parents'; rejuvenation after a device copy (which recomputes only
the copied slots and keeps the partials of slots that kept their own state)
leaves exactly the particles that a fresh upload of the same states and a full
refresh leave; and those states match the oracle."""
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


def _setup(seed=12, T=18, S=150, R=4, N=14, kmax=6):
    rng = np.random.default_rng(seed)
    codes = rng.integers(0, 21, size=(T, S)).astype(np.int8)
    batch = sim.prior_particles(rng, N, T, S, num_rates=R, invar=True, k_range=(1, kmax), max_classes=kmax)
    return codes, batch, R


@pytest.mark.parametrize("order", ["systematic", "any"])
def test_copy_particles_gathers_parent_states(gpu, order):
    _say("case device copies: 14 particles, %s parent order" % order)
    codes, batch, R = _setup()
    N, kmax = batch.n, 6
    rng = np.random.default_rng(3)
    if order == "systematic":
        parents = np.sort(rng.integers(0, N, N))
    else:
        parents = rng.integers(0, N, N)
    parents[0] = N - 1  # a slot that is both copied over and a parent of others
    with Engine(batch.children.shape[1] + 1, codes.shape[1], R, N, kmax, workspaces=2) as eng:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        before, _ = eng.refresh_likelihood()
        states = [eng.particle_state(p) for p in range(N)]
        params = [eng.particle_params(p) for p in range(N)]
        eng.copy_particles(parents)
        after, _ = eng.refresh_likelihood()
        assert np.array_equal(after, before[parents])
        for i, p in enumerate(parents):
            K, lg, asg = eng.particle_state(i)
            assert K == states[p][0] and np.array_equal(lg, states[p][1]) and np.array_equal(asg, states[p][2])
            for a, b in zip(eng.particle_params(i), params[p]):
                assert np.array_equal(np.asarray(a), np.asarray(b))


def test_rejuvenation_after_device_copy_equals_fresh_upload(gpu):
    _say("case rejuvenation after device copy: 14 particles, 3 MCMC iterations")
    codes, batch, R = _setup(seed=13)
    N, kmax = batch.n, 6
    T = batch.children.shape[1] + 1
    parents = np.sort(np.random.default_rng(8).integers(0, N, N))
    kept = int(np.sum(parents == np.arange(N)))
    assert 0 < kept < N
    out = {}
    for how in ("device_copy", "fresh_upload"):
        with Engine(T, codes.shape[1], R, N, kmax, workspaces=2) as eng:
            eng.set_alignment(codes)
            eng.set_particles(batch)
            eng.refresh_likelihood()
            if how == "device_copy":
                eng.copy_particles(parents)
            else:
                eng.set_particles(batch.take(parents))
            lnl, stats = eng.rejuvenate(0.7, 3, seed=21, invariant=True, dpp_alpha=1.0)
            out[how] = (lnl, eng.snapshot(range(N), batch))
    a, b = out["device_copy"], out["fresh_upload"]
    assert np.array_equal(a[0], b[0])
    for f in ("children", "branch_len", "base", "pinv", "shape"):
        assert np.array_equal(getattr(a[1], f), getattr(b[1], f)), f
    for i in (0, N // 2, N - 1):
        want = orc.particle_lnl(codes, a[1].particle(i), R)
        assert a[0][i] == pytest.approx(want, rel=1e-10)
