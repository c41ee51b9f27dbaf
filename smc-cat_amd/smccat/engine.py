"""Python view of the engine's C-ABI (numpy in, numpy out).

`ParticleBatch` holds particle states in the layout smc_set_particles takes;
`Engine` owns one libsmccat_hip engine.  Method names follow the reference's
Particle / SMC vocabulary (refreshLikelihood, computeNextStep, ...).
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


@dataclass
class ParticleBatch:
    """Particle states, one row per particle (SerializedParticle fields, parsed).

    children   int32 [N, T-1, 2]   children of internal node T+i (node ids as Tree.cpp)
    branch_len f64   [N, 2T-1]     by node id (root entry ignored)
    base       f64   [N, 190]      log-exchangeabilities
    num_classes int32 [N]
    logits     f64   [N, Kmax, 20]
    assignment int32 [N, S]
    pinv       f64   [N]
    shape      f64   [N]
    """
    children: np.ndarray
    branch_len: np.ndarray
    base: np.ndarray
    num_classes: np.ndarray
    logits: np.ndarray
    assignment: np.ndarray
    pinv: np.ndarray
    shape: np.ndarray

    @property
    def n(self):
        return self.children.shape[0]

    def take(self, idx):
        idx = np.asarray(idx)
        return ParticleBatch(*(np.ascontiguousarray(getattr(self, f)[idx]) for f in
                               ("children", "branch_len", "base", "num_classes", "logits",
                                "assignment", "pinv", "shape")))

    def particle(self, i):
        """Single particle as a dict (for the oracle)."""
        k = int(self.num_classes[i])
        return dict(children=self.children[i], branch_len=self.branch_len[i], base=self.base[i],
                    num_classes=k, logits=self.logits[i, :k], assignment=self.assignment[i],
                    pinv=float(self.pinv[i]), shape=float(self.shape[i]))


class EngineError(RuntimeError):
    pass


class Engine:
    def __init__(self, taxa, sites, num_rates=1, max_particles=1, max_classes=1,
                 precision="fp64", device=0, partial_slots=0, pmat_classes=0, workspaces=1):
        self.lib = _lib.load()
        cfg = _lib.EngineConfig(device, taxa, sites, num_rates, max_particles, max_classes,
                                _lib.SMC_PREC_FP64 if precision == "fp64" else _lib.SMC_PREC_FP32,
                                partial_slots, pmat_classes, workspaces, 0)
        h = ctypes.c_void_p()
        rc = self.lib.smc_engine_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise EngineError("smc_engine_create failed with status %d" % rc)
        self.h = h
        self.taxa, self.sites, self.num_rates = taxa, sites, num_rates
        self.max_particles, self.max_classes = max_particles, max_classes
        self.precision = precision

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.smc_engine_last_error(self.h)
            raise EngineError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def close(self):
        if getattr(self, "h", None):
            self.lib.smc_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def device_bytes(self):
        return int(self.lib.smc_engine_device_bytes(self.h))

    def set_alignment(self, codes):
        codes = np.ascontiguousarray(codes, dtype=np.int8)
        assert codes.shape == (self.taxa, self.sites)
        self._check(self.lib.smc_set_alignment(self.h, codes.ctypes.data_as(ctypes.POINTER(ctypes.c_int8))),
                    "smc_set_alignment")

    def set_particles(self, batch, first=0):
        n = batch.n
        T, S, K = self.taxa, self.sites, self.max_classes
        ch = np.ascontiguousarray(batch.children, dtype=np.int32).reshape(n, T - 1, 2)
        bl = np.ascontiguousarray(batch.branch_len, dtype=np.float64).reshape(n, 2 * T - 1)
        base = np.ascontiguousarray(batch.base, dtype=np.float64).reshape(n, 190)
        nc = np.ascontiguousarray(batch.num_classes, dtype=np.int32).reshape(n)
        lg = np.zeros((n, K, 20))
        kk = min(K, batch.logits.shape[1])
        lg[:, :kk] = batch.logits[:, :kk]
        asg = np.ascontiguousarray(batch.assignment, dtype=np.int32).reshape(n, S)
        pinv = np.ascontiguousarray(batch.pinv, dtype=np.float64).reshape(n)
        shape = np.ascontiguousarray(batch.shape, dtype=np.float64).reshape(n)
        self._check(self.lib.smc_set_particles(self.h, first, n, _ip(ch), _dp(bl), _dp(base), _ip(nc),
                                               _dp(lg), _ip(asg), _dp(pinv), _dp(shape)),
                    "smc_set_particles")

    def refresh_likelihood(self, first=0, count=None):
        """refreshLikelihood(true) + lnLikelihood() for slots [first, first+count)."""
        if count is None:
            count = self.max_particles - first
        lnl = np.empty(count)
        nf = np.empty(count, dtype=np.uint8)
        self._check(self.lib.smc_refresh_likelihood(self.h, first, count, _dp(lnl),
                                                    nf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))),
                    "smc_refresh_likelihood")
        return lnl, nf

    def compute_next_step(self, raw_weights, raw_lnl, target_ess, last_temp):
        """computeNextStep: returns (new_raw_weights, normalized, ess, current_temp)."""
        w = np.array(raw_weights, dtype=np.float64)
        l = np.ascontiguousarray(raw_lnl, dtype=np.float64)
        norm = np.empty_like(w)
        ess = ctypes.c_double()
        temp = ctypes.c_double()
        self._check(self.lib.smc_compute_next_step(self.h, len(w), _dp(w), _dp(l), _dp(norm),
                                                   ctypes.byref(ess), int(target_ess),
                                                   ctypes.byref(temp), float(last_temp)),
                    "smc_compute_next_step")
        return w, norm, ess.value, temp.value

    def systematic_resample(self, normalized, u):
        w = np.ascontiguousarray(normalized, dtype=np.float64)
        out = np.empty(len(w), dtype=np.int32)
        self._check(self.lib.smc_systematic_resample(self.h, len(w), _dp(w), float(u), _ip(out)),
                    "smc_systematic_resample")
        return out

    def site_class_lnl(self, particles, sites, cand_logits, ncand=None, tempering=1.0):
        """tempering * log L_site under candidate profiles; cand_logits [count, n_cand, 20]."""
        particles = np.ascontiguousarray(particles, dtype=np.int32)
        sites = np.ascontiguousarray(sites, dtype=np.int32)
        cl = np.ascontiguousarray(cand_logits, dtype=np.float64)
        count, n_cand = cl.shape[0], cl.shape[1]
        if ncand is None:
            ncand = np.full(count, n_cand, dtype=np.int32)
        ncand = np.ascontiguousarray(ncand, dtype=np.int32)
        out = np.empty((count, n_cand))
        self._check(self.lib.smc_site_class_lnl(self.h, count, _ip(particles), _ip(sites), n_cand,
                                                _ip(ncand), _dp(cl), float(tempering), _dp(out)),
                    "smc_site_class_lnl")
        return out

    def gibbs_partition_move(self, slots, tempering, seed, num_iter=0, dpp_alpha=0.01):
        """Particle::gibbsPartitionMove for several slots; returns (moved, created, erased)."""
        slots = np.ascontiguousarray(slots, dtype=np.int32)
        stats = (ctypes.c_int64 * 3)()
        self._check(self.lib.smc_gibbs_partition_move(self.h, len(slots), _ip(slots), float(tempering),
                                                      int(seed), int(num_iter), float(dpp_alpha), stats),
                    "smc_gibbs_partition_move")
        return tuple(int(v) for v in stats)

    def particle_state(self, slot):
        """(num_classes, logits [K, 20], assignment [S]) of a slot."""
        nc = ctypes.c_int32()
        lg = np.zeros((self.max_classes, 20))
        asg = np.zeros(self.sites, dtype=np.int32)
        self._check(self.lib.smc_get_particle_state(self.h, int(slot), ctypes.byref(nc), _dp(lg), _ip(asg)),
                    "smc_get_particle_state")
        return nc.value, lg[:nc.value].copy(), asg

    def update_params(self, batch, first=0):
        """New topology / branch lengths / base / pinv / shape (classes untouched)."""
        n, T = batch.n, self.taxa
        ch = np.ascontiguousarray(batch.children, dtype=np.int32).reshape(n, T - 1, 2)
        bl = np.ascontiguousarray(batch.branch_len, dtype=np.float64).reshape(n, 2 * T - 1)
        base = np.ascontiguousarray(batch.base, dtype=np.float64).reshape(n, 190)
        pinv = np.ascontiguousarray(batch.pinv, dtype=np.float64).reshape(n)
        shape = np.ascontiguousarray(batch.shape, dtype=np.float64).reshape(n)
        self._check(self.lib.smc_update_params(self.h, first, n, _ip(ch), _dp(bl), _dp(base), _dp(pinv),
                                               _dp(shape)), "smc_update_params")

    def refresh_incremental(self, node_flags, eig_flags=None, first=0):
        """refreshLikelihood(false): node_flags [count, 2T-1] bit0 updateCL, bit1 updateTP."""
        nf = np.ascontiguousarray(node_flags, dtype=np.uint8)
        count = nf.shape[0]
        ef = None if eig_flags is None else np.ascontiguousarray(eig_flags, dtype=np.uint8)
        lnl = np.empty(count)
        bad = np.empty(count, dtype=np.uint8)
        bp = ctypes.POINTER(ctypes.c_uint8)
        self._check(self.lib.smc_refresh_incremental(self.h, first, count, nf.ctypes.data_as(bp),
                                                     ef.ctypes.data_as(bp) if ef is not None else None,
                                                     _dp(lnl), bad.ctypes.data_as(bp)),
                    "smc_refresh_incremental")
        return lnl, bad

    def copy_particles(self, parents):
        """Slot i of [0, len(parents)) takes the state of slot parents[i] (on the device)."""
        par = np.ascontiguousarray(parents, dtype=np.int32)
        self._check(self.lib.smc_copy_particles(self.h, len(par), _ip(par)), "smc_copy_particles")

    def accept_reject(self, accept, first=0):
        acc = np.ascontiguousarray(accept, dtype=np.uint8)
        self._check(self.lib.smc_accept_reject(self.h, first, len(acc),
                                               acc.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))),
                    "smc_accept_reject")

    def _rejuv_options(self, iterations, seed, using_lg, invariant, gibbs_probability, dpp_alpha,
                       proposal, stream_keys, deltas):
        return _lib.RejuvenationOptions(int(iterations), int(using_lg), int(invariant), 0, int(seed),
                                        deltas.get("scale_delta", 1.0),
                                        deltas.get("subtree_scale_delta", 1.0),
                                        deltas.get("shape_delta", 1.0), deltas.get("rate_delta", 0.25),
                                        deltas.get("rate_full_scale", 0.5),
                                        deltas.get("invar_alpha", 100.0), float(gibbs_probability),
                                        float(dpp_alpha), None, proposal.h if proposal is not None else None,
                                        0, 0, stream_keys.ctypes.data if stream_keys is not None else None)

    def rejuvenate(self, tempering, iterations, seed, first=0, count=None, using_lg=False,
                   invariant=True, gibbs_probability=0.0, dpp_alpha=0.01, proposal=None,
                   stream_keys=None, **deltas):
        """MCMC rejuvenation of slots [first, first+count); returns (lnl, stats[4, 2]).
        proposal: a ProposalNetwork makes the CAT move the learned one; stream_keys
        [count] (optional) keys the slots' random streams instead of the slot ids."""
        if count is None:
            count = self.max_particles - first
        keys = None
        if stream_keys is not None:
            keys = np.ascontiguousarray(stream_keys, dtype=np.int32)
            assert keys.shape == (count,)
        o = self._rejuv_options(iterations, seed, using_lg, invariant, gibbs_probability, dpp_alpha, proposal,
                                keys, deltas)
        lnl = np.empty(count)
        stats = (ctypes.c_int64 * 8)()
        self._check(self.lib.smc_rejuvenate(self.h, first, count, float(tempering), ctypes.byref(o),
                                            _dp(lnl), stats), "smc_rejuvenate")
        return lnl, np.array(list(stats)).reshape(4, 2)

    def rejuvenation_steps(self, iterations, seed, keys, using_lg=False, invariant=True):
        """The number of MH steps rejuvenate() makes for each stream key."""
        keys = np.ascontiguousarray(keys, dtype=np.int32)
        out = np.empty(len(keys), dtype=np.int32)
        o = self._rejuv_options(iterations, seed, using_lg, invariant, 0.0, 0.01, None, None, {})
        self._check(self.lib.smc_rejuvenation_steps(self.h, ctypes.byref(o), len(keys), _ip(keys), _ip(out)),
                    "smc_rejuvenation_steps")
        return out

    def run_sampler(self, num_particles, iterations, seed, using_lg=False, invariant=True, cat=False,
                    threshold=0.6, alg8_probability=0.1, batch=0, max_iterations=0, dpp_alpha=0.01,
                    proposal=None, proposal_warmup=1, train_epochs=8, train_lr=2e-3, on_iteration=None):
        """The SMC loop over slots [0, num_particles) (which hold the initial
        population).  proposal: a ProposalNetwork for the learned CAT move;
        on_iteration(iteration, temp, ess) is called after every reweighting.
        Returns (result dict, final lnl, final raw log weights)."""
        cb = _lib.ITERATION_CALLBACK((lambda u, it, t, e: on_iteration(it, t, e)) if on_iteration
                                     else (lambda u, it, t, e: None))
        o = _lib.SamplerOptions(int(num_particles), int(iterations), float(threshold),
                                float(alg8_probability), int(using_lg), int(invariant), int(cat),
                                int(batch), int(max_iterations), 0, int(seed), float(dpp_alpha),
                                proposal.h if proposal is not None else None, int(proposal_warmup),
                                int(train_epochs), float(train_lr),
                                ctypes.cast(cb, ctypes.c_void_p) if on_iteration else None, None)
        r = _lib.SamplerResult()
        lnl = np.empty(num_particles)
        w = np.empty(num_particles)
        self._check(self.lib.smc_run_sampler(self.h, ctypes.byref(o), ctypes.byref(r), _dp(lnl), _dp(w)),
                    "smc_run_sampler")
        res = {k: getattr(r, k) for k, _ in _lib.SamplerResult._fields_ if k not in ("move_stats", "cat_stats")}
        res["move_stats"] = np.array(list(r.move_stats)).reshape(4, 2)
        res["cat_stats"] = np.array(list(r.cat_stats))
        return res, lnl, w

    def tuning(self):
        t = _lib.Tuning()
        self._check(self.lib.smc_engine_get_tuning(self.h, ctypes.byref(t)), "smc_engine_get_tuning")
        return {"blocks_per_cu": t.blocks_per_cu, "unroll": list(t.unroll), "xcd_groups": t.xcd_groups,
                "unroll_runtime": t.unroll_runtime, "chain_max_nodes": t.chain_max_nodes}

    def set_tuning(self, blocks_per_cu, unroll, xcd_groups=8, unroll_runtime=None, chain_max_nodes=None):
        cur = self.tuning()
        if unroll_runtime is None:
            unroll_runtime = cur["unroll_runtime"]
        if chain_max_nodes is None:
            chain_max_nodes = cur["chain_max_nodes"]
        t = _lib.Tuning(int(blocks_per_cu), (ctypes.c_int32 * 3)(*unroll), 0.0, 0, int(xcd_groups),
                        int(unroll_runtime), int(chain_max_nodes))
        self._check(self.lib.smc_engine_set_tuning(self.h, ctypes.byref(t)), "smc_engine_set_tuning")

    def set_chain_max_nodes(self, n):
        """Incremental refreshes run a particle's dirty set as chains up to n stored nodes."""
        t = self.tuning()
        self.set_tuning(t["blocks_per_cu"], t["unroll"], t["xcd_groups"], t["unroll_runtime"], n)

    def autotune(self, first=0, count=None, sweeps=3):
        """Measured search over the sweep's launch shape; returns the chosen shape."""
        if count is None:
            count = self.max_particles - first
        t = _lib.Tuning()
        self._check(self.lib.smc_engine_autotune(self.h, int(first), int(count), int(sweeps), ctypes.byref(t)),
                    "smc_engine_autotune")
        return {"blocks_per_cu": t.blocks_per_cu, "unroll": list(t.unroll), "xcd_groups": t.xcd_groups,
                "unroll_runtime": t.unroll_runtime, "ms_per_sweep": t.ms_per_sweep,
                "configs_tried": t.configs_tried}

    def set_graphs(self, enable):
        self._check(self.lib.smc_engine_set_graphs(self.h, int(bool(enable))), "smc_engine_set_graphs")

    def graph_stats(self):
        c, l = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.smc_engine_graph_stats(self.h, ctypes.byref(c), ctypes.byref(l)),
                    "smc_engine_graph_stats")
        return {"captures": c.value, "launches": l.value}

    PHASES = ("set_particles", "refresh", "update_params", "refresh_incremental", "accept_reject",
              "propose", "cat_move", "site_class")

    def phase_stats(self, reset=False):
        """{phase: (host ms, calls)} of the engine's entry points."""
        ms = np.zeros(len(self.PHASES))
        calls = np.zeros(len(self.PHASES), dtype=np.int64)
        self._check(self.lib.smc_engine_phase_stats(self.h, _dp(ms), calls.ctypes.data_as(
            ctypes.POINTER(ctypes.c_int64))), "smc_engine_phase_stats")
        if reset:
            self._check(self.lib.smc_engine_reset_phase_stats(self.h), "smc_engine_reset_phase_stats")
        return {k: (float(ms[i]), int(calls[i])) for i, k in enumerate(self.PHASES)}

    def ln_prior(self, slot, using_lg=False, dpp_alpha=0.01):
        """Particle::lnPrior of an uploaded slot."""
        out = ctypes.c_double()
        self._check(self.lib.smc_ln_prior(self.h, int(slot), int(using_lg), float(dpp_alpha),
                                          ctypes.byref(out)), "smc_ln_prior")
        return out.value

    def particle_params(self, slot):
        """(children [T-1, 2], branch_len [2T-1], base [190], pinv, shape) of a slot."""
        T = self.taxa
        ch = np.zeros((T - 1, 2), dtype=np.int32)
        bl = np.zeros(2 * T - 1)
        base = np.zeros(190)
        pv, sh = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.smc_get_particle_params(self.h, int(slot), _ip(ch), _dp(bl), _dp(base),
                                                     ctypes.byref(pv), ctypes.byref(sh)),
                    "smc_get_particle_params")
        return ch, bl, base, pv.value, sh.value

    def snapshot(self, slots, batch_template):
        """ParticleBatch of the engine's current state for the given slots."""
        out = batch_template.take(np.zeros(len(slots), dtype=int))
        out.logits = np.zeros((len(slots), self.max_classes, 20))
        for j, s in enumerate(slots):
            ch, bl, base, pv, sh = self.particle_params(s)
            K, lg, asg = self.particle_state(s)
            out.children[j], out.branch_len[j], out.base[j] = ch, bl, base
            out.pinv[j], out.shape[j], out.num_classes[j] = pv, sh, K
            out.logits[j] = 0.0
            out.logits[j, :K] = lg
            out.assignment[j] = asg
        return out

    def set_prune_variant(self, variant):
        """'mfma' (matrix cores, default: short dirty paths run as chains), 'mfma_levels'
        (dirty nodes launched level by level), 'mfma_walk' (one launch per incremental
        refresh: each tile of a particle's dirty nodes walked in post-order) or 'valu'."""
        v = {"mfma": _lib.SMC_PRUNE_MFMA, "valu": _lib.SMC_PRUNE_VALU,
             "mfma_levels": _lib.SMC_PRUNE_MFMA_LEVELS, "mfma_walk": _lib.SMC_PRUNE_MFMA_WALK}[variant]
        self._check(self.lib.smc_engine_set_variant(self.h, _lib.SMC_K_PRUNE, v), "smc_engine_set_variant")

    def proposal_tuning(self, n, using_lg=False):
        """(var log shape, var log branch length, mean pInvar, shrinkage diagonal,
        rate Cholesky [190, 190] or None) of slots [0, n)."""
        out = np.zeros(4)
        chol = None if using_lg else np.zeros((190, 190))
        self._check(self.lib.smc_proposal_tuning(self.h, int(n), 1 if using_lg else 0, _dp(out),
                                                 None if chol is None else _dp(chol)), "smc_proposal_tuning")
        return out[0], out[1], out[2], out[3], chol

    def split_posteriors(self, weights):
        """(words [U, W] uint64, posterior [U], mean branch length [U]) of the distinct
        splits of slots [0, len(weights)) (computeSplitPosteriors)."""
        w = np.ascontiguousarray(weights, dtype=np.float64)
        W = (self.taxa + 63) // 64
        n = ctypes.c_int64()
        self._check(self.lib.smc_split_posteriors(self.h, len(w), _dp(w), 0, ctypes.byref(n), None, None, None),
                    "smc_split_posteriors")
        words = np.zeros((n.value, W), dtype=np.uint64)
        post = np.zeros(n.value)
        ln = np.zeros(n.value)
        self._check(self.lib.smc_split_posteriors(self.h, len(w), _dp(w), n.value, ctypes.byref(n),
                                                  words.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                                  _dp(post), _dp(ln)), "smc_split_posteriors")
        return words, post, ln

    def consensus(self, weights):
        """(words [k, W], posterior [k], mean length [k]) of the greedy consensus splits."""
        w = np.ascontiguousarray(weights, dtype=np.float64)
        W = (self.taxa + 63) // 64
        m = 2 * self.taxa - 3
        words = np.zeros((m, W), dtype=np.uint64)
        post = np.zeros(m)
        ln = np.zeros(m)
        k = ctypes.c_int32()
        self._check(self.lib.smc_consensus(self.h, len(w), _dp(w), ctypes.byref(k),
                                           words.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), _dp(post),
                                           _dp(ln)), "smc_consensus")
        return words[:k.value], post[:k.value], ln[:k.value]

    def set_patterns(self, compress):
        """Merge identical columns of a class into one weighted slot (default on) for
        particles uploaded afterwards."""
        self._check(self.lib.smc_engine_set_patterns(self.h, 1 if compress else 0), "smc_engine_set_patterns")

    def pattern_count(self, first=0, count=None):
        """Slots (site patterns within classes) the sweep of slots [first, first+count) prunes."""
        if count is None:
            count = self.max_particles - first
        return int(self.lib.smc_pattern_count(self.h, int(first), int(count)))

    def set_mixture(self, mode):
        """'reference' (Particle.cpp:653-670 exactly, default) or 'logspace'."""
        m = {"reference": _lib.SMC_MIX_REFERENCE, "logspace": _lib.SMC_MIX_LOGSPACE}[mode]
        self._check(self.lib.smc_engine_set_mixture(self.h, m), "smc_engine_set_mixture")

    def set_lane(self, on=True):
        """Whole-tree steps of lockstep rejuvenation on a stream of their own
        (default on; the results are the same either way)."""
        self._check(self.lib.smc_engine_set_lane(self.h, 1 if on else 0), "smc_engine_set_lane")

    def set_timing(self, on=True):
        self._check(self.lib.smc_engine_set_timing(self.h, 1 if on else 0), "smc_engine_set_timing")

    def reset_stats(self):
        self._check(self.lib.smc_engine_reset_stats(self.h), "smc_engine_reset_stats")

    def kernel_stats(self):
        out = {}
        for fam, name in enumerate(_lib.KERNEL_FAMILIES):
            n = ctypes.c_int64()
            ms = ctypes.c_double()
            self._check(self.lib.smc_engine_kernel_stats(self.h, fam, ctypes.byref(n), ctypes.byref(ms)),
                        "smc_engine_kernel_stats")
            out[name] = (int(n.value), float(ms.value))
        return out


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _lp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


class ProposalNetwork:
    """The learned CAT-allocation proposal of one engine (smc_mlp_* C-ABI).

    Examples are (header [44], K, label, class entries [K, 24]) as in
    include/smccat_hip.h; collect_gibbs records the Gibbs move's own draws."""

    HDR, CLS = 44, 24

    def __init__(self, engine, seed=0):
        self.engine = engine
        self.lib = engine.lib
        self.h = ctypes.c_void_p()
        engine._check(self.lib.smc_mlp_create(engine.h, int(seed), ctypes.byref(self.h)), "smc_mlp_create")
        self.num_params = int(self.lib.smc_mlp_num_params())

    def close(self):
        if self.h:
            self.lib.smc_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        self.engine._check(rc, what)

    @property
    def params(self):
        th = np.empty(self.num_params, dtype=np.float32)
        self._check(self.lib.smc_mlp_get_params(self.h, _fp(th)), "smc_mlp_get_params")
        return th

    @params.setter
    def params(self, theta):
        th = np.ascontiguousarray(theta, dtype=np.float32)
        assert th.shape == (self.num_params,)
        self._check(self.lib.smc_mlp_set_params(self.h, _fp(th)), "smc_mlp_set_params")

    def add_examples(self, header, num_classes, label, cls_list):
        header = np.ascontiguousarray(header, dtype=np.float32)
        K = np.ascontiguousarray(num_classes, dtype=np.int32)
        lab = np.ascontiguousarray(label, dtype=np.int32)
        off = np.concatenate([[0], np.cumsum([len(c) for c in cls_list])[:-1]]).astype(np.int64)
        cls = np.ascontiguousarray(np.concatenate(cls_list), dtype=np.float32)
        self._check(self.lib.smc_mlp_add_examples(self.h, len(K), _fp(header), _ip(K), _ip(lab), _lp(off),
                                                  _fp(cls), len(cls)), "smc_mlp_add_examples")

    def collect_gibbs(self, slots, tempering, seed, num_iter=0, dpp_alpha=0.01):
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        st = np.zeros(3, dtype=np.int64)
        self._check(self.lib.smc_mlp_collect_gibbs(self.h, len(sl), _ip(sl), float(tempering), int(seed),
                                                   int(num_iter), float(dpp_alpha), _lp(st)),
                    "smc_mlp_collect_gibbs")
        return st

    @property
    def num_examples(self):
        return int(self.lib.smc_mlp_num_examples(self.h))

    def clear_examples(self):
        self._check(self.lib.smc_mlp_clear_examples(self.h), "smc_mlp_clear_examples")

    def example(self, i):
        hdr = np.empty(self.HDR, dtype=np.float32)
        K, lab = ctypes.c_int32(), ctypes.c_int32()
        self._check(self.lib.smc_mlp_get_example(self.h, int(i), _fp(hdr), ctypes.byref(K), ctypes.byref(lab),
                                                 None), "smc_mlp_get_example")
        cls = np.empty((K.value, self.CLS), dtype=np.float32)
        self._check(self.lib.smc_mlp_get_example(self.h, int(i), None, None, None, _fp(cls)),
                    "smc_mlp_get_example")
        return hdr, K.value, lab.value, cls

    def loss_grad(self, first, n):
        g = np.empty(self.num_params, dtype=np.float32)
        loss = ctypes.c_double()
        self._check(self.lib.smc_mlp_loss_grad(self.h, int(first), int(n), ctypes.byref(loss), _fp(g)),
                    "smc_mlp_loss_grad")
        return loss.value, g

    def infer(self, first, n, out_stride):
        out = np.empty((n, out_stride), dtype=np.float32)
        self._check(self.lib.smc_mlp_infer(self.h, int(first), int(n), int(out_stride), _fp(out)),
                    "smc_mlp_infer")
        return out

    def train(self, epochs, batch=4096, lr=1e-3, seed=0):
        losses = np.empty(max(epochs, 1))
        self._check(self.lib.smc_mlp_train(self.h, int(epochs), int(batch), float(lr), int(seed), _dp(losses)),
                    "smc_mlp_train")
        return losses[:epochs]

    def move(self, slots, tempering, seed, num_iter=0, dpp_alpha=0.01, proposal_tempering=-1.0):
        """The learned allocation move; returns stats [updates, proposed, accepted,
        created, erased]."""
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        st = np.zeros(5, dtype=np.int64)
        self._check(self.lib.smc_learned_partition_move(self.h, len(sl), _ip(sl), float(tempering), int(seed),
                                                        int(num_iter), float(dpp_alpha),
                                                        float(proposal_tempering), _lp(st)),
                    "smc_learned_partition_move")
        return st


def discretize_gamma(shape, num):
    lib = _lib.load()
    out = np.empty(num)
    rc = lib.smc_discretize_gamma(float(shape), int(num), _dp(out))
    if rc != 0:
        raise EngineError("smc_discretize_gamma failed (%d)" % rc)
    return out


def lg_base_matrix():
    lib = _lib.load()
    out = np.empty(190)
    lib.smc_lg_base_matrix(_dp(out))
    return out
