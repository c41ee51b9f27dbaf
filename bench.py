"""Benchmark of the particle-likelihood hot path on one MI355X.

Headline (the default run): C4 -- 200 taxa x 10,000 sites, GTR+G4+I, 4,096
particles with CAT class counts K ~ U{5..120}, FP32 pruning on the matrix cores with
FP64 eigensystems, back-transforms and reductions.  One step = one pass of the hot
path over the whole population held in HBM: eigen-decomposition of every class rate
matrix, P(t) for every (node, rate, class), level-ordered pruning with the root
mixture fused, the per-particle lnL reduction (Particle::refreshLikelihood(true) for
every particle, Particle.cpp:530-675), then computeNextStep (Miscellaneous.hpp:42-109)
and systematic resampling (main.cpp:179-195).  Partials are resident for about 240
particles at a time, so a sweep runs in waves; each sweep is one replay of its
captured graph.  `value` is node.pattern.category updates per second.

Extra keys of the same JSON line:
  c2_fp64          the plain FP64 drop-in on C2 (64 taxa x 2,000 sites, 1,024 particles)
  rejuvenation_c4  SMC ESS/sec at the C4 shape through the cost of one rejuvenation
                   step per particle: plain FP64 (one particle at a time, Gibbs CAT move)
                   against FP32 + lockstep batches with the Gibbs move and with the learned
                   move (network trained on >= 1M recorded Gibbs draws before timing)
  smc_ess          whole SMC runs (ESS at T = 1 / wall seconds) on a small CAT workload
  cpu_baseline     the CPU FP64 restatement of the reference loop on the host's cores
Synthetic data (simulated alignments, prior-drawn particle states).  Prints one JSON line.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "smc-cat_amd"))

import numpy as np  # noqa: E402

METRIC = "pruned node·pattern·category updates/s; SMC ESS/sec vs plain FP64, 1 MI355X"
HBM_PEAK_GBS = 8000.0
MFMA_PEAK = {"fp64": 78.6, "fp32": 157.3}
# The launch shape the autotuner picked for a full sweep of each workload on MI355X
# (smc_engine_autotune over one wave); --autotune searches again.
TUNED = {("c4", "fp32"): (64, [2, 2, 2], 8, 2), ("c2", "fp64"): (32, [1, 1, 2], 8, 2),
         ("c3", "fp32"): (64, [2, 2, 2], 8, 2)}


def progress(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def algorithmic_bytes_per_column(T, word):
    """Persisted-partials traffic per (pattern, rate) column of one full sweep:
    T-2 internal non-root children read 20 words, T tip children read a 1-byte
    code, T-1 internal nodes write 20 words + an 8-byte log-scale."""
    return (T - 2) * 20 * word + T * 1 + (T - 1) * (20 * word + 8)


def profiled_value(name, config, precision, particles, default_particles):
    """A figure measured with rocprofv3 for this exact workload, read from the
    machine-readable lines of the committed profile summary, or None:
    traffic_bytes_per_launch (HBM bytes per pruning launch, FETCH_SIZE/WRITE_SIZE
    calibrated against a plain copy), mfma_busy (MfmaUtil over the pruning
    dispatches), mfma_tflops (MfmaFlops / time)."""
    if particles != default_particles:
        return None
    files = {"c2": "prune_mfma_c2_64taxa_2000sites_1024particles.md",
             "c4": "prune_mfma_c4_200taxa_10000sites_4096particles.md"}
    if config not in files:
        return None
    try:
        with open(os.path.join(REPO, "profiles", files[config])) as f:
            for line in f:
                m = re.match(r"(\w+)\[(\w+)\] = ([0-9.e+]+)", line.strip())
                if m and m.group(1) == name and m.group(2) == precision:
                    return float(m.group(3))
    except OSError:
        return None
    return None


def host_info():
    """Host cores and CPU model (the CPU baseline's machine)."""
    info = {"nproc": None, "cpu_count": os.cpu_count(), "model": None}
    try:
        info["nproc"] = int(subprocess.run(["nproc"], stdout=subprocess.PIPE, text=True, timeout=10).stdout)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemAvailable:"):
                    info["mem_available_gb"] = int(line.split()[1]) / 1e6
    except OSError:
        pass
    return info


def cpu_baseline(codes, batch, R, budget_s, bytes_per_particle):
    """The CPU FP64 oracle (the reference's loop order: node -> child -> class -> rate
    -> member site, 20x20 matvec, per-(site, rate) max-rescale; one particle per
    OpenMP thread, as main.cpp:295) on every host core the memory allows."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_binding as orc
    T, S = codes.shape
    info = host_info()
    cores = info["nproc"] or os.cpu_count() or 1
    # host memory: the box caps a command near 270 GB; keep the oracle's buffers under 150 GB
    mem_cap = max(1, int(150e9 // bytes_per_particle))
    threads = max(1, min(cores, mem_cap, batch.n))
    ps = [batch.particle(i) for i in range(batch.n)]
    done, t0 = 0, time.perf_counter()
    used = threads
    while True:
        chunk = [ps[(done + k) % len(ps)] for k in range(threads)]
        _, used = orc.lnl_batch(codes, chunk, R, threads=threads)
        done += len(chunk)
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": done * (T - 1) * S * R / dt, "unit": "updates/s", "cores": int(used), "kind": "port",
            "sample": "%d particles (%d taxa x %d sites, K_r=%d, their own K) through the CPU FP64 oracle in "
                      "reference loop order, one particle per OpenMP thread, %.1f s; threads = min(nproc %s, "
                      "%d by host memory at %.2f GB per particle)" %
                      (done, T, S, R, dt, cores, mem_cap, bytes_per_particle / 1e9),
            "host": info}


def sweep_bench(args, config, precision, steps, warmup, autotune, fp32_error=0, tag=""):
    """Full-likelihood sweeps of a configuration's whole population: updates/s,
    the roofline of the pruning kernel and the per-family kernel times."""
    from smccat import Engine, lg_base_matrix
    from smccat import simulate as sim

    codes, cfg = sim.simulate_config(config)
    T, S, R = cfg["taxa"], cfg["sites"], cfg["rates"]
    N = args.particles or cfg["particles"]
    kr = cfg["k_range"]
    batch = sim.prior_particles(np.random.default_rng(1), N, T, S, num_rates=R, invar=cfg["invar"], lg=cfg["lg"],
                                lg_table=lg_base_matrix() if cfg["lg"] else None, k_range=kr)
    word = 8 if precision == "fp64" else 4
    slots = args.partial_slots
    if slots == 0:  # partials of one wave within ~160 GB of HBM
        per_particle = (T - 1) * R * S * (20 * word + 4)
        slots = int(min(N, max(1, 160e9 // per_particle)))
    eng = Engine(T, S, R, N, kr[1], precision, partial_slots=slots,
                 pmat_classes=int(slots * (kr[0] + kr[1]) // 2 * 1.25) if kr[1] > 1 else 0)
    eng.set_prune_variant(args.variant)
    mixture = args.mixture or ("logspace" if config in ("c3", "c4") else "reference")
    eng.set_mixture(mixture)
    eng.set_alignment(codes)
    eng.set_particles(batch)
    target = int(0.6 * N)
    progress("%ssweep workload %s %s: %d particles uploaded, waves of %d" % (tag, config, precision, N, slots))

    def step():
        lnl, nf = eng.refresh_likelihood()
        w = np.full(N, -np.log(N))
        w, norm, ess, temp = eng.compute_next_step(w, lnl, target, 0.0)
        eng.systematic_resample(norm, 0.5)
        return lnl, nf, ess, temp

    tune = None
    if args.variant == "mfma":
        if args.tuning:
            g, u0, u1, u2, rt = (int(x) for x in args.tuning.split(","))
            eng.set_tuning(g, [u0, u1, u2], 8, unroll_runtime=rt)
        elif autotune:
            tune = eng.autotune(0, slots, sweeps=2 if slots == N else 1)
            progress("%sautotune: %s" % (tag, tune))
        elif (config, precision) in TUNED:
            g, u, x, rt = TUNED[(config, precision)]
            eng.set_tuning(g, u, x, unroll_runtime=rt)
    for _ in range(warmup):
        step()
    eng.set_graphs(not args.no_graphs)
    t0 = time.perf_counter()
    for _ in range(steps):
        lnl, nf, ess, temp = step()
    dt = time.perf_counter() - t0
    progress("%ssweep: %d steps in %.2f s" % (tag, steps, dt))
    # per-kernel durations (HIP events on the engine's stream, the stream the kernels
    # run on): the same kernels and launch shapes, launched eagerly so each launch is timed
    n_timed = 1 if N > 2048 else max(1, min(steps, 3))
    eng.set_timing(True)
    eng.reset_stats()
    for _ in range(n_timed):
        step()
    stats = eng.kernel_stats()
    eng.set_timing(False)
    shape = tune or eng.tuning()
    graphs = eng.graph_stats()
    patterns = eng.pattern_count(0, N)  # site patterns within classes: the columns a sweep prunes
    eng.close()

    updates = patterns * (T - 1) * R
    bytes_sweep = patterns * R * algorithmic_bytes_per_column(T, word)
    n_prune, ms_prune = stats["prune"]
    prune_ms = ms_prune / n_timed
    achieved = bytes_sweep / (prune_ms * 1e-3) / 1e9
    traffic = profiled_value("traffic_bytes_per_launch", config, precision, N, cfg["particles"])
    out = {
        "value": updates * steps / dt,
        "ms_per_step": dt * 1e3 / steps,
        "workload": "%s: %d taxa x %d sites, %s+G%d%s%s, %d particles; step = full refreshLikelihood(true) of "
                    "every particle + computeNextStep + systematic resampling" %
                    (config, T, S, "LG" if cfg["lg"] else "GTR", R, "+I" if cfg["invar"] else "",
                     " CAT K~U{%d..%d}" % kr if kr[1] > 1 else "", N),
        "config": {"taxa": T, "sites": S, "rates": R, "particles": N, "partial_slots": slots, "precision": precision,
                   "prune_kernel": args.variant, "mixture": mixture, "updates_per_step": updates,
                   "patterns_per_step": patterns, "columns_per_step": N * S},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "mfma": {"busy": profiled_value("mfma_busy", config, precision, N, cfg["particles"]),
                              "achieved": profiled_value("mfma_tflops", config, precision, N, cfg["particles"]),
                              "peak": MFMA_PEAK[precision], "unit": "TFLOP/s",
                              "source": "rocprofv3 --pmc MfmaUtil MfmaFlops (committed profile)"},
                     "kernel": "%s<%s>" % ("k_prune_mfma" if args.variant == "mfma" else "k_prune",
                                           "double" if word == 8 else "float"),
                     "launches_per_step": n_prune / n_timed,
                     "avg_launch_ms": ms_prune / max(n_prune, 1),
                     "algorithmic_bytes_per_step": bytes_sweep,
                     "bytes_per_update": bytes_sweep / updates},
        "kernel_ms_per_step": {k: v[1] / n_timed for k, v in stats.items()},
        "launch_shape": shape,
        "graphs": graphs,
        "check": {"lnl_finite": bool(np.all(np.isfinite(lnl)) and not nf.any()),
                  "nonfinite_particles": int(np.sum(~np.isfinite(lnl))),
                  "lnl_mean_finite": float(np.mean(lnl[np.isfinite(lnl)])) if np.isfinite(lnl).any() else None,
                  "ess": ess, "temp": temp},
    }
    if precision == "fp32" and fp32_error:
        # per-particle error of the fast path against the plain FP64 engine (which
        # matches the FP64 oracle within 1e-10) on a sample of the same particles
        idx = np.arange(0, N, max(1, N // fp32_error))[:fp32_error]
        e64 = Engine(T, S, R, len(idx), kr[1], "fp64", partial_slots=min(len(idx), max(1, slots // 2)))
        e64.set_mixture(mixture)
        e64.set_alignment(codes)
        e64.set_particles(batch.take(idx))
        ref, _ = e64.refresh_likelihood()
        e64.close()
        ok = np.isfinite(ref) & np.isfinite(lnl[idx])
        rel = np.abs(lnl[idx][ok] - ref[ok]) / np.abs(ref[ok])
        out["fp32_error_vs_fp64"] = {"particles": int(len(idx)), "max_rel": float(rel.max()),
                                     "median_rel": float(np.median(rel)),
                                     "max_abs": float(np.abs(lnl[idx][ok] - ref[ok]).max()),
                                     "nonfinite_in_both": int(np.sum(~np.isfinite(ref) & ~np.isfinite(lnl[idx])))}
    return out, (codes, batch, R, T, S)


def ess_run(codes, batch, R, precision, batch_size, learned, iterations, seed, dpp_alpha, kmax,
            alg8=0.1, warmup=1, epochs=8):
    """One SMC run (main.cpp:143-336) on the engine; returns its summary."""
    from smccat import Engine, ProposalNetwork
    T, S = codes.shape
    N = batch.n
    eng = Engine(T, S, R, N, kmax, precision, workspaces=2)
    net = None
    try:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        if learned:
            net = ProposalNetwork(eng, seed=seed)
        res, lnl, _ = eng.run_sampler(N, iterations, seed=seed, using_lg=True, invariant=False, cat=True,
                                      alg8_probability=alg8, batch=batch_size, dpp_alpha=dpp_alpha,
                                      proposal=net, proposal_warmup=warmup, train_epochs=epochs)
    finally:
        if net is not None:
            net.close()
        eng.close()
    return {"seconds": res["seconds"], "final_ess": res["final_ess"],
            "ess_per_sec": res["final_ess"] / res["seconds"],
            "mean_ess": res["ess_sum"] / max(res["iterations"], 1),
            "mean_ess_per_sec": res["ess_sum"] / max(res["iterations"], 1) / res["seconds"],
            "iterations": res["iterations"], "resamples": res["resamples"],
            "rejuvenation_seconds": res["rejuvenation_seconds"], "training_seconds": res["training_seconds"],
            "training_examples": int(res["cat_stats"][2])}


def ess_compare(args):
    """Whole SMC runs: the plain FP64 path (one particle at a time, as the
    reference's per-thread loop; Gibbs CAT move) against FP32 + batched + learned
    proposal on the same CAT data and initial population."""
    from smccat import lg_base_matrix
    from smccat import simulate as sim
    T, S, N = args.ess_taxa, args.ess_sites, args.ess_particles
    codes, cfg = sim.simulate_config("c3", taxa=T, sites=S)
    R = cfg["rates"]
    kr = (5, 60)
    batch = sim.prior_particles(np.random.default_rng(11), N, T, S, num_rates=R, lg=True,
                                lg_table=lg_base_matrix(), k_range=kr, max_classes=kr[1] + 8)
    out = {"workload": "CAT data as c3 (DP-CAT, 20 profiles ~ Dirichlet(0.5), LG+G4) at %d taxa x %d sites; %d "
                       "particles, K ~ U{%d..%d}; %d rejuvenation iterations; CAT move probability 0.1; DP alpha "
                       "1.0" % (T, S, N, kr[0], kr[1], args.ess_iterations),
           "metric": "ESS at T = 1 / wall seconds of the SMC loop (initial likelihoods included); mean ESS after "
                     "each reweight / wall seconds beside it"}
    for name, (prec, bs, learned) in {"plain_fp64": ("fp64", 1, False),
                                      "fp32_batched_learned": ("fp32", 0, True)}.items():
        vals = []
        for seed in range(1, args.ess_seeds + 1):
            vals.append(ess_run(codes, batch, R, prec, bs, learned, args.ess_iterations, seed, 1.0, kr[1] + 8))
            progress("ess %s seed %d: %.1f s, final ESS %.1f, %d iterations" %
                     (name, seed, vals[-1]["seconds"], vals[-1]["final_ess"], vals[-1]["iterations"]))
        out[name] = {k: float(np.mean([v[k] for v in vals])) for k in vals[0]}
        out[name]["seeds"] = args.ess_seeds
    out["ratio_ess_per_sec"] = out["fp32_batched_learned"]["ess_per_sec"] / out["plain_fp64"]["ess_per_sec"]
    out["ratio_mean_ess_per_sec"] = (out["fp32_batched_learned"]["mean_ess_per_sec"] /
                                     out["plain_fp64"]["mean_ess_per_sec"])
    return out


def rejuv_compare(args):
    """Cost of one rejuvenation step per particle at the C4 shape (200 taxa x
    10,000 sites, GTR+G4+I, CAT K ~ U{5..120}; r MCMC iterations, then with
    probability 0.1 the CAT move, main.cpp:296-311).  Every path runs the same moves
    on the same target, so at equal ESS per SMC step the ratio of these costs is the
    ratio of ESS/sec.  Paths:
      plain_fp64            one particle at a time (the reference's per-thread loop),
                            FP64, the Gibbs CAT move
      fp32_batched_gibbs    FP32, lockstep batches of 150 resident particles keyed by
                            chain length (the sampler's shape): the C4 population
                            of 4,096 chains sorted by length into batches (27 full
                            ones), of which 8 spread over the order are timed (1,200
                            particles); the
                            Gibbs CAT move in one lockstep of the ~410 particles that
                            draw it at the C4 population
      fp32_batched_learned  the same with the learned CAT move, its network trained
                            on >= 1M Gibbs draws recorded before the timed runs."""
    from smccat import Engine, ProposalNetwork
    from smccat import simulate as sim
    codes, cfg = sim.simulate_config("c4")
    T, S = codes.shape
    R, kr = cfg["rates"], cfg["k_range"]
    kmax = kr[1] + 8
    nb, nbatch, it, tau = args.rj_particles, args.rj_batches, args.rj_iterations, args.rj_temp
    popn = nb * nbatch
    # the population's chains keyed by length (the sampler's order) into lockstep
    # batches of nb; the timed batches are spread evenly over that order
    npop = max(args.rj_population, popn)
    pop_all = sim.prior_particles(np.random.default_rng(4), npop, T, S, num_rates=R, invar=True, k_range=kr,
                                  max_classes=kmax)
    out = {"workload": "c4 data (%d taxa x %d sites, GTR+G4+I, DP-CAT), K ~ U{%d..%d}, %d MCMC iterations, CAT "
                       "move probability 0.1, DP alpha 1.0, tempering %.2f" % (T, S, kr[0], kr[1], it, tau)}
    # ---- fast paths: FP32, batches of nb particles resident, keys sorted by chain length
    eng = Engine(T, S, R, popn, kmax, "fp32", workspaces=2, partial_slots=nb,
                 pmat_classes=int(nb * (kr[0] + kr[1]) / 2 * 1.4))
    eng.set_mixture("logspace")
    eng.set_alignment(codes)
    full = eng.rejuvenation_steps(it, 5, np.arange(npop, dtype=np.int32), invariant=True)
    order = np.argsort(-full, kind="stable").astype(np.int32)
    nbat_all = npop // nb  # full batches (a last, partial one is not timed)
    chosen = [int((j + 0.5) * nbat_all / nbatch) for j in range(nbatch)]
    keys = np.concatenate([order[b * nb:(b + 1) * nb] for b in chosen]).astype(np.int32)
    assert len(keys) == popn, "timed batches must be full"
    pop = pop_all.take(keys)
    eng.set_particles(pop)
    net = ProposalNetwork(eng, seed=1)
    # training pairs: Gibbs draws recorded on the timed particles (the CAT move scores
    # single sites from the tips and needs no resident partials)
    t0 = time.perf_counter()
    k = 0
    while net.num_examples < args.rj_train_pairs:
        net.collect_gibbs(range(popn), tau, seed=300 + k, num_iter=args.rj_collect_iters, dpp_alpha=1.0)
        k += 1
    t_col = time.perf_counter() - t0
    t0 = time.perf_counter()
    losses = net.train(args.rj_epochs, batch=4096, lr=2e-3, seed=1)
    t_train = time.perf_counter() - t0
    n_pairs = net.num_examples
    progress("rejuv: %d training pairs recorded in %.1f s, %d epochs in %.1f s (loss %.3f -> %.3f)" %
             (n_pairs, t_col, args.rj_epochs, t_train, losses[0], losses[-1]))
    eng.set_particles(pop)  # the timed runs start from the initial population
    steps = eng.rejuvenation_steps(it, 5, keys, invariant=True)
    t0 = time.perf_counter()
    stats_tot = np.zeros((4, 2), dtype=np.int64)
    for b in range(nbatch):
        _, st = eng.rejuvenate(tau, it, seed=5, first=b * nb, count=nb, invariant=True, gibbs_probability=0.0,
                               dpp_alpha=1.0, stream_keys=keys[b * nb:(b + 1) * nb])
        stats_tot += st
    t_mcmc = time.perf_counter() - t0
    s_step = t_mcmc / float(steps.sum())
    # lockstep efficiency: MH steps over (batch size x rounds), rounds being the
    # longest chain of each batch, for the timed batches and for the whole population
    rounds = sum(int(steps[b * nb:(b + 1) * nb].max()) for b in range(nbatch))
    fsort = full[order]
    slots_all = sum(len(fsort[b:b + nb]) * int(fsort[b:b + nb].max()) for b in range(0, len(fsort), nb))
    lockstep = {"rounds": rounds, "efficiency": float(steps.sum()) / (nb * rounds),
                "population": npop, "batches_in_population": nbat_all, "timed_batches": chosen,
                "efficiency_population": float(full.sum()) / slots_all}
    progress("rejuv fast MCMC: %d particles in batches of %d: %.1f ms per particle, %.1f us per MH step" %
             (popn, nb, t_mcmc / popn * 1e3, s_step * 1e6))
    # CAT moves: the sampler moves every particle that draws one (probability 0.1)
    # together, after the MCMC of all batches: at the C4 population that is
    # about 410 particles per rejuvenation; each particle is charged 0.1 moves
    n_cat = min(popn, int(round(0.1 * 4096)))
    slots = list(range(n_cat))
    cat = {}
    gi = max(1, S // (2 * args.rj_gibbs_fraction))
    t0 = time.perf_counter()
    eng.gibbs_partition_move(slots, tau, seed=6, num_iter=gi, dpp_alpha=1.0)
    cat["gibbs"] = (time.perf_counter() - t0) * (S / 2.0) / gi / n_cat
    t0 = time.perf_counter()
    mst = net.move(slots, tau, seed=6, dpp_alpha=1.0)
    cat["learned"] = (time.perf_counter() - t0) / n_cat
    acc = float(mst[2] / max(mst[1], 1))
    t0 = time.perf_counter()
    eng.refresh_likelihood(0, nb)
    t_ref = (time.perf_counter() - t0) / nb
    net.close()
    eng.close()
    mean_steps = float(full.mean())  # the population's mean chain
    fast = {}
    for name in ("gibbs", "learned"):
        per = s_step * mean_steps + 0.1 * cat[name] + t_ref
        fast[name] = {"mcmc_seconds_per_step": s_step, "mcmc_seconds_per_particle": s_step * mean_steps,
                      "cat_move_seconds_per_move": cat[name], "refresh_seconds_per_particle": t_ref,
                      "seconds_per_particle": per}
    fast["learned"].update({"training_pairs": n_pairs, "collect_seconds": t_col, "train_seconds": t_train,
                            "train_loss_first_last": [float(losses[0]), float(losses[-1])],
                            "acceptance": acc, "proposals": int(mst[1]), "accepted": int(mst[2])})
    fast["gibbs"]["gibbs_iterations_timed"] = gi
    progress("rejuv fast CAT move per particle: Gibbs %.3f s, learned %.3f s (acceptance %.2f)" %
             (cat["gibbs"], cat["learned"], acc))
    out["mean_steps"] = mean_steps
    out["timed_particles"] = popn
    out["lockstep"] = lockstep
    out["mh_stats_proposed_accepted"] = stats_tot.tolist()
    out["fp32_batched_gibbs"] = fast["gibbs"]
    out["fp32_batched_learned"] = fast["learned"]
    # ---- plain path: FP64, one particle at a time, on particles spread over the
    # population (a particle's cost per step depends on its move mix and classes)
    npl = max(1, min(args.rj_plain, npop))
    pick = np.linspace(0, npop - 1, npl).round().astype(np.int32)
    sub = pop_all.take(pick)
    eng = Engine(T, S, R, npl, kmax, "fp64", workspaces=2)
    eng.set_mixture("logspace")
    eng.set_alignment(codes)
    eng.set_particles(sub)
    t0 = time.perf_counter()
    for i in range(npl):
        eng.rejuvenate(tau, it, seed=5, first=i, count=1, invariant=True, gibbs_probability=0.0,
                       stream_keys=pick[i:i + 1])
    t_pl = time.perf_counter() - t0
    steps_pl = eng.rejuvenation_steps(it, 5, pick, invariant=True)
    s_plain = t_pl / float(steps_pl.sum())
    t0 = time.perf_counter()
    eng.gibbs_partition_move([0], tau, seed=7, num_iter=gi, dpp_alpha=1.0)
    eng.refresh_likelihood(0, 1)
    t_gibbs = (time.perf_counter() - t0) * (S / 2.0) / gi
    eng.close()
    per_plain = s_plain * mean_steps + 0.1 * t_gibbs
    out["plain_fp64"] = {"particles_sampled": npl, "sample": pick.tolist(), "mcmc_seconds_per_step": s_plain,
                         "mcmc_seconds_per_particle": s_plain * mean_steps, "gibbs_seconds_per_move": t_gibbs,
                         "seconds_per_particle": per_plain}
    progress("rejuv plain: %.1f us per MH step, Gibbs move %.2f s" % (s_plain * 1e6, t_gibbs))
    # MCMC parts priced at the population's mean chain length (sampled chains differ by chance)
    out["ratio_fp32_batched_learned"] = per_plain / fast["learned"]["seconds_per_particle"]
    out["ratio_fp32_batched_gibbs"] = per_plain / fast["gibbs"]["seconds_per_particle"]
    out["note"] = ("ratios are plain seconds per particle over fast seconds per particle at the population's mean "
                   "MH step count; equal ESS per SMC step makes them the ESS/sec ratios")
    progress("rejuv ratio: learned %.2fx, gibbs %.2fx" % (out["ratio_fp32_batched_learned"],
                                                           out["ratio_fp32_batched_gibbs"]))
    return out


def c5_bench(args):
    """Learned CAT-allocation proposal on c3 data: Gibbs draws recorded as
    training pairs, GPU training, and the learned move against the Gibbs move."""
    from smccat import Engine, ProposalNetwork, lg_base_matrix
    from smccat import simulate as sim
    codes, cfg = sim.simulate_config("c3")
    T, S = codes.shape
    R, N, kr = cfg["rates"], args.c5_particles or cfg["particles"], cfg["k_range"]
    batch = sim.prior_particles(np.random.default_rng(5), N, T, S, num_rates=R, lg=True,
                                lg_table=lg_base_matrix(), k_range=kr, max_classes=kr[1] + 8)
    out = {"workload": "c3 data (%d taxa x %d sites, LG+G4, DP-CAT), %d particles K ~ U{%d..%d}, DP alpha 1.0, "
                       "tempering 1" % (T, S, N, kr[0], kr[1])}
    eng = Engine(T, S, R, N, kr[1] + 8, "fp32", partial_slots=8)
    eng.set_alignment(codes)
    eng.set_particles(batch)
    net, crp, ev = ProposalNetwork(eng, seed=1), ProposalNetwork(eng, seed=2), ProposalNetwork(eng, seed=3)
    t0 = time.perf_counter()
    rounds = 0
    while net.num_examples < args.c5_examples:
        net.collect_gibbs(range(N), 1.0, seed=100 + rounds, dpp_alpha=1.0)
        rounds += 1
    t_collect = time.perf_counter() - t0
    n = net.num_examples
    progress("c5 collected %d examples in %.1f s" % (n, t_collect))
    ev.collect_gibbs(range(N), 1.0, seed=900, num_iter=100, dpp_alpha=1.0)  # held out, never trained on
    hold = ev.num_examples
    ev.params = np.zeros(net.num_params, dtype=np.float32)
    crp_loss, _ = ev.loss_grad(0, hold)
    ev.params = net.params
    before, _ = ev.loss_grad(0, hold)
    t0 = time.perf_counter()
    losses = net.train(args.c5_epochs, batch=4096, lr=2e-3, seed=1)
    t_train = time.perf_counter() - t0
    ev.params = net.params
    after, _ = ev.loss_grad(0, hold)
    out.update({"examples": n, "heldout_examples": hold, "collect_seconds": t_collect,
                "train": {"epochs": args.c5_epochs, "seconds": t_train, "examples_per_sec": n * args.c5_epochs / t_train,
                          "epoch_loss": [float(x) for x in losses]},
                "heldout_loss": {"crp_only": crp_loss, "untrained": before, "trained": after}})
    crp.params = np.zeros(net.num_params, dtype=np.float32)
    iters = args.c5_move_iters
    moves = {}
    for name, fn in (("gibbs", lambda: eng.gibbs_partition_move(range(N), 1.0, seed=7, num_iter=iters,
                                                                   dpp_alpha=1.0)),
                     ("learned", lambda: net.move(range(N), 1.0, seed=7, num_iter=iters, dpp_alpha=1.0)),
                     ("crp_proposal", lambda: crp.move(range(N), 1.0, seed=7, num_iter=iters, dpp_alpha=1.0))):
        t0 = time.perf_counter()
        st = fn()
        dt = time.perf_counter() - t0
        moves[name] = {"seconds": dt, "site_updates_per_sec": N * iters / dt,
                       "site_moves": int(st[0] if name == "gibbs" else st[2])}
        if name != "gibbs":
            moves[name]["acceptance"] = float(st[2] / max(st[1], 1))
    out["moves"] = moves
    out["speedup_site_updates"] = moves["learned"]["site_updates_per_sec"] / moves["gibbs"]["site_updates_per_sec"]
    for x in (net, crp, ev):
        x.close()
    eng.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", default="headline", choices=["headline", "sweep", "rejuv", "c5"],
                    help="headline: the C4 updates/s line with every extra key; sweep: one "
                         "configuration's sweep only; rejuv: the rejuvenation comparison; c5: the "
                         "learned-proposal benchmark")
    ap.add_argument("--config", default="c4", choices=["c1", "c2", "c3", "c4"])
    ap.add_argument("--precision", default="", choices=["", "fp64", "fp32"],
                    help="default: fp32 for c3/c4, fp64 for c1/c2")
    ap.add_argument("--partial-slots", type=int, default=0)
    ap.add_argument("--particles", type=int, default=0)
    ap.add_argument("--variant", default="mfma", choices=["mfma", "valu"])
    ap.add_argument("--mixture", default="", choices=["", "reference", "logspace"])
    ap.add_argument("--autotune", action="store_true", help="search the launch shape instead of the tuned one")
    ap.add_argument("--tuning", default="",
                    help="fixed launch shape G,U0,U1,U2,RT: workgroups per CU, column-block unrolls by stored "
                         "children, roots' unroll (default: the tuned shape of the workload)")
    ap.add_argument("--no-graphs", action="store_true", help="time eager launches instead of graph replays")
    ap.add_argument("--fp32-error", type=int, default=32)
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--no-rejuv", action="store_true")
    ap.add_argument("--no-ess", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--rj-particles", type=int, default=150, help="particles per resident lockstep batch")
    ap.add_argument("--rj-batches", type=int, default=8,
                    help="resident lockstep batches timed, spread over the population's batches")
    ap.add_argument("--rj-population", type=int, default=4096,
                    help="population whose chains are keyed by length into lockstep batches (0: the timed "
                         "particles alone)")
    ap.add_argument("--rj-plain", type=int, default=12, help="particles of the plain FP64 path, spread over the population")
    ap.add_argument("--rj-iterations", type=int, default=16)
    ap.add_argument("--rj-temp", type=float, default=0.5)
    ap.add_argument("--rj-train-pairs", type=int, default=1_000_000)
    ap.add_argument("--rj-collect-iters", type=int, default=600)
    ap.add_argument("--rj-epochs", type=int, default=4)
    ap.add_argument("--rj-gibbs-fraction", type=int, default=10,
                    help="Gibbs moves are timed over 1/this of their S/2 site updates and scaled")
    ap.add_argument("--ess-taxa", type=int, default=16)
    ap.add_argument("--ess-sites", type=int, default=200)
    ap.add_argument("--ess-particles", type=int, default=64)
    ap.add_argument("--ess-iterations", type=int, default=4)
    ap.add_argument("--ess-seeds", type=int, default=2)
    ap.add_argument("--c5-particles", type=int, default=0)
    ap.add_argument("--c5-examples", type=int, default=1_000_000)
    ap.add_argument("--c5-epochs", type=int, default=6)
    ap.add_argument("--c5-move-iters", type=int, default=500)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(particles would shard across GPUs by slot)", file=sys.stderr)
        sys.exit(2)
    if args.mode == "rejuv":
        print(json.dumps({"metric": "rejuvenation step cost per particle at the C4 shape",
                          "rejuvenation_c4": rejuv_compare(args)}), flush=True)
        return
    if args.mode == "c5":
        print(json.dumps({"metric": "learned CAT-allocation proposal: training and move throughput",
                          "c5": c5_bench(args)}), flush=True)
        return
    precision = args.precision or ("fp32" if args.config in ("c3", "c4") else "fp64")
    head, (codes, batch, R, T, S) = sweep_bench(args, args.config, precision, args.steps, args.warmup,
                                                args.autotune, fp32_error=args.fp32_error)
    out = {"metric": METRIC, "value": head["value"], "unit": "updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64" if precision == "fp64" else "f32",
           "data": "synthetic", "config": dict(head["config"], workload=head["workload"])}
    for k in ("roofline", "kernel_ms_per_step", "launch_shape", "graphs", "check", "fp32_error_vs_fp64"):
        if k in head:
            out[k] = head[k]
    if args.mode == "headline" and not args.no_c2 and args.config != "c2":
        c2, _ = sweep_bench(args, "c2", "fp64", args.steps, args.warmup, args.autotune, tag="c2: ")
        out["c2_fp64"] = {k: c2[k] for k in ("value", "ms_per_step", "workload", "roofline", "kernel_ms_per_step",
                                             "launch_shape", "check")}
    if not args.no_cpu_baseline:
        # the same configuration through the CPU restatement, on a bounded sample
        from smccat import simulate as sim
        cfg = sim.CONFIGS[args.config]
        kr = cfg["k_range"]
        sample = batch.take(np.arange(min(batch.n, 512)))
        per_particle = 8.0 * (20 * S * R * (2 * T - 1) + 400 * R * (2 * T - 1) * max(1, (kr[0] + kr[1]) // 2))
        out["cpu_baseline"] = cpu_baseline(codes, sample, R, args.cpu_seconds, per_particle)
        progress("cpu baseline: %.3g updates/s on %d threads" % (out["cpu_baseline"]["value"],
                                                                 out["cpu_baseline"]["cores"]))
    else:
        out["cpu_baseline"] = None
    if args.mode == "headline" and not args.no_rejuv:
        out["rejuvenation_c4"] = rejuv_compare(args)
    if args.mode == "headline" and not args.no_ess:
        out["smc_ess"] = ess_compare(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
