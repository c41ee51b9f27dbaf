"""Synthetic workloads: random trees, simulated amino-acid alignments and
prior-drawn particle states for the benchmark configurations.

* random_tree follows the reference's random tree (Tree.cpp:10-77): start from
  a root with two tips, then repeatedly split the pendant edge of a uniformly
  chosen tip (a Yule-process topology); every node draws an Exp(rate 10)
  length (Tree.cpp:343-353) and the root's is 0.  Node ids: tips 0..T-1,
  internal nodes T.. in creation order (root first).
* simulate_alignment evolves sites down the tree under a reversible model
  with per-site class (CAT profile), gamma rate category and invariant class.
* prior_particles draws particle states as Particle::initialize does
  (Particle.cpp:81-158): GTR log-exchangeabilities ~ N(0,1) or the LG table,
  19 N(0,1) logits + 0 per class, pInvar ~ U(0,1), shape ~ Gamma(3, scale 2).
"""
import numpy as np

from .engine import ParticleBatch

NS = 20

CONFIGS = {
    # benchmark workloads C1-C5: taxa, sites, rates, invariant, lg, particles, true profiles
    "c1": dict(taxa=16, sites=200, rates=4, invar=False, lg=True, particles=64, shape=0.8,
               profiles=1, k_range=(1, 1)),
    "c2": dict(taxa=64, sites=2000, rates=4, invar=True, lg=False, particles=1024, shape=0.8,
               profiles=1, k_range=(1, 1), pinv_true=0.1),
    "c3": dict(taxa=100, sites=5000, rates=4, invar=False, lg=True, particles=512, shape=0.8,
               profiles=20, k_range=(5, 60), dirichlet=0.5),
    "c4": dict(taxa=200, sites=10000, rates=4, invar=True, lg=False, particles=4096, shape=0.8,
               profiles=30, k_range=(5, 120), pinv_true=0.1),
}


def coords190():
    out = []
    for i in range(1, 20):
        for j in range(i):
            out.append((i, j))
    return out


_COORDS = coords190()


def random_tree(rng, taxa, mean_len=0.1):
    """Returns (children [T-1,2] int32, branch_len [2T-1], root id).

    The root is internal node T; the internal node created when tip i joins is
    T + i - 1 and takes children (i, r), r the split tip, in r's old child slot."""
    T = taxa
    children = np.zeros((T - 1, 2), dtype=np.int32)
    parent = np.zeros(2 * T - 1, dtype=np.int64)
    slot = np.zeros(2 * T - 1, dtype=np.int64)
    children[0] = (0, 1)
    parent[0] = parent[1] = T
    slot[1] = 1
    u = rng.random(max(T - 2, 0))
    for i in range(2, T):
        r = int(u[i - 2] * i)
        new = T + i - 1
        a, sl = parent[r], slot[r]
        children[a - T, sl] = new
        parent[new], slot[new] = a, sl
        children[new - T] = (i, r)
        parent[i], slot[i] = new, 0
        parent[r], slot[r] = new, 1
    blen = rng.exponential(mean_len, size=2 * T - 1)
    blen[T] = 0.0
    return children, blen, T


def build_q(base, logits):
    pi = np.exp(logits)
    pi = pi / pi.sum()
    R = np.zeros((NS, NS))
    for c, (i, j) in enumerate(_COORDS):
        R[i, j] = R[j, i] = np.exp(base[c])
    Q = R * pi[None, :]
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(axis=1))
    mu = float(np.sum(-np.diag(Q) * pi))
    return Q / mu, pi


def eigen(Q, pi):
    sp = np.sqrt(pi)
    S = sp[:, None] * Q / sp[None, :]
    S = 0.5 * (S + S.T)
    lam, U = np.linalg.eigh(S)
    return U / sp[:, None], lam, U.T * sp[None, :]


def pmatrix_from_eigen(A, lam, B, t):
    P = (A * np.exp(lam * t)[None, :]) @ B
    P = np.clip(P, 0.0, None)
    return P / P.sum(axis=1, keepdims=True)


def pmatrix(Q, pi, t):
    return pmatrix_from_eigen(*eigen(Q, pi), t)


def gamma_rates(shape, k):
    from scipy import stats
    if k == 1:
        return np.ones(1)
    q = (2 * np.arange(k) + 1) / (2.0 * k)
    v = stats.gamma.ppf(q, shape, scale=1.0 / shape)
    return v * (k / v.sum())


def simulate_alignment(rng, taxa, sites, children, blen, root, profiles_base, profiles_logits,
                       site_profile, rates, pinv=0.0):
    """Evolve `sites` columns; returns int8 codes [T, S] (0..19)."""
    T = taxa
    codes = np.zeros((T, sites), dtype=np.int8)
    rate_cat = rng.integers(0, len(rates), size=sites)
    invariant = rng.random(sites) < pinv
    states = np.zeros((2 * T - 1, sites), dtype=np.int64)
    # pre-order
    order = [root]
    k = 0
    while k < len(order):
        n = order[k]
        k += 1
        if n >= T:
            order.extend(int(c) for c in children[n - T])
    for prof in np.unique(site_profile):
        Q, pi = build_q(profiles_base[prof], profiles_logits[prof])
        A, lam, B = eigen(Q, pi)
        for rc in range(len(rates)):
            cols = np.nonzero((site_profile == prof) & (rate_cat == rc) & ~invariant)[0]
            if cols.size == 0:
                continue
            states[root, cols] = rng.choice(NS, size=cols.size, p=pi)
            for n in order:
                if n < T:
                    continue
                for c in children[n - T]:
                    P = pmatrix_from_eigen(A, lam, B, blen[c] * rates[rc])
                    cum = np.cumsum(P, axis=1)
                    u = rng.random(cols.size)
                    par = states[n, cols]
                    states[c, cols] = np.minimum((u[:, None] > cum[par]).sum(axis=1), NS - 1)
        inv_cols = np.nonzero((site_profile == prof) & invariant)[0]
        if inv_cols.size:
            st = rng.choice(NS, size=inv_cols.size, p=pi)
            states[:, inv_cols] = st[None, :]
    codes[:, :] = states[:T].astype(np.int8)
    return codes


def simulate_config(name, seed=None, sites=None, taxa=None):
    """Alignment for a benchmark configuration (data seed = config number)."""
    cfg = dict(CONFIGS[name])
    if sites:
        cfg["sites"] = sites
    if taxa:
        cfg["taxa"] = taxa
    seed = int(name[1:]) if seed is None else seed
    rng = np.random.default_rng(seed)
    T, S = cfg["taxa"], cfg["sites"]
    children, blen, root = random_tree(rng, T)
    nprof = cfg["profiles"]
    if cfg["lg"]:
        from .engine import lg_base_matrix
        try:
            lgv = lg_base_matrix()
        except RuntimeError:
            lgv = None
        # true LG rates for simulation: the published exchangeabilities themselves
        base_true = np.log(lgv) if lgv is not None else rng.normal(size=190)
    else:
        base_true = np.log(rng.dirichlet(np.ones(190)))
    pbase = np.tile(base_true, (nprof, 1))
    if nprof == 1:
        plog = np.log(rng.dirichlet(np.ones(NS)))[None, :]
        site_profile = np.zeros(S, dtype=np.int64)
    else:
        a = cfg.get("dirichlet", 1.0)
        plog = np.log(rng.dirichlet(np.full(NS, a), size=nprof) + 1e-300)
        # sites spread uniformly over the nprof true profiles
        site_profile = rng.integers(0, nprof, size=S)
    rates = gamma_rates(cfg["shape"], cfg["rates"])
    codes = simulate_alignment(rng, T, S, children, blen, root, pbase, plog, site_profile, rates,
                               cfg.get("pinv_true", 0.0))
    return codes, cfg


def prior_particles(rng, n, taxa, sites, num_rates=1, invar=False, lg=False, k_range=(1, 1),
                    max_classes=None, lg_table=None):
    """N particle states drawn as Particle::initialize would (Particle.cpp:81-158),
    except that CAT classes get K ~ U{k_range} with sites assigned uniformly
    (the reference's alpha = 0.01 CRP would almost always give K = 1)."""
    T = taxa
    Kmax = max_classes or k_range[1]
    children = np.zeros((n, T - 1, 2), dtype=np.int32)
    blen = np.zeros((n, 2 * T - 1))
    base = np.zeros((n, 190))
    ncls = np.zeros(n, dtype=np.int32)
    logits = np.zeros((n, Kmax, NS))
    asg = np.zeros((n, sites), dtype=np.int32)
    pinv = np.zeros(n)
    shape = np.ones(n)
    for i in range(n):
        ch, bl, _ = random_tree(rng, T)
        children[i] = ch
        blen[i] = bl
        if lg:
            base[i] = lg_table if lg_table is not None else rng.normal(size=190)
        else:
            base[i] = rng.normal(size=190)
        if invar:
            pinv[i] = rng.random()
        if num_rates > 1:
            shape[i] = rng.gamma(3.0, 2.0)
        K = int(rng.integers(k_range[0], k_range[1] + 1))
        ncls[i] = K
        logits[i, :K, :19] = rng.normal(size=(K, 19))
        if K == 1:
            asg[i] = 0
        else:
            a = rng.integers(0, K, size=sites)
            a[:K] = np.arange(K)          # every class non-empty
            rng.shuffle(a)
            asg[i] = a
    return ParticleBatch(children, blen, base, ncls, logits, asg, pinv, shape)
