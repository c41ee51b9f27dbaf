"""Independent likelihood checks for tiny trees (test helper).

The site likelihood is computed by explicit enumeration of every assignment
of states to the internal nodes (np.einsum with optimize=False sums the full
product tensor, i.e. no pruning), with P(t) from scipy.linalg.expm -- neither
the oracle's eigen-decomposition nor its pruning order is involved.  The
mixture formula is the reference's (Particle.cpp:653-670): rate weights 1
(Q4) and the invariant term p * pi[c] * isInvariant.
"""
import string

import numpy as np
import scipy.linalg

NS = 20


def coords190():
    return [(i, j) for i in range(1, 20) for j in range(i)]


def build_q(base, logits):
    pi = np.exp(logits)
    pi = pi / pi.sum()
    R = np.zeros((NS, NS))
    for c, (i, j) in enumerate(coords190()):
        R[i, j] = R[j, i] = np.exp(base[c])
    Q = R * pi[None, :]
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(axis=1))
    mu = float(np.sum(-np.diag(Q) * pi))
    return Q / mu, pi


def site_likelihood_rate(codes_col, children, blen, Q, pi, rate):
    """Sum over internal states of pi_root * prod_edges P(parent, child) for one site."""
    T = len(codes_col)
    internal = list(range(T, 2 * T - 1))
    parent = {}
    for k, n in enumerate(internal):
        for c in children[k]:
            parent[int(c)] = n
    root = [n for n in internal if n not in parent][0]
    letter = {n: string.ascii_letters[i] for i, n in enumerate(internal)}
    ops, subs = [pi], [letter[root]]
    for v, u in parent.items():
        P = scipy.linalg.expm(Q * blen[v] * rate)
        if v >= T:
            ops.append(P)
            subs.append(letter[u] + letter[v])
        else:
            c = int(codes_col[v])
            ops.append(P[:, c] if c < 20 else P.sum(axis=1))
            subs.append(letter[u])
    return float(np.einsum(",".join(subs) + "->", *ops, optimize=False))


def lnl(codes, children, blen, base, logits, assignment, pinv, rates):
    T, S = codes.shape
    eig = [build_q(base, lg) for lg in logits]
    tot = 0.0
    for s in range(S):
        Q, pi = eig[assignment[s]]
        col = codes[:, s]
        L = sum(site_likelihood_rate(col, children, blen, Q, pi, r) for r in rates)
        inv = 1.0 if np.all(col == col[0]) else 0.0
        ic = int(col[0]) if (inv == 1.0 and col[0] < 20) else 0
        tot += np.log(L * (1.0 - pinv) + pinv * pi[ic] * inv)
    return tot


def lnl_logdomain(codes, children, blen, base, logits, assignment, pinv, rates):
    """The same likelihood enumerated entirely in log space (logsumexp over all
    internal-state assignments, log P from scipy expm): valid where the site
    likelihood underflows double precision."""
    from scipy.special import logsumexp
    T, S = codes.shape
    internal = list(range(T, 2 * T - 1))
    parent = {}
    for k, n in enumerate(internal):
        for c in children[k]:
            parent[int(c)] = n
    root = [n for n in internal if n not in parent][0]
    order = internal  # enumerate joint internal states
    import itertools
    tot = 0.0
    for s in range(S):
        lg = logits[assignment[s]]
        Q, pi = build_q(base, lg)
        logpi = lg - logsumexp(lg)
        col = codes[:, s]
        terms = []
        for r in rates:
            logP = {v: np.log(np.clip(scipy.linalg.expm(Q * blen[v] * r), 1e-320, None))
                    for v in parent}
            vals = []
            for states in itertools.product(range(NS), repeat=len(order)):
                st = dict(zip(order, states))
                acc = logpi[st[root]]
                for v, u in parent.items():
                    if v >= T:
                        acc += logP[v][st[u], st[v]]
                    else:
                        c = int(col[v])
                        acc += logP[v][st[u], c] if c < 20 else logsumexp(logP[v][st[u]])
                vals.append(acc)
            terms.append(logsumexp(vals))
        l = logsumexp(terms)
        inv = 1.0 if np.all(col == col[0]) else 0.0
        ic = int(col[0]) if (inv == 1.0 and col[0] < 20) else 0
        a = l + np.log1p(-pinv)
        b = np.log(pinv * np.exp(logpi[ic]) * inv) if inv and pinv > 0 else -np.inf
        tot += np.logaddexp(a, b)
    return tot
