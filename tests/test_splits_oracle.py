"""The split oracle (oracle/splits_oracle.py) pinned by closed forms: a binary tree
on T taxa has 2T - 3 distinct splits (T trivial ones), all pairwise compatible, with
lengths summing to the tree length; identical trees give posterior 1 and their own
split set as the consensus; a two-tree mixture gives w, 1 - w and 1 for shared
splits; and the four-point condition equals its definition by brute force over
the subsets of a 5-taxon set."""
import itertools
import sys

import numpy as np

import splits_oracle as so
from smccat import simulate as sim


def test_split_counts_lengths_compatibility():
    rng = np.random.default_rng(1)
    for T in (3, 5, 12, 70):
        ch, bl, root = sim.random_tree(rng, T)
        sp = so.splits(ch, bl, T)
        assert len(sp) == 2 * T - 3
        assert sum(1 for s in sp if bin(s).count("1") in (1, T - 1)) == T
        assert abs(sum(sp.values()) - sum(bl[n] for n in range(2 * T - 1) if n != root)) < 1e-12
        keys = list(sp)
        assert all(so.compatible(a, b, T) for a, b in itertools.combinations(keys, 2))
        assert all(s & 1 for s in keys)


def test_identical_trees_and_mixture():
    rng = np.random.default_rng(2)
    T = 9
    t1 = sim.random_tree(rng, T)[:2]
    t2 = sim.random_tree(rng, T)[:2]
    post = so.posteriors([t1] * 5, [0.2] * 5, T)
    assert all(abs(p - 1.0) < 1e-12 for p, _ in post.values())
    order = sorted(post, key=lambda s: -post[s][0])
    assert set(so.greedy_consensus(order, T)) == set(so.splits(*t1, T))
    w = 0.3
    post = so.posteriors([t1, t2], [w, 1 - w], T)
    s1, s2 = set(so.splits(*t1, T)), set(so.splits(*t2, T))
    for s, (p, _) in post.items():
        want = (w if s in s1 else 0.0) + (1 - w if s in s2 else 0.0)
        assert abs(p - want) < 1e-12


def _all_rooted_trees(T):
    """Every rooted binary tree on T taxa as children arrays (tip i joins the pendant
    edge of any earlier node), with unit branch lengths."""
    out = []

    def grow(edges, children, nxt, i):
        if i == T:
            ch = np.zeros((T - 1, 2), dtype=np.int32)
            for n, (a, b) in children.items():
                ch[n - T] = (a, b)
            out.append((ch, np.ones(2 * T - 1)))
            return
        for child in list(edges):  # split the edge above `child`
            par = edges[child]
            ch2 = dict(children)
            e2 = dict(edges)
            if par is not None:
                a, b = ch2[par]
                ch2[par] = (nxt, b) if a == child else (a, nxt)
            ch2[nxt] = (i, child)
            e2[nxt], e2[i], e2[child] = par, nxt, nxt
            grow(e2, ch2, nxt + 1, i + 1)

    grow({0: T, 1: T, T: None}, {T: (0, 1)}, T + 1, 2)
    return out


def test_four_point_condition_against_tree_enumeration():
    """Two splits are compatible exactly when some binary tree holds both: checked
    over all 105 rooted binary trees on 5 taxa and every pair of splits."""
    T = 5
    trees = _all_rooted_trees(T)
    assert len(trees) == 105
    together, seen = set(), set()
    for ch, bl in trees:
        sp = list(so.splits(ch, bl, T))
        seen.update(sp)
        together.update((a, b) for a in sp for b in sp)
    assert len(seen) == 2 ** (T - 1) - 1  # every bipartition with taxon 0 on the set side
    for a in seen:
        for b in seen:
            assert so.compatible(a, b, T) == ((a, b) in together), (bin(a), bin(b))
