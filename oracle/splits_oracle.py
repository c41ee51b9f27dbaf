"""CPU restatement of the split machinery (test infrastructure only; the product
never imports it).

  splits(children, blen, T)       Tree::computeSplits + getSplitSet + getSplitBranchMap
                                  (Tree.cpp:485-535): the tips below every non-root node,
                                  flipped when they do not hold taxon 0; lengths summed per
                                  split (the root's two children give one split)
  posteriors(trees, weights, T)   computeSplitPosteriors (Miscellaneous.hpp:115-150) and the
                                  weighted mean branch length per split (main.cpp:367-383)
  compatible(a, b, T)             the four-point condition of main.cpp:353-357
  greedy_consensus(post, T)       main.cpp:344-360: decreasing posterior, keep a split when
                                  compatible with every kept one

Splits are Python ints, bit t = taxon t.  Ties in posterior are broken by the caller's
order (the reference iterates an unordered_map, so its tie order is unspecified).
"""


def splits(children, blen, T):
    nn = 2 * T - 1
    sets = [0] * nn
    for t in range(T):
        sets[t] = 1 << t
    parent = [-1] * nn
    for k in range(T - 1):
        for c in children[k]:
            parent[int(c)] = T + k
    root = [n for n in range(T, nn) if parent[n] < 0][0]
    order = []
    stack = [root]
    while stack:  # pre-order of the internal nodes; reversed, children come first
        n = stack.pop()
        order.append(n)
        for c in children[n - T]:
            if int(c) >= T:
                stack.append(int(c))
    for n in reversed(order):
        a, b = (int(c) for c in children[n - T])
        sets[n] = sets[a] | sets[b]
    full = (1 << T) - 1
    out = {}
    for n in range(nn):
        if n == root:
            continue
        s = sets[n] if sets[n] & 1 else full & ~sets[n]
        out[s] = out.get(s, 0.0) + float(blen[n])
    return out


def posteriors(trees, weights, T):
    post, wlen = {}, {}
    for (children, blen), w in zip(trees, weights):
        for s, ln in splits(children, blen, T).items():
            post[s] = post.get(s, 0.0) + w
            wlen[s] = wlen.get(s, 0.0) + w * ln
    return {s: (post[s], wlen[s] / post[s] if post[s] > 0 else 0.0) for s in post}


def compatible(a, b, T):
    full = (1 << T) - 1
    na, nb = full & ~a, full & ~b
    return (a & b) == 0 or (a & nb) == 0 or (na & b) == 0 or (na & nb) == 0


def greedy_consensus(ordered_splits, T):
    kept = []
    for s in ordered_splits:
        if len(kept) >= 2 * T - 3:
            break
        if all(compatible(s, k, T) for k in kept):
            kept.append(s)
    return kept
