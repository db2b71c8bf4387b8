"""Brute-force UniFrac / Faith PD straight from the definitions (pure Python).

Independent of every compiled path: no balanced-parentheses index caches, no
postorder batching, no stripes.  For every non-root node it takes the set of
observed leaves below it; the node's per-sample value is the sum of those
leaves' (normalised) counts.  Shear and collapse need no code here: nodes with
no observed leaf contribute zero to every metric, and a chain of single-child
nodes has one leaf set, so its lengths add up exactly as collapse adds them.
The one place collapse shows is bypass_tips: a node whose observed leaf set is
a single leaf is folded into that tip and skipped with it.

Definitions: Lozupone & Knight 2005 (unweighted), Lozupone et al. 2007
(weighted), Chen et al. 2012 (generalized), Chang et al. 2011 (VAW).
"""
import math


def nodes_from_bp(structure, lengths, names):
    """[(length, name, leaf_names)] for every non-root node."""
    stack, out = [], []
    children = {}
    for i, s in enumerate(structure):
        if s:
            stack.append(i)
            children[i] = []
        else:
            o = stack.pop()
            if stack:
                children[stack[-1]].append(o)
            out.append((o, stack[-1] if stack else None))
    # leaf sets, children before parents
    leaves = {}
    for o, _ in out:
        if not children[o]:
            leaves[o] = {names[o]}
        else:
            s = set()
            for c in children[o]:
                s |= leaves[c]
            leaves[o] = s
    return [(lengths[o], names[o], leaves[o]) for o, parent in out if parent is not None]


def distances(structure, lengths, names, table, method, alpha=1.0, vaw=False, bypass_tips=False, normalize=True):
    """table: dict sample -> dict obs -> count.  Returns dict (a, b) -> distance for a < b."""
    samples = list(table)
    totals = {s: sum(table[s].values()) for s in samples}
    observed = set()
    for s in samples:
        observed |= set(table[s])
    nodes = []
    for length, name, leafset in nodes_from_bp(structure, lengths, names):
        obs = leafset & observed
        if not obs:
            continue
        if bypass_tips and len(obs) == 1:
            continue
        nodes.append((length, obs))

    def value(s, obs, norm):
        v = sum(table[s].get(o, 0.0) for o in obs)
        return v / totals[s] if norm else v

    out = {}
    for ia in range(len(samples)):
        for ib in range(ia + 1, len(samples)):
            a, b = samples[ia], samples[ib]
            num = den = 0.0
            m = totals[a] + totals[b]
            for L, obs in nodes:
                u, v = value(a, obs, normalize), value(b, obs, normalize)
                w = 1.0
                if vaw:
                    mi = value(a, obs, False) + value(b, obs, False)
                    q = math.sqrt(mi * (m - mi))
                    if not q > 0:
                        continue
                    w = 1.0 / q
                if method.startswith("unweighted"):
                    pa, pb = u > 0, v > 0
                    num += L * w * (pa != pb)
                    den += L * w * (pa or pb)
                elif method.startswith("weighted"):
                    num += L * w * abs(u - v)
                    den += L * w * (u + v)
                elif method == "generalized":
                    s1, d1 = (u + v) * w, abs(u - v) * w
                    if s1 == 0:
                        continue
                    p = s1 ** alpha * L
                    num += p * d1 / s1
                    den += p
            if method in ("unweighted", "weighted_normalized", "generalized"):
                out[(ia, ib)] = num / den if den != 0 else float("nan")
            else:
                out[(ia, ib)] = num
    return out


def faith_pd(structure, lengths, names, table):
    samples = list(table)
    observed = set()
    for s in samples:
        observed |= set(table[s])
    res = []
    nodes = nodes_from_bp(structure, lengths, names)
    for s in samples:
        res.append(sum(L for L, _, leafset in nodes if any(table[s].get(o, 0) > 0 for o in leafset & observed)))
    return res
