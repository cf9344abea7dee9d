"""Small deterministic graphs for the CPU tests (edge lists as numpy arrays)."""
import numpy as np


def complete(n):
    s, d = np.triu_indices(n, 1)
    return s.astype(np.int32), d.astype(np.int32), n


def star(m):
    """centre 0, leaves 1..m"""
    return np.zeros(m, np.int32), np.arange(1, m + 1, dtype=np.int32), m + 1


def path(n):
    a = np.arange(n - 1, dtype=np.int32)
    return a, a + 1, n


def cycle(n):
    a = np.arange(n, dtype=np.int32)
    return a, (a + 1) % n, n


def gnp(n, p, seed):
    rng = np.random.default_rng(seed)
    s, d = np.triu_indices(n, 1)
    keep = rng.random(len(s)) < p
    return s[keep].astype(np.int32), d[keep].astype(np.int32), n


def permute(src, dst, n, seed):
    perm = np.random.default_rng(seed).permutation(n).astype(np.int32)
    return perm[src], perm[dst], n


def py_brute(src, dst, n, k, lut):
    """Pure-Python brute force: clique count and motif histogram (dict)."""
    import itertools
    adj = [set() for _ in range(n)]
    for a, b in zip(src.tolist(), dst.tolist()):
        if a != b:
            adj[a].add(b)
            adj[b].add(a)
    cliques = 0
    hist = {}
    for S in itertools.combinations(range(n), k):
        if all(b in adj[a] for a, b in itertools.combinations(S, 2)):
            cliques += 1
        # connectivity
        seen = {S[0]}
        todo = [S[0]]
        while todo:
            x = todo.pop()
            for y in S:
                if y not in seen and y in adj[x]:
                    seen.add(y)
                    todo.append(y)
        if len(seen) != k:
            continue
        # any order whose first two vertices are adjacent
        for a, b in itertools.permutations(S, 2):
            if b in adj[a]:
                break
        rest = [x for x in S if x not in (a, b)]
        order = [a, b] + rest
        q = 0
        for s_ in range(2, k):
            for t in range(s_):
                if order[t] in adj[order[s_]]:
                    q |= 1 << ((s_ - 2) * (s_ + 1) // 2 + t)
        c = int(lut[q])
        hist[c] = hist.get(c, 0) + 1
    return cliques, hist
