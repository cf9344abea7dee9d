"""Sequential model of the device enumeration algorithm (dumato_amd/csrc).

It follows the same levels, job ranges, donation rule and leaf counting as the
HIP kernels, one step at a time, so the tests can check on the CPU that the
restated algorithm gives the oracle's counts under ANY donation schedule
(the conservation property the on-device work stealing relies on).

Motif state per level l (prefix id[0..l]): U_l = sorted union of N(id[i]) for
i <= l restricted to ids > id[0], each element with the mask pm of the lists
that contain it; E_l = the elements that pass the canonical rule (valid).
Clique state per level l: C_l = N+(id[0]) & ... & N+(id[l]).
A job is (level, prefix ids, lo, hi): process children E_level[lo:hi].
"""
from collections import deque

import numpy as np


def quick_off(s):
    return (s - 2) * (s + 1) // 2


class MotifModel:
    def __init__(self, g, K, lut, ncls):
        self.g, self.K, self.lut, self.ncls = g, K, lut, ncls
        self.nb = [set(g.neighbours(v).tolist()) for v in range(g.n)]

    def root_level(self, v):
        U = [(int(u), 1, True) for u in self.g.neighbours(v) if u > v]
        return U

    def forward(self, U, c, l1, v0):
        """U_{l1} from U_{l1-1} and child c (placed at position l1)."""
        nc = self.nb[c]
        old = {u: (pm, val) for (u, pm, val) in U}
        out = []
        for u in sorted(set(old) | {x for x in nc if x > v0}):
            if u in old:
                pm, val = old[u]
                if u in nc:
                    pm |= 1 << l1
                out.append((u, pm, val and u > c))
            else:
                out.append((u, 1 << l1, True))
        return out

    def rebuild(self, ids):
        v0 = ids[0]
        U = [self.root_level(v0)]
        ind = [0, 0]
        for l in range(1, len(ids)):
            c = ids[l]
            pmc = [pm for (u, pm, val) in U[l - 1] if u == c][0]
            if l >= 2:
                ind.append(ind[l - 1] + (pmc << quick_off(l)))
            U.append(self.forward(U[l - 1], c, l, v0))
        return U, ind[:max(len(ids), 2)]

    def leaf_phase(self, ids, U, ind, lo, hi, hist, donate, push):
        """children E[lo:hi] of the parent level l = K-3, descending."""
        K = self.K
        l = len(ids) - 1
        E = [(u, pm) for (u, pm, val) in U if val]
        idx = {u: (pm, val) for (u, pm, val) in U}
        pos_in_E = {u: j for j, (u, pm) in enumerate(E)}
        v0 = ids[0]
        n_gt = {}
        for (u, pm) in E[hi:]:
            n_gt[pm] = n_gt.get(pm, 0) + 1
        i = hi - 1
        s_off = quick_off(l + 2)
        while i >= lo:
            if donate() and i - lo >= 1:
                mid = (lo + i + 1) // 2
                push((l, list(ids), lo, mid))
                # keep [mid, i]: n_gt is unchanged (it counts indices > i)
                lo = mid
            c, pmc = E[i]
            indc = ind[l] + ((pmc << quick_off(l + 1)) if l + 1 >= 2 else 0)
            n_new = 0
            n_int = {}
            for u in self.nb[c]:
                if u <= v0:
                    continue
                if u not in idx:
                    n_new += 1
                elif idx[u][1] and pos_in_E[u] > i:
                    pm = idx[u][0]
                    n_int[pm] = n_int.get(pm, 0) + 1
            top = 1 << (l + 1)
            if n_new:
                hist[self.lut[indc + (top << s_off)]] += n_new
            for pm in set(n_gt) | set(n_int):
                a = n_int.get(pm, 0)
                b = n_gt.get(pm, 0) - a
                if a:
                    hist[self.lut[indc + ((pm | top) << s_off)]] += a
                if b:
                    hist[self.lut[indc + (pm << s_off)]] += b
            n_gt[pmc] = n_gt.get(pmc, 0) + 1
            i -= 1

    def run_job(self, job, hist, donate, push):
        lvl, ids, lo, hi = job
        K = self.K
        U, ind = self.rebuild(ids)
        # stack of (ids, U, ind, cursor, end) from level lvl upward
        E = [(u, pm) for (u, pm, val) in U[lvl] if val]
        if lvl == K - 3:
            self.leaf_phase(ids, U[lvl], ind, lo, min(hi, len(E)), hist, donate, push)
            return
        stack = [[list(ids), U[lvl], list(ind), lo, min(hi, len(E)), E]]
        while stack:
            top = stack[-1]
            # donation: lowest level with >= 2 children left
            if donate():
                for fr in stack:
                    if fr[4] - fr[3] >= 2:
                        mid = (fr[3] + fr[4]) // 2
                        push((len(fr[0]) - 1, list(fr[0]), mid, fr[4]))
                        fr[4] = mid
                        break
            fids, fU, find, cur, end, fE = top
            if cur >= end:
                stack.pop()
                continue
            top[3] += 1
            c, pmc = fE[cur]
            l1 = len(fids)
            nids = fids + [c]
            nind = find + ([find[l1 - 1] + (pmc << quick_off(l1))] if l1 >= 2 else [])
            nind = nind[:max(l1 + 1, 2)]
            if len(nind) < 2:
                nind = [0, 0]
            nU = self.forward(fU, c, l1, fids[0])
            if l1 == K - 3:
                nE = [(u, pm) for (u, pm, val) in nU if val]
                self.leaf_phase(nids, nU, nind, 0, len(nE), hist, donate, push)
            else:
                nE = [(u, pm) for (u, pm, val) in nU if val]
                stack.append([nids, nU, nind, 0, len(nE), nE])

    def run(self, donate_prob=0.0, seed=0, roots=None):
        rng = np.random.default_rng(seed)
        hist = np.zeros(self.ncls, dtype=np.int64)
        q = deque((0, [v], 0, 1 << 30) for v in (range(self.g.n) if roots is None else roots))
        donate = (lambda: rng.random() < donate_prob) if donate_prob > 0 else (lambda: False)
        while q:
            self.run_job(q.popleft(), hist, donate, q.append)
        return int(hist.sum()), hist


class CliqueModel:
    def __init__(self, go, K):
        """go: oriented CSR (neighbours with larger ids only)."""
        self.g, self.K = go, K
        self.nb = [go.neighbours(v).tolist() for v in range(go.n)]
        self.nbs = [set(x) for x in self.nb]

    def rebuild(self, ids):
        C = [self.nb[ids[0]]]
        for l in range(1, len(ids)):
            C.append([u for u in C[l - 1] if u in self.nbs[ids[l]]])
        return C

    def run_job(self, job, donate, push):
        lvl, ids, lo, hi = job
        K = self.K
        if K == 2:
            return len(self.nb[ids[0]])
        C = self.rebuild(ids)
        count = 0
        stack = [[list(ids), C[lvl], lo, min(hi, len(C[lvl]))]]
        while stack:
            if donate():
                for fr in stack:
                    if fr[3] - fr[2] >= 2:
                        mid = (fr[2] + fr[3]) // 2
                        push((len(fr[0]) - 1, list(fr[0]), mid, fr[3]))
                        fr[3] = mid
                        break
            top = stack[-1]
            fids, fC, cur, end = top
            if cur >= end:
                stack.pop()
                continue
            top[2] += 1
            c = fC[cur]
            later = fC[cur + 1:]
            nC = [u for u in later if u in self.nbs[c]]
            if len(fids) == K - 2:  # leaf phase: count |N+(c) & C_l[i+1:]|
                count += len(nC)
            else:
                stack.append([fids + [c], nC, 0, len(nC)])
        return count

    def run(self, donate_prob=0.0, seed=0):
        rng = np.random.default_rng(seed)
        q = deque((0, [v], 0, 1 << 30) for v in range(self.g.n))
        donate = (lambda: rng.random() < donate_prob) if donate_prob > 0 else (lambda: False)
        total = 0
        while q:
            total += self.run_job(q.popleft(), donate, q.append)
        return total
