"""Keypoint-graph BP problems shared by the tests: the 848x480 workload
(keypoints drawn uniformly on the random-dot pair) and tiny explicit graphs."""
from __future__ import annotations

import numpy as np


def workload(n: int = 2000, w: int = 848, h: int = 480, seed: int = 1, kp_seed: int = 17):
    """The random-dot fronto pair and n keypoints drawn uniformly (integer
    pixels) inside the 9x7 census border."""
    from doubleeye_amd import synth
    L, R, gt = synth.fronto_pair(w, h, seed)
    rng = np.random.default_rng(kp_seed)
    xs = rng.integers(4, w - 4, n)
    ys = rng.integers(3, h - 3, n)
    return L, R, np.stack([xs, ys], 1).astype(np.int32), gt


def random_tree(rng, n: int, max_lab: int = 4, dmax: int = 20, cmax: int = 10):
    """A random tree (node k hangs from a random earlier node) with 1..max_lab
    labels per node, integer disparities and integer unary costs."""
    nb = [[] for _ in range(n)]
    for k in range(1, n):
        p = int(rng.integers(0, k))
        nb[k].append(p)
        nb[p].append(k)
    off = np.zeros(n + 1, np.int32)
    adj = []
    for i in range(n):
        adj += sorted(nb[i])
        off[i + 1] = len(adj)
    nc = rng.integers(1, max_lab + 1, n).astype(np.int32)
    cd = np.full((n, 8), -1, np.int32)
    cc = np.zeros((n, 8), np.float32)
    for i in range(n):
        cd[i, :nc[i]] = rng.choice(dmax + 1, nc[i], replace=False)
        cc[i, :nc[i]] = rng.integers(0, cmax + 1, nc[i])
    return cd, cc, nc, off, np.array(adj, np.int32)


def random_graph(rng, n: int, deg: int = 3, max_lab: int = 8, dmax: int = 60, cmax: int = 40):
    """A random symmetric graph with loops, sorted CSR lists."""
    nb = [set() for _ in range(n)]
    for i in range(n):
        for j in rng.choice(n, min(deg, n - 1), replace=False):
            if j != i:
                nb[i].add(int(j))
                nb[int(j)].add(i)
    off = np.zeros(n + 1, np.int32)
    adj = []
    for i in range(n):
        adj += sorted(nb[i])
        off[i + 1] = len(adj)
    nc = rng.integers(0, max_lab + 1, n).astype(np.int32)
    cd = np.full((n, 8), -1, np.int32)
    cc = np.zeros((n, 8), np.float32)
    for i in range(n):
        cd[i, :nc[i]] = rng.choice(dmax + 1, nc[i], replace=False)
        cc[i, :nc[i]] = rng.integers(0, cmax + 1, nc[i])
    return cd, cc, nc, off, np.array(adj, np.int32)
