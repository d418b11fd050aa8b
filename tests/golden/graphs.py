"""Named small graphs used as known-answer cases (edge lists, each edge once)."""
import itertools

import numpy as np


def _arr(edges):
    U = np.array([a for a, _ in edges], dtype=np.int32)
    V = np.array([b for _, b in edges], dtype=np.int32)
    return U, V


def complete(n):
    return n, *_arr(list(itertools.combinations(range(n), 2)))


def complete_bipartite(a, b):
    return a + b, *_arr([(i, a + j) for i in range(a) for j in range(b)])


def prism():
    e = [(0, 1), (1, 2), (2, 0), (3, 4), (4, 5), (5, 3), (0, 3), (1, 4), (2, 5)]
    return 6, *_arr(e)


def rook4():
    e = [(a, b) for a in range(16) for b in range(a + 1, 16) if a // 4 == b // 4 or a % 4 == b % 4]
    return 16, *_arr(e)


def shrikhande():
    steps = {(0, 1), (0, 3), (1, 0), (3, 0), (1, 1), (3, 3)}
    e = [(a, b) for a in range(16) for b in range(a + 1, 16)
         if ((b // 4 - a // 4) % 4, (b % 4 - a % 4) % 4) in steps]
    return 16, *_arr(e)


def petersen():
    outer = [(i, (i + 1) % 5) for i in range(5)]
    spokes = [(i, i + 5) for i in range(5)]
    inner = [(5 + i, 5 + (i + 2) % 5) for i in range(5)]
    return 10, *_arr(outer + spokes + inner)


def path(n):
    return n, *_arr([(i, i + 1) for i in range(n - 1)])


def star(n):
    return n, *_arr([(0, i) for i in range(1, n)])


def both_directions(g):
    n, U, V = g
    return n, np.concatenate([U, V]), np.concatenate([V, U])


def named_graphs():
    return {
        "K3": complete(3), "K4": complete(4), "K5": complete(5), "K33": complete_bipartite(3, 3),
        "prism": prism(), "rook4": rook4(), "shrikhande": shrikhande(), "petersen": petersen(),
        "path6": path(6), "star7": star(7),
        "K3_both": both_directions(complete(3)), "K33_both": both_directions(complete_bipartite(3, 3)),
    }
