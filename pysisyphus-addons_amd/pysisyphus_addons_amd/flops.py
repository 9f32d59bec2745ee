"""Algorithmic FP64 operation counts of the 3c2e kernel, per primitive triplet
(one primitive pair of the shell pair x one primitive of the aux shell).

Counted (an FMA counts 2, every add/mul 1; exp/sqrt/divide 1):
  * scaled Boys values (-alpha)^N F_N(T), N = 0..L: 7-term Taylor (14) plus
    the power and scale (2) per order;
  * powers (2 X_d)^a, a = 0..L: 1 each;
  * closed-form R_tuv for orders lc..L: 3 per innermost term plus 2 per
    middle-loop term;
  * the Hobson transform W[tuv][m] = sum_terms c R: 2 per solid-harmonic term;
  * the bra contraction acc[mu nu][m] += Et[mu nu][tuv] W[tuv][m]:
    2 nsph(la) nsph(lb) nsph(lc) nherm(la+lb).
Work done once per primitive pair and shared by a chunk of aux shells (the
spherical bra matrix) is not counted, so the figure is a lower bound of the
kernel's work.
"""

from functools import lru_cache

import numpy as np


def nherm(L):
    return (L + 1) * (L + 2) * (L + 3) // 6


def herm_list(L):
    out = []
    for s in range(L + 1):
        for t in range(s, -1, -1):
            for u in range(s - t, -1, -1):
                out.append((t, u, s - t - u))
    return out


@lru_cache(maxsize=None)
def sph_nterms(L):
    """Number of Cartesian terms of each real solid harmonic (m = -L..L)."""
    from math import comb

    counts = []
    for m in range(-L, L + 1):
        am = abs(m)
        vm2 = 1 if m < 0 else 0
        mons = set()
        for t in range((L - am) // 2 + 1):
            for u in range(t + 1):
                for v2 in range(vm2, am + 1, 2):
                    lx = 2 * t + am - 2 * u - v2
                    ly = 2 * u + v2
                    mons.add((lx, ly))
        # coefficients of distinct (t,u,v) can cancel only for m=0 symmetric sums;
        # the kernel table stores the non-zero ones, counted here from exact sums
        tot = {}
        for t in range((L - am) // 2 + 1):
            for u in range(t + 1):
                for v2 in range(vm2, am + 1, 2):
                    lx = 2 * t + am - 2 * u - v2
                    ly = 2 * u + v2
                    sgn = -1 if ((t + (v2 - vm2) // 2) & 1) else 1
                    c = sgn * 0.25 ** t * comb(L, t) * comb(L - t, am + t) * comb(t, u) * comb(am, v2)
                    tot[(lx, ly)] = tot.get((lx, ly), 0.0) + c
        counts.append(sum(1 for v in tot.values() if v != 0.0))
    return tuple(counts)


@lru_cache(maxsize=None)
def flops_per_prim_triplet(la, lb, lc):
    L = la + lb + lc
    H = nherm(la + lb)
    f = (L + 1) * 16 + 3 * (L + 1)
    for (t, u, v) in herm_list(L):
        if t + u + v < lc:
            continue
        inner = 0
        middle = 0
        for i in range(t // 2 + 1):
            for j in range(u // 2 + 1):
                middle += 1
                inner += v // 2 + 1
        f += 3 * inner + 2 * middle
    f += 2 * H * sum(sph_nterms(lc))
    f += 2 * (2 * la + 1) * (2 * lb + 1) * (2 * lc + 1) * H
    return f


def class_index(la, lb, lc):
    return (la * (la + 1) // 2 + lb) * 6 + lc


def total_flops(class_prim_triplets):
    """Sum over classes of primitive-triplet counts x flops (stats layout)."""
    tot = 0.0
    for la in range(5):
        for lb in range(la + 1):
            for lc in range(6):
                n = class_prim_triplets[class_index(la, lb, lc)]
                if n:
                    tot += n * flops_per_prim_triplet(la, lb, lc)
    return tot
