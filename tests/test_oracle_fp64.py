"""Pins the oracle's FP64 SAT (oracle/oibvh_oracle.c, sat_f64: the 17-axis
test of utils.cpp:71-169 evaluated in double from widened FP32 inputs)
against exact arithmetic, independently of anything the oracle wrote.

Two input families make every double intermediate of the SAT exact:

* (a) a small integer grid (coordinates in [-2, 2]): FP32 is exact there
  too, and exact ties (touching triangles) are frequent;
* (b) integer coordinates in [-2^10, 2^10] arranged to touch or nearly
  touch. After the translation to P1 every coordinate and edge is at most
  2^11 in magnitude, cross products of edges at most 2^23, the g/h axes at
  most 2^35 and their dot products with points at most 3 * 2^46 < 2^53: all
  integers that double holds exactly, while FP32 (24-bit significand) rounds
  every product above 2^24.

On both, the expected outcome is the SAT evaluated in exact integer
arithmetic, and (for non-degenerate triangles) an independent exact
predicate: the triangles intersect iff some convex combination of P's
vertices equals one of Q's, decided by an exact simplex over Fractions. The
second check pins the axis list itself (a dropped axis lets separated pairs
through) and the closed comparison (touching counts as intersecting).
"""
from fractions import Fraction

import numpy as np
import pytest

LIM = 1 << 10


# ----------------------------------------------------------------- exact SAT
def _sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def _cross(a, b):
    return (a[1] * b[2] - b[1] * a[2], a[2] * b[0] - b[2] * a[0], a[0] * b[1] - b[0] * a[1])


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def sat_axes(P, Q):
    """The 17 axes of utils.cpp:97-169 (n1, m1, e_i x f_j, e_i x n1, f_j x m1)
    in exact integer arithmetic, P1-translated frame. Returns the indices of
    the separating axes and whether some axis has an exact zero gap between
    the two projections (touching)."""
    P1, P2, P3 = P
    Q1, Q2, Q3 = Q
    p1 = (0, 0, 0)
    p2, p3 = _sub(P2, P1), _sub(P3, P1)
    q1, q2, q3 = _sub(Q1, P1), _sub(Q2, P1), _sub(Q3, P1)
    e = [_sub(p2, p1), _sub(p3, p2), _sub(p1, p3)]
    f = [_sub(q2, q1), _sub(q3, q2), _sub(q1, q3)]
    n1, m1 = _cross(e[0], e[1]), _cross(f[0], f[1])
    axes = [n1, m1] + [_cross(a, b) for a in e for b in f]
    axes += [_cross(a, n1) for a in e] + [_cross(b, m1) for b in f]
    sep, touching = [], False
    for k, ax in enumerate(axes):
        pp = [_dot(ax, p) for p in (p1, p2, p3)]
        qq = [_dot(ax, q) for q in (q1, q2, q3)]
        if min(pp) > max(qq) or min(qq) > max(pp):
            sep.append(k)
        elif min(pp) == max(qq) or min(qq) == max(pp):
            touching = True
    return sep, touching


def exact_sat(P, Q):
    """(intersect, touching) of the exact 17-axis SAT"""
    sep, touching = sat_axes(P, Q)
    return not sep, touching


# --------------------------------------------------- independent exact test
def _feasible(A, b):
    """Is {x >= 0 : A x = b} non-empty? Phase 1 of the simplex method over
    Fractions with Bland's rule (b >= 0)."""
    m, n = len(A), len(A[0])
    T = [[Fraction(v) for v in A[i]] + [Fraction(int(i == k)) for k in range(m)] + [Fraction(b[i])]
         for i in range(m)]
    basis = [n + i for i in range(m)]
    cost = [0] * n + [1] * m
    while True:
        enter = None
        for j in range(n + m):
            if j in basis:
                continue
            r = cost[j] - sum(cost[basis[i]] * T[i][j] for i in range(m))
            if r < 0:
                enter = j
                break
        if enter is None:
            break
        best = None
        for i in range(m):
            if T[i][enter] > 0:
                ratio = T[i][-1] / T[i][enter]
                if best is None or ratio < best[0] or (ratio == best[0] and basis[i] < basis[best[1]]):
                    best = (ratio, i)
        r = best[1]
        pv = T[r][enter]
        T[r] = [v / pv for v in T[r]]
        for i in range(m):
            if i != r and T[i][enter] != 0:
                fac = T[i][enter]
                T[i] = [a - fac * c for a, c in zip(T[i], T[r])]
        basis[r] = enter
    return sum(cost[basis[i]] * T[i][-1] for i in range(m)) == 0


def exact_intersect(P, Q):
    """Closed triangles P, Q share a point: lambda, mu >= 0 with sum 1 each and
    sum lambda_i P_i = sum mu_j Q_j."""
    A = [[1, 1, 1, 0, 0, 0], [0, 0, 0, 1, 1, 1]]
    for k in range(3):
        A.append([P[0][k], P[1][k], P[2][k], -Q[0][k], -Q[1][k], -Q[2][k]])
    return _feasible(A, [1, 1, 0, 0, 0])


def _degenerate(T):
    return _cross(_sub(T[1], T[0]), _sub(T[2], T[0])) == (0, 0, 0)


# ----------------------------------------------------------------- inputs
def grid_pairs(n, seed):
    """(a) coordinates in [-2, 2]; a third of the pairs share a vertex or lie
    in one plane, so ties are frequent."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        P = [tuple(int(v) for v in rng.integers(-2, 3, 3)) for _ in range(3)]
        Q = [tuple(int(v) for v in rng.integers(-2, 3, 3)) for _ in range(3)]
        if k % 3 == 1:
            Q[int(rng.integers(0, 3))] = P[int(rng.integers(0, 3))]
        elif k % 3 == 2:
            for T in (P, Q):
                for i in range(3):
                    T[i] = (T[i][0], T[i][1], 0)
        out.append((P, Q))
    return out


def _lattice_triangle(rng, d):
    """P1 plus edge vectors that are multiples of d, inside [-LIM, LIM]^3"""
    while True:
        P1 = rng.integers(-LIM // 2, LIM // 2 + 1, 3)
        u = rng.integers(-LIM // (2 * d), LIM // (2 * d) + 1, 3) * d
        v = rng.integers(-LIM // (2 * d), LIM // (2 * d) + 1, 3) * d
        P = [P1, P1 + u, P1 + v]
        if np.abs(np.array(P)).max() <= LIM and np.any(np.cross(u, v) != 0):
            return P, u, v


def near_touching_pairs(n, seed):
    """(b) integer coordinates in [-2^10, 2^10] (exact in double), where Q
    touches P at a point of P (a vertex, an edge point or an interior point
    on the lattice P1 + (a u + b v) / d), or lies in P's plane, or is pushed
    one unit off such a contact."""
    rng = np.random.default_rng(seed)
    d = 4
    out = []
    while len(out) < n:
        P, u, v = _lattice_triangle(rng, d)
        a = int(rng.integers(0, d + 1))
        b = int(rng.integers(0, d + 1 - a))
        c = P[0] + (a * u + b * v) // d  # a point of P (exact: u, v are multiples of d)
        kind = len(out) % 4
        w1 = rng.integers(-300, 301, 3)
        w2 = rng.integers(-300, 301, 3)
        if kind == 0:  # Q has a vertex at c, its other vertices anywhere
            Q = [c, c + w1, c + w2]
        elif kind == 1:  # Q's edge passes through c (c is the midpoint of Q1 Q2)
            Q = [c + w1, c - w1, c + w2]
        elif kind == 2:  # Q in P's plane, vertices on P's lattice
            co = rng.integers(-d, 2 * d + 1, (2, 2))
            Q = [c, P[0] + (co[0, 0] * u + co[0, 1] * v) // d, P[0] + (co[1, 0] * u + co[1, 1] * v) // d]
        else:  # as kind 0, the contact vertex pushed one unit along an axis
            Q = [c + np.eye(3, dtype=np.int64)[int(rng.integers(0, 3))] * int(rng.choice([-1, 1])),
                 c + w1, c + w2]
        Q = [np.asarray(q, np.int64) for q in Q]
        if max(np.abs(np.array(Q)).max(), np.abs(np.array(P)).max()) > LIM:
            continue
        perm = rng.permutation(3)
        Q = [Q[i] for i in perm]
        out.append(([tuple(int(x) for x in p) for p in P], [tuple(int(x) for x in q) for q in Q]))
    return out


def _f32(T):
    return np.asarray(T, np.float32)


# ----------------------------------------------------------------- tests
def test_exact_sat_equals_exact_intersection():
    """The exact 17-axis SAT decides closed-triangle intersection: it agrees
    with the barycentric LP on every non-degenerate pair of both families."""
    pairs = grid_pairs(1500, 0x5A7) + near_touching_pairs(400, 0x5A8)
    checked = touching = 0
    for P, Q in pairs:
        if _degenerate(P) or _degenerate(Q):
            continue
        s, t = exact_sat(P, Q)
        assert s == exact_intersect(P, Q), (P, Q)
        checked += 1
        touching += s and t
    assert checked > 1200 and touching > 100


def test_fp64_sat_exact_on_small_grid(orc):
    pairs = grid_pairs(3000, 0x5A9)
    hits = ties = 0
    sole = set()
    for P, Q in pairs:
        sep, tie = sat_axes(P, Q)
        want = not sep
        if len(sep) == 1:
            sole.add(sep[0])
        assert orc.tri_intersect(_f32(P), _f32(Q), fp64=True) == want, (P, Q)
        # on this grid FP32 is exact too
        assert orc.tri_intersect(_f32(P), _f32(Q)) == want, (P, Q)
        hits += want
        ties += want and tie
    assert 300 < hits < 2700 and ties > 200
    # every axis is the only separating one for some pair, so an oracle
    # missing any axis (or testing a wrong one) fails above
    assert sole == set(range(17)), sorted(sole)


def test_fp64_sat_exact_on_10bit_coordinates(orc):
    pairs = near_touching_pairs(3000, 0x5AA)
    f32_wrong = hits = ties = 0
    for P, Q in pairs:
        want, tie = exact_sat(P, Q)
        assert orc.tri_intersect(_f32(P), _f32(Q), fp64=True) == want, (P, Q)
        f32_wrong += orc.tri_intersect(_f32(P), _f32(Q)) != want
        hits += want
        ties += want and tie
    # the family separates FP32 from exact arithmetic (a double evaluation
    # that rounded like FP32 anywhere would fail above), and it holds both
    # outcomes and exact contacts
    assert f32_wrong >= 10, f32_wrong
    assert 300 < hits < 2900 and ties > 300


def test_fp64_sat_special_cases(orc):
    """The SAT special cases, in the FP64 form."""
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    t = lambda A, B: orc.tri_intersect(A, B, fp64=True)  # noqa: E731
    assert t(P, P)
    assert not t(P, P + np.float32(5))
    assert t(P, np.array([[1, 0, 0], [2, 0, 1], [2, 1, -1]], np.float32))  # shared vertex only
    assert t(P, np.array([[0.2, 0.2, 0], [2, 0.2, 0], [0.2, 2, 0]], np.float32))  # coplanar overlap
    assert not t(P, np.array([[0.8, 0.8, 0], [2, 0.8, 0], [0.8, 2, 0]], np.float32))  # coplanar, apart
    assert t(P, np.array([[0.2, 0.2, -1], [0.2, 0.2, 1], [3, 3, 0.5]], np.float32))  # piercing
    assert t(P, np.array([[0.1, 0.1, 0]] * 3, np.float32))  # zero-area triangle: zero axes never separate
    # a point triangle outside P along its normal: only n1 separates it
    assert not t(P, np.array([[0.1, 0.1, 0.5]] * 3, np.float32))
    # edge-edge contact at one point, planes crossing
    assert t(P, np.array([[0.5, 0.5, 0], [0.5, 0.5, 1], [1.5, 1.5, 0.5]], np.float32))


@pytest.mark.parametrize("scale", [1.0, 2.0 ** -20, 2.0 ** 20])
def test_fp64_sat_scale_invariant(orc, scale):
    """Every axis comparison is between quantities of one degree, so scaling
    all coordinates by a power of two (exact in both precisions) keeps the
    FP64 outcome; the 10-bit family must keep its exact answers."""
    for P, Q in near_touching_pairs(300, 0x5AB):
        want, _ = exact_sat(P, Q)
        got = orc.tri_intersect(_f32(P) * np.float32(scale), _f32(Q) * np.float32(scale), fp64=True)
        assert got == want, (P, Q, scale)
