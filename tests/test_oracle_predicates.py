"""Pins the oracle's geometric predicates independently of its own formulas: exact rational
arithmetic (fractions) on inputs away from rounding boundaries, closed-form special cases
read off the reference (edge.rs, rect.rs, circle.rs, simple_polygon.rs, util/fpa.rs), and
brute force.  A dropped term, a swapped operand or a flipped comparison fails here."""
import ctypes as C
import math
import random
from fractions import Fraction as Fr

import numpy as np
import pytest

from oracle import oracle as O

L = O.lib()
f32 = np.float32


def arr(*v):
    a = np.array(v, dtype=np.float32)
    return a, a.ctypes.data_as(O.f32p)


def edge_edge(e1, e2):
    a, pa = arr(*e1)
    b, pb = arr(*e2)
    return L.or_edge_edge(pa, pb) == 1


def rect_edge(r, e):
    a, pa = arr(*r)
    b, pb = arr(*e)
    return L.or_rect_edge(pa, pb) == 1


def circle_edge(c, e):
    a, pa = arr(*c)
    b, pb = arr(*e)
    return L.or_circle_edge(pa, pb) == 1


def circle_rect(c, r):
    a, pa = arr(*c)
    b, pb = arr(*r)
    return L.or_circle_rect(pa, pb) == 1


def rnd32(rng, lo, hi):
    return float(f32(rng.uniform(lo, hi)))


# ------------------------------------------------------------------------ edge-edge
def exact_tu(e1, e2):
    x1, y1, x2, y2 = map(Fr, e1)
    x3, y3, x4, y4 = map(Fr, e2)
    den = (x2 - x1) * (y4 - y3) - (y2 - y1) * (x4 - x3)
    if den == 0:
        return None
    t = ((x2 - x4) * (y4 - y3) - (y2 - y4) * (x4 - x3)) / den
    u = ((x2 - x4) * (y2 - y1) - (y2 - y4) * (x2 - x1)) / den
    return t, u


def test_edge_edge_vs_exact():
    rng = random.Random(1)
    n_checked = n_hit = 0
    for _ in range(20000):
        e1 = [rnd32(rng, -100, 100) for _ in range(4)]
        e2 = [rnd32(rng, -100, 100) for _ in range(4)]
        tu = exact_tu(e1, e2)
        if tu is None:
            continue
        t, u = tu
        margin = Fr(1, 1000)
        if min(abs(t), abs(t - 1), abs(u), abs(u - 1)) < margin:
            continue          # too close to a segment end for f32 rounding
        expect = 0 < t < 1 and 0 < u < 1
        assert edge_edge(e1, e2) == expect, (e1, e2)
        n_checked += 1
        n_hit += expect
    assert n_checked > 15000 and 1000 < n_hit < n_checked


def test_edge_edge_special_cases():
    # parallel and collinear-overlapping segments are not collisions (edge.rs:212-215)
    assert not edge_edge((0, 0, 10, 0), (0, 1, 10, 1))
    assert not edge_edge((0, 0, 10, 0), (5, 0, 15, 0))
    # endpoints touching: t, u in [0, 1] inclusive (edge.rs:219)
    assert edge_edge((0, 0, 10, 0), (10, 0, 10, 10))
    assert edge_edge((0, 0, 10, 10), (0, 10, 5, 5))
    # T-junction and a near miss
    assert edge_edge((0, 0, 10, 0), (5, -5, 5, 0))
    assert not edge_edge((0, 0, 10, 0), (5, 0.5, 5, 5))
    # disjoint bounding boxes exit early
    assert not edge_edge((0, 0, 1, 1), (2, 2, 3, 3))


# ------------------------------------------------------------------------ rect-edge
def exact_seg_rect(r, e):
    """Liang-Barsky clip of the closed segment against the closed rectangle, in rationals."""
    x0, y0, x1, y1 = map(Fr, e)
    lo, hi = Fr(0), Fr(1)
    dx, dy = x1 - x0, y1 - y0
    for p, q in ((-dx, x0 - Fr(r[0])), (dx, Fr(r[2]) - x0), (-dy, y0 - Fr(r[1])), (dy, Fr(r[3]) - y0)):
        if p == 0:
            if q < 0:
                return False
        else:
            t = q / p
            if p < 0:
                lo = max(lo, t)
            else:
                hi = min(hi, t)
    return lo <= hi


def test_rect_edge_vs_exact():
    rng = random.Random(2)
    n = 0
    for _ in range(20000):
        r = sorted([rnd32(rng, -50, 50) for _ in range(2)]), sorted([rnd32(rng, -50, 50) for _ in range(2)])
        rect = (r[0][0], r[1][0], r[0][1], r[1][1])
        e = [rnd32(rng, -80, 80) for _ in range(4)]
        # skip near-tangent cases: distance of any corner to the line below a margin
        x0, y0, x1, y1 = map(Fr, e)
        dx, dy = x1 - x0, y1 - y0
        L2 = dx * dx + dy * dy
        if L2 == 0:
            continue
        near = False
        for cx, cy in ((rect[2], rect[3]), (rect[0], rect[3]), (rect[0], rect[1]), (rect[2], rect[1])):
            cr = (Fr(cx) - x0) * dy - (Fr(cy) - y0) * dx
            if cr * cr < L2 * Fr(1, 10 ** 6):
                near = True
        if near or min(abs(Fr(v) - Fr(w)) for v in e[0::2] for w in (rect[0], rect[2])) < Fr(1, 10 ** 4):
            continue
        if min(abs(Fr(v) - Fr(w)) for v in e[1::2] for w in (rect[1], rect[3])) < Fr(1, 10 ** 4):
            continue
        assert rect_edge(rect, e) == exact_seg_rect(rect, e), (rect, e)
        n += 1
    assert n > 10000


def test_rect_edge_special_cases():
    r = (0, 0, 10, 10)
    assert rect_edge(r, (5, 5, 6, 6))            # inside
    assert rect_edge(r, (-5, 5, 15, 5))          # crossing
    assert rect_edge(r, (10, -5, 10, 15))        # along a side (inclusive bounds)
    assert rect_edge(r, (-5, 10, 10, -5))        # diagonal through the lower-left corner region
    assert not rect_edge(r, (11, 0, 20, 10))     # outside the x range
    assert not rect_edge(r, (-5, 6, 6, 17))      # clips the corner region but misses


# ------------------------------------------------------------------------ circles
def exact_seg_dist2(p, e):
    px, py = map(Fr, p)
    x1, y1, x2, y2 = map(Fr, e)
    c, d = x2 - x1, y2 - y1
    l2 = c * c + d * d
    t = ((px - x1) * c + (py - y1) * d) / l2 if l2 else Fr(-1)
    t = min(max(t, Fr(0)), Fr(1))
    qx, qy = x1 + t * c, y1 + t * d
    return (px - qx) ** 2 + (py - qy) ** 2


def test_circle_edge_and_rect_vs_exact():
    rng = random.Random(3)
    n = 0
    for _ in range(20000):
        c = (rnd32(rng, -20, 20), rnd32(rng, -20, 20), rnd32(rng, 0.5, 15))
        e = [rnd32(rng, -30, 30) for _ in range(4)]
        d2 = exact_seg_dist2(c[:2], e)
        r2 = Fr(c[2]) ** 2
        if abs(d2 - r2) > r2 * Fr(1, 10 ** 4):
            assert circle_edge(c, e) == (d2 <= r2), (c, e)
            n += 1
        x = sorted([rnd32(rng, -20, 20) for _ in range(2)])
        y = sorted([rnd32(rng, -20, 20) for _ in range(2)])
        rect = (x[0], y[0], x[1], y[1])
        nx = min(max(Fr(c[0]), Fr(x[0])), Fr(x[1]))
        ny = min(max(Fr(c[1]), Fr(y[0])), Fr(y[1]))
        dd = (nx - Fr(c[0])) ** 2 + (ny - Fr(c[1])) ** 2
        if abs(dd - r2) > r2 * Fr(1, 10 ** 4):
            assert circle_rect(c, rect) == (dd <= r2)
    assert n > 10000
    # zero-length edge: the closest point is its start (param = -1, edge.rs:76-81)
    assert circle_edge((0, 0, 1), (0.5, 0.5, 0.5, 0.5))
    assert not circle_edge((0, 0, 1), (2, 0, 2, 0))


# ------------------------------------------------------------------------ FPA
def test_fpa_boundaries():
    eq = lambda a, b: L.or_fpa_eq(C.c_float(a), C.c_float(b)) == 1
    eps = float(np.finfo(np.float32).eps)
    assert eq(1.0, 1.0)
    assert eq(0.0, eps) and eq(0.0, -eps)                     # |a-b| <= EPSILON
    assert not eq(0.0, float(np.nextafter(f32(eps), f32(1)) * 2))
    big = f32(1000.0)
    four = float(np.frombuffer((np.frombuffer(big.tobytes(), np.int32) + 4).tobytes(), np.float32)[0])
    five = float(np.frombuffer((np.frombuffer(big.tobytes(), np.int32) + 5).tobytes(), np.float32)[0])
    assert eq(float(big), four) and not eq(float(big), five)  # ulps <= 4
    assert eq(four, float(big))
    # across the sign: the bit difference wraps and saturates, so only the epsilon test can pass
    assert eq(1e-30, -1e-30)
    assert not eq(1.0, -1.0)


# ------------------------------------------------------------------------ point in polygon
def pip(poly, p):
    a = np.ascontiguousarray(poly, dtype=np.float32)
    return L.or_pip(a.ctypes.data_as(O.f32p), len(a), C.c_float(p[0]), C.c_float(p[1]))


def test_pip_convex_vs_halfplanes():
    rng = random.Random(4)
    for trial in range(30):
        n = rng.randint(3, 12)
        ang = sorted(rng.uniform(0, 2 * math.pi) for _ in range(n))
        R = rng.uniform(5, 50)
        poly = np.array([[R * math.cos(a), R * math.sin(a)] for a in ang], dtype=np.float32)
        for _ in range(300):
            p = (rnd32(rng, -1.2 * R, 1.2 * R), rnd32(rng, -1.2 * R, 1.2 * R))
            # exact side tests against every (CCW) edge
            inside, near = True, False
            for i in range(n):
                x1, y1 = map(Fr, map(float, poly[i]))
                x2, y2 = map(Fr, map(float, poly[(i + 1) % n]))
                cr = (x2 - x1) * (Fr(p[1]) - y1) - (y2 - y1) * (Fr(p[0]) - x1)
                l = (x2 - x1) ** 2 + (y2 - y1) ** 2
                if cr * cr < l * Fr(1, 10 ** 6):
                    near = True
                inside &= cr > 0
            # rays passing within FPA distance of a vertex take the vertex branch
            if near or any(abs(float(v[1]) - p[1]) < 1e-4 for v in poly):
                continue
            assert pip(poly, p) == (1 if inside else 0), (poly, p)


def test_pip_vertex_on_ray():
    sq = np.array([[0, 0], [10, 0], [10, 10], [0, 10]], dtype=np.float32)
    assert pip(sq, (5, 5)) == 1
    diamond = np.array([[5, 0], [10, 5], [5, 10], [0, 5]], dtype=np.float32)
    # the ray from (2, 5) passes exactly through the vertex (10, 5): counted once (edge below)
    assert pip(diamond, (2, 5)) == 1
    assert pip(diamond, (-1, 5)) == 0 and pip(diamond, (11, 5)) == 0


# ------------------------------------------------------------------------ quadrant test
def test_edge_quadrants_superset():
    """Edge::collides_with_quadrants flags every quadrant the edge touches (the reference's
    debug_assert, qt_traits.rs:87-102)."""
    rng = random.Random(5)
    for _ in range(20000):
        r = (0.0, 0.0, 64.0, 64.0)
        e = [rnd32(rng, -10, 74) for _ in range(4)]
        a, pa = arr(*e)
        b, pb = arr(*r)
        mask = L.or_edge_quadrants(pa, pb)
        qs = [(32, 32, 64, 64), (0, 32, 32, 64), (0, 0, 32, 32), (32, 0, 64, 32)]
        for k, q in enumerate(qs):
            if rect_edge(q, e):
                assert mask >> k & 1, (e, k, mask)


def quad_mask(e, r):
    a, pa = arr(*e)
    b, pb = arr(*r)
    m = L.or_edge_quadrants(pa, pb)
    return [m >> k & 1 for k in range(4)]


def test_edge_quadrants_hand_derived():
    """Masks of Edge::collides_with_quadrants (qt_traits.rs:19-106) worked out by hand on
    r = [0,4]^2: quadrants q0 [2,4]x[2,4], q1 [0,2]x[2,4], q2 [0,2]^2, q3 [2,4]x[0,2]; sides
    top (4,4)->(0,4), left (0,4)->(0,0), bottom (0,0)->(4,0), right (4,0)->(4,4); bisectors
    h (0,2)->(4,2), v (2,0)->(2,4).  In edge_intersection_half u = 1 at e2's start, so u > 0.5
    is e2's first half."""
    r = (0.0, 0.0, 4.0, 4.0)
    # Vertical edge on the v bisector: all four bbox tests overlap (x = 2 is in every
    # quadrant's x range), no end point inside.  Left, right and v are parallel (no
    # intersection).  Top is crossed at (2, 4): u = 0.5, Halfway, sets q0 and q1; bottom at
    # (2, 0): Halfway, sets q2 and q3.  Without the Halfway arm q3 would stay unset.
    assert quad_mask((2, -1, 2, 5), r) == [1, 1, 1, 1]
    # The same on the h bisector: left crossed at (0, 2) and right at (4, 2), both Halfway.
    assert quad_mask((-1, 2, 5, 2), r) == [1, 1, 1, 1]
    # Vertical at x = 3: q1 and q2 fail the bbox test (3 > 2); top at (3, 4) has u = 0.75
    # (first half: q0), bottom at (3, 0) u = 0.25 (second half: q3).
    assert quad_mask((3, 5, 3, -1), r) == [1, 0, 0, 1]
    # Through the top-left corner only: q0, q2, q3 fail the bbox test; left is crossed at its
    # start (0, 4), u = 1: first half, q1.  The edge touches q1 in one point (rects are closed).
    assert quad_mask((-1, 3, 1, 5), r) == [0, 1, 0, 0]
    # The diagonal through the centre: corners (0,0) and (4,4) set q2 and q0 through left,
    # right, bottom; the h bisector is crossed at its midpoint (u = -12 / -24 = 0.5 exactly),
    # Halfway: q1 and q3 as well (they touch the edge at (2, 2)).
    assert quad_mask((-1, -1, 5, 5), r) == [1, 1, 1, 1]


def test_edge_quadrants_rounding_overwrite():
    """A flag that the bbox test set to Some(false) is overwritten with Some(true) by a later
    half-intersection (qt_traits.rs:110-140: the guard only asks that SOME target is unset).
    The edge from (-1, a) to (5, b), a = 2 - 2^-22, b = 2 - 2^-23, lies below y = 2, so q0
    and q1 fail the bbox test (Some(false)); q2 and q3 are unset.  The right side call
    (targets q3, q0) computes y2 - y4 = b - 4 = -2 - 2^-23, which rounds (tie to even) to -2,
    so u_nom = (x2 - x4)(y2 - y1) - (y2 - y4)(x2 - x1) = 1 * 2^-23 + 12 rounds to 12, and
    u = 12 / 24 = 0.5 exactly: Halfway sets q3 AND q0.  The left call (targets q1, q2) gives
    u = 0.5 - 3 * 2^-25 (second half: q2 only).  So the reference reports q0 although the edge
    does not touch it; an implementation that only filled unset flags would not."""
    a, b = 2 - 2 ** -22, 2 - 2 ** -23
    assert float(f32(a)) == a and float(f32(b)) == b and float(f32(b - 4)) == -2.0
    assert quad_mask((-1.0, a, 5.0, b), (0.0, 0.0, 4.0, 4.0)) == [1, 0, 1, 1]
    # the same along x: the top call (targets q0, q1) stays in the second half for this edge,
    # the bottom call (targets q2, q3) rounds to Halfway and sets q3
    assert quad_mask((a, 5.0, b, -1.0), (0.0, 0.0, 4.0, 4.0)) == [0, 1, 1, 1]


def exact_seg_meets_rect(e, q, inset):
    """Exact (rationals): does the closed segment e meet the closed rectangle q shrunk by
    `inset` on every side (inset < 0 grows it)?  Liang-Barsky clipping."""
    x1, y1, x2, y2 = map(Fr, e)
    lo_x, lo_y, hi_x, hi_y = Fr(q[0]) + inset, Fr(q[1]) + inset, Fr(q[2]) - inset, Fr(q[3]) - inset
    t0, t1 = Fr(0), Fr(1)
    for p, d in ((-(x2 - x1), x1 - lo_x), (x2 - x1, hi_x - x1), (-(y2 - y1), y1 - lo_y), (y2 - y1, hi_y - y1)):
        if p == 0:
            if d < 0:
                return False
            continue
        t = d / p
        if p < 0:
            t0 = max(t0, t)
        else:
            t1 = min(t1, t)
        if t0 > t1:
            return False
    return True


def test_edge_quadrants_equal_exact_masks():
    """Away from rounding, the mask is exactly the set of quadrants the edge touches (not a
    superset): for edges that either pass through a quadrant's interior by a margin or stay a
    margin away from it, for every quadrant, the oracle's mask equals the exact one."""
    rng = random.Random(8)
    r = (0.0, 0.0, 64.0, 64.0)
    qs = [(32, 32, 64, 64), (0, 32, 32, 64), (0, 0, 32, 32), (32, 0, 64, 32)]
    eps = Fr(1, 1 << 10)
    checked = 0
    for _ in range(6000):
        e = [rnd32(rng, -10, 74) for _ in range(4)]
        if rng.random() < 0.3:                        # short edges inside and across quadrants
            e[2], e[3] = float(f32(e[0] + rng.uniform(-12, 12))), float(f32(e[1] + rng.uniform(-12, 12)))
        exact = []
        for q in qs:
            if exact_seg_meets_rect(e, q, eps):
                exact.append(1)
            elif not exact_seg_meets_rect(e, q, -eps):
                exact.append(0)
            else:
                exact = None
                break
        if exact is None:
            continue
        checked += 1
        assert quad_mask(e, r) == exact, (e, exact)
    assert checked > 5500


# ------------------------------------------------------------------------ transformation
def test_transform_matrix_closed_form():
    dt = np.array([[0.0, 3.0, 4.0], [math.pi / 2, 0, 0], [1.234, -5, 6]], dtype=np.float32)
    m = O.compose(dt)
    assert list(m[0]) == [1.0, -0.0, 3.0, 0.0, 1.0, 4.0]
    for row, (r, tx, ty) in zip(m, dt):
        s, c = math.sin(float(r)), math.cos(float(r))
        assert abs(row[0] - c) < 1e-6 and abs(row[1] + s) < 1e-6 and abs(row[3] - s) < 1e-6
        assert row[2] == tx and row[5] == ty and row[0] == row[4] and row[1] == -row[3]


def test_libm_sincosf_agrees_with_sinf_cosf():
    """Rust's f32::sin_cos is sinf + cosf; a host that fuses them into sincosf must agree."""
    libm = C.CDLL("libm.so.6")
    libm.sinf.restype = libm.cosf.restype = C.c_float
    libm.sinf.argtypes = libm.cosf.argtypes = [C.c_float]
    libm.sincosf.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    rng = np.random.default_rng(6)
    s, c = C.c_float(), C.c_float()
    for x in rng.uniform(0, 2 * math.pi, 20000).astype(np.float32):
        libm.sincosf(C.c_float(x), C.byref(s), C.byref(c))
        assert s.value == libm.sinf(C.c_float(x)) and c.value == libm.cosf(C.c_float(x))
