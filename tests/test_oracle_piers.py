"""Pins the oracle's generate_piers (fail_fast/piers.rs:25-226) by properties that do not come
from its own f32 code:

* every pier lies inside its polygon (end points inside, no crossing with an edge: exact
  integer arithmetic) and, lengthened back by 1/CLIPPING_TRIM about its midpoint, ends on the
  polygon's boundary (piers.rs:128-158);
* every pier runs along one of the 90 ray angles k*pi/90 (piers.rs:100-122: the vertical base
  ray rotated by linspace(0, pi, 91)[..90]);
* every pier minimises the reference's loss (piers.rs:168-196) over all clips of all
  200 x 90 rays, recomputed here in float64 with numpy from the reference's definitions (rays,
  clips, the 100 x 100 lattice of unrepresented points, pole distances, forfeit distance and
  radius of influence), up to a relative 1e-4: f32 sums of ~5000 squared distances cannot
  separate clips closer than that, so near-ties may go either way."""
import numpy as np
import pytest

from jagua_amd import synth as S
from oracle import oracle as O

from test_oracle_query_pins import inside, proper_cross, to_int

RAYS_PER_ANGLE, N_ANGLES, N_POINTS, TRIM, ACTION = 200, 90, 100, 0.999, 0.10


def centroid64(v):
    """SPolygon::centroid (simple_polygon.rs:173-192): area-weighted, in f64."""
    x, y = v[:, 0], v[:, 1]
    xn, yn = np.roll(x, -1), np.roll(y, -1)
    cr = x * yn - xn * y
    a = cr.sum() / 2
    return np.array([((x + xn) * cr).sum() / (6 * a), ((y + yn) * cr).sum() / (6 * a)])


def clips64(v):
    """All clips of all rays (piers.rs:100-158) in f64, and the rays' angle index per clip.
    Rays through a polygon vertex are dropped (their f32 pairing of hits is a rounding matter)."""
    lo, hi = v.min(0), v.max(0)
    w, h = hi - lo
    d = (w - h) / 2 if h < w else 0.0
    e = (h - w) / 2 if w < h else 0.0
    ex_lo, ex_hi = lo - np.array([e, d]), hi + np.array([e, d])          # inflate_to_square
    side = ex_hi[0] - ex_lo[0]
    c = centroid64(v)
    base = np.array([[c[0], c[1] - 2 * (ex_hi[1] - ex_lo[1])], [c[0], c[1] + 2 * (ex_hi[1] - ex_lo[1])]])
    xs = ex_lo[0] + (side / RAYS_PER_ANGLE) * np.arange(RAYS_PER_ANGLE)
    ang = np.linspace(0.0, np.pi, N_ANGLES + 1)[:N_ANGLES]
    A, B = v, np.roll(v, -1, axis=0)
    out, kind = [], []
    for k, t in enumerate(ang):
        cs, sn = np.cos(t), np.sin(t)
        R = np.array([[cs, -sn], [sn, cs]])
        p0 = (base[0][None] + np.stack([xs, np.zeros_like(xs)], 1)) @ R.T      # [rays, 2]
        p1 = (base[1][None] + np.stack([xs, np.zeros_like(xs)], 1)) @ R.T
        r = p1 - p0
        s = B - A
        den = r[:, None, 0] * s[None, :, 1] - r[:, None, 1] * s[None, :, 0]
        q = A[None] - p0[:, None]
        with np.errstate(divide="ignore", invalid="ignore"):
            tr = (q[..., 0] * s[None, :, 1] - q[..., 1] * s[None, :, 0]) / den   # along the ray
            te = (q[..., 0] * r[:, None, 1] - q[..., 1] * r[:, None, 0]) / den   # along the edge
        hit = (den != 0) & (te >= 0) & (te <= 1) & (tr >= 0) & (tr <= 1)
        near_vertex = (den != 0) & (tr >= 0) & (tr <= 1) & ((np.abs(te) < 1e-9) | (np.abs(te - 1) < 1e-9))
        for i in range(len(xs)):
            if near_vertex[i].any():
                continue
            ts = np.sort(tr[i][hit[i]])
            for j in range(0, len(ts) - 1, 2):
                a, b = p0[i] + ts[j] * r[i], p0[i] + ts[j + 1] * r[i]
                mid, dv = (a + b) / 2, (b - a) * (TRIM / 2)
                out.append(np.concatenate([mid - dv, mid + dv]))
                kind.append(k)
    return np.array(out), np.array(kind), (ex_lo, ex_hi), (lo, hi)


def seg_point_dist(E, P):
    """[len(E), len(P)] distances from points to segments E (x1 y1 x2 y2), f64."""
    a, b = E[:, None, :2], E[:, None, 2:]
    ab = b - a
    l2 = np.maximum((ab * ab).sum(-1), 1e-300)
    t = np.clip(((P[None] - a) * ab).sum(-1) / l2, 0, 1)
    d = P[None] - (a + t[..., None] * ab)
    return np.sqrt((d * d).sum(-1))


def inside64(v, P):
    x, y = P[:, 0][:, None], P[:, 1][:, None]
    ax, ay, bx, by = v[:, 0][None], v[:, 1][None], np.roll(v[:, 0], -1)[None], np.roll(v[:, 1], -1)[None]
    cond = (ay > y) != (by > y)
    with np.errstate(divide="ignore", invalid="ignore"):
        xc = ax + (y - ay) * (bx - ax) / (by - ay)
    return ((cond & (x < xc)).sum(1) % 2) == 1


def loss64(E, P, d_rays, d_poles, radius):
    out = np.empty(len(E))
    for c0 in range(0, len(E), 1024):
        d = np.minimum(d_rays[None], seg_point_dist(E[c0:c0 + 1024], P))
        s = np.where(d < radius, np.minimum(d_poles[None], d), d_poles[None])
        out[c0:c0 + 1024] = (s * s).sum(1)
    return out


# polygons whose f64 argmin changes with the radius of influence (0.05, 0.1 or 0.2 of the
# square's side) and with the pole distances, so that a wrong constant or term is caught
@pytest.mark.parametrize("seed,nv,shape,n_piers", [(700, 9, 2, 2), (704, 30, 0, 2), (705, 40, 2, 3)])
def test_piers_properties_and_f64_argmin(seed, nv, shape, n_piers):
    poly = S.polygon(seed, nv, nv, 10.0, 30.0, shape)
    v32 = O.spolygon_new(poly)[0]
    poles, piers = O.surrogate(poly, n_ff_poles=2, n_ff_piers=n_piers)
    assert len(piers) == n_piers
    v = v32.astype(np.float64)
    vi = [(to_int(x), to_int(y)) for x, y in v32]
    diam = np.linalg.norm(v.max(0) - v.min(0))
    # inside and ending on the boundary
    for p in piers:
        a, b = (to_int(p[0]), to_int(p[1])), (to_int(p[2]), to_int(p[3]))
        assert inside(vi, a) and inside(vi, b)
        for j in range(len(vi)):
            assert not proper_cross(a, b, vi[j], vi[(j + 1) % len(vi)])
        mid, half = (p[:2] + p[2:]) / 2.0, (p[2:] - p[:2]) / 2.0 / TRIM
        ends = np.stack([mid - half, mid + half]).astype(np.float64)
        dist = seg_point_dist(np.concatenate([v, np.roll(v, -1, axis=0)], 1), ends).min(0)
        assert (dist < 1e-4 * diam).all(), dist
        # along a ray angle: direction (-sin t, cos t) for t = k pi / 90
        t = np.arctan2(-(p[2] - p[0]), p[3] - p[1]) % np.pi
        k = np.round(t / (np.pi / N_ANGLES))
        assert abs(t - k * np.pi / N_ANGLES) < 1e-4 or abs(t - np.pi) < 1e-4
    # the loss over every clip of every ray, in f64
    E, _, (ex_lo, ex_hi), (lo, hi) = clips64(v)
    assert len(E) > 10000
    gx = np.linspace(ex_lo[0], ex_hi[0], N_POINTS)
    gy = np.linspace(ex_lo[1], ex_hi[1], N_POINTS)
    P = np.stack(np.meshgrid(gx, gy, indexing="ij"), -1).reshape(-1, 2)
    rel = poles[:2].astype(np.float64)
    d2 = ((P[:, None] - rel[None, :, :2]) ** 2).sum(-1)
    P = P[inside64(v, P) & (d2 > rel[None, :, 2] ** 2).all(1)]
    d2 = ((P[:, None] - rel[None, :, :2]) ** 2).sum(-1)
    forfeit = (hi[0] - lo[0]) * (hi[1] - lo[1])               # sqrt(w^2 h^2)
    d_poles = np.minimum(forfeit, (np.sqrt(d2) - rel[None, :, 2]).min(1))
    radius = ACTION * (ex_hi[0] - ex_lo[0])
    d_rays = np.full(len(P), forfeit)
    for p in piers.astype(np.float64):
        losses = loss64(E, P, d_rays, d_poles, radius)
        mine = loss64(p[None], P, d_rays, d_poles, radius)[0]
        assert mine <= losses.min() * (1 + 1e-4), (mine, losses.min())
        # and it is one of the clips (the f32 clip of the same ray)
        assert np.abs(E - p[None]).max(1).min() < 1e-3 * diam
        d_rays = np.minimum(d_rays, seg_point_dist(p[None], P)[0])
