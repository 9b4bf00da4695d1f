"""GPU pole searches (jg_spolygon_new_batch) against the host construction and the oracle,
bit for bit: POIs and surrogate poles of stars, value-noise rings and large star-shaped
containers, under the default and a fixed-count pole configuration."""
import numpy as np
import pytest

import jagua_amd as J
from jagua_amd import synth as S

from helpers import assert_same_item

pytestmark = pytest.mark.gpu


def _polys():
    ps = [S.polygon(10_000 + i, 12, 24, 2.0, 30.0, 0) for i in range(60)]
    ps += [S.polygon(20_000 + i, 8, 256, 10.0, 30.0, 2) for i in range(60)]
    ps += [S.polygon(30_000 + i, 64, 1024, 6.0, 12.0, 1) for i in range(24)]
    ps += [S.config3_layout(3, l).container for l in (0, 2)]
    return ps


def _same(a, b):
    assert np.array_equal(a.vertices.view(np.uint32), b.vertices.view(np.uint32))
    ia, ib = a.info(), b.info()
    for k in ("area", "diameter"):
        assert np.float32(getattr(ia, k)).view(np.uint32) == np.float32(getattr(ib, k)).view(np.uint32), k
    assert np.array_equal(np.array(ia.bbox[:], np.float32).view(np.uint32), np.array(ib.bbox[:], np.float32).view(np.uint32))
    assert np.array_equal(np.array(ia.poi[:], np.float32).view(np.uint32), np.array(ib.poi[:], np.float32).view(np.uint32))
    assert np.array_equal(a.poles.view(np.uint32), b.poles.view(np.uint32))
    assert np.array_equal(a.piers.view(np.uint32), b.piers.view(np.uint32))


@pytest.mark.parametrize("cfg", [None, J.SPSurrogateConfig(((32, 0.0), (32, 0.0), (32, 0.0)), n_ff_poles=2)],
                         ids=["default", "32poles"])
def test_batch_matches_host(cuda, cfg):
    ps = _polys()
    got, n_host = J.SPolygon.new_batch(ps, cfg)
    assert n_host == 0
    for p, g in zip(ps, got):
        _same(g, J.SPolygon.new(p).generate_surrogate(cfg))


def test_batch_poi_only_and_oracle(cuda):
    ps = _polys()[::7]
    got, _ = J.SPolygon.new_batch(ps, with_surrogate=False)
    for p, g in zip(ps, got):
        h = J.SPolygon.new(p)
        assert np.array_equal(np.array(g.info().poi[:], np.float32), np.array(h.info().poi[:], np.float32))
        assert g.info().n_poles == 0
    got, _ = J.SPolygon.new_batch(ps[:12])
    for p, g in zip(ps[:12], got):
        assert_same_item(g, p)


def test_batch_errors(cuda):
    ps = [S.polygon(1, 12, 12, 5.0, 9.0, 0), np.array([[0, 0], [1, 0], [2, 0]], np.float32)]
    with pytest.raises(J.JaguaError, match="polygon 1"):
        J.SPolygon.new_batch(ps)
    got, n = J.SPolygon.new_batch([])
    assert got == [] and n == 0


def _comb(teeth=20):
    """A comb (teeth [2k, 2k+1] x [5, 20] on a base [0, 2 teeth - 1] x [0, 5]) whose horizontal
    rays cross its boundary 2 * teeth times."""
    pts = [(0.0, 0.0), (2.0 * teeth - 1.0, 0.0)]
    for k in range(teeth - 1, -1, -1):
        if k < teeth - 1:
            pts.append((2.0 * k + 1.0, 5.0))
        pts += [(2.0 * k + 1.0, 20.0), (2.0 * k, 20.0)]
        if k > 0:
            pts.append((2.0 * k, 5.0))
    return np.array(pts, np.float32)


@pytest.mark.parametrize("n_piers,n_ff", [(1, 1), (3, 2)])
def test_batch_piers_match_host(cuda, n_piers, n_ff):
    """generate_piers on the GPU (ray clipping, the point lattice, sequential loss sums, argmin
    with ties to the first clip, distance updates) equals the host routine bit for bit.  Two
    stars have rays crossing their boundary 40 and 48 times; the comb's pole search outgrows
    the device frontier, so that polygon is finished on the host."""
    import concurrent.futures as cf
    ps = [S.polygon(40_000 + i, 8, 48, 5.0, 30.0, i % 3) for i in range(20)] + [_comb()]
    ps += [S.polygon(50_000 + i, 8, 256, 5.0, 30.0, i % 3) for i in (324, 429)]
    cfg = J.SPSurrogateConfig(n_ff_poles=n_ff, n_ff_piers=n_piers)
    got, n_host = J.SPolygon.new_batch(ps, cfg)
    assert n_host == 1
    with cf.ThreadPoolExecutor(8) as ex:
        ref = list(ex.map(lambda p: J.SPolygon.new(p).generate_surrogate(cfg), ps))
    for g, r in zip(got, ref):
        assert g.piers.shape == (n_piers, 4)
        _same(g, r)
    assert_same_item(got[0], ps[0], n_ff_poles=n_ff, n_ff_piers=n_piers)
