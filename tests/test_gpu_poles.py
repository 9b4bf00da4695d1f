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
