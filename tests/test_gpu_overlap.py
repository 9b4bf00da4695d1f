"""GPU pole-overlap forward and backward against the FP64 oracle and against central finite
differences (FP64) of the same quantity; deterministic run to run.

Tolerances: Q within 1e-5 relative; gradient within 1e-4 relative of FP64 both norm-wise
(||g - g64|| / ||g64||) and as the max componentwise error over the gradient scale
(max |g - g64| / max |g64|).  Central differences at h = 1e-3 carry an O(h^2) truncation
error of about 1e-3 relative on this workload (it shrinks 100x per 10x smaller h), so they are
checked at 2e-3 there and at 1e-4 with h = 1e-5."""
import numpy as np
import pytest
import torch

import jagua_amd as J
from jagua_amd import synth as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def errors(g, g64):
    glob = np.linalg.norm(g - g64) / np.linalg.norm(g64)
    scaled = np.abs(g - g64).max() / np.abs(g64).max()
    big = np.abs(g64) >= 1e-3 * np.abs(g64).max()
    comp = np.max(np.abs(g - g64)[big] / np.abs(g64)[big])
    return glob, scaled, comp


@pytest.mark.parametrize("n_items", [256, 4096])
def test_overlap_vs_fp64(cuda, n_items):
    w = S.config4_workload(seed=5, n_items=n_items, width=2000.0 * (n_items / 4096) ** 0.5,
                           height=800.0 * (n_items / 4096) ** 0.5)
    ov = J.PoleOverlap(w["poles"], w["off"], w["pairs"], w["cedges"])
    pose = torch.from_numpy(w["pose"]).to(cuda)
    q = ov.forward(pose).item()
    g = ov.backward(pose).cpu().numpy().astype(np.float64)
    q64, g64 = O.overlap64(w["pose"].astype(np.float64), w["poles"], w["off"], w["pairs"], w["cedges"])
    assert q64 > 0
    glob, scaled, comp = errors(g, g64)
    print(f"n_items={n_items} pairs={len(w['pairs'])} Q rel err {abs(q - q64) / q64:.2e} grad rel err: "
          f"L2 {glob:.2e}, max/scale {scaled:.2e}, max componentwise (|g64| >= 1e-3 max) {comp:.2e}")
    assert abs(q - q64) / q64 < 1e-5, (q, q64)
    assert glob < 1e-4 and scaled < 1e-4
    # determinism
    assert ov.forward(pose).item() == q
    assert np.array_equal(ov.backward(pose).cpu().numpy().astype(np.float64), g)
    # the one-pass form equals the two passes bit for bit
    q2, g2 = ov.forward_backward(pose)
    assert q2.item() == q and np.array_equal(g2.cpu().numpy().astype(np.float64), g)


def test_overlap_items_of_many_poles(cuda):
    """Items of 60-150 poles: pairs whose pole sets span several 64-pole chunks on both sides."""
    w = S.config4_workload(seed=9, n_items=96, width=160.0, height=80.0, poles_lo=60, poles_hi=150)
    n = np.diff(w["off"].astype(np.int64))
    assert n.max() > 128 and len(w["pairs"]) > 0
    ov = J.PoleOverlap(w["poles"], w["off"], w["pairs"], w["cedges"])
    pose = torch.from_numpy(w["pose"]).to(cuda)
    q, g = ov.forward_backward(pose)
    q, g = q.item(), g.cpu().numpy().astype(np.float64)
    q64, g64 = O.overlap64(w["pose"].astype(np.float64), w["poles"], w["off"], w["pairs"], w["cedges"])
    glob, scaled, comp = errors(g, g64)
    print(f"poles {n.min()}-{n.max()}, pairs {len(w['pairs'])}: Q rel err {abs(q - q64) / q64:.2e}, grad L2 "
          f"{glob:.2e}, max/scale {scaled:.2e}")
    assert abs(q - q64) / q64 < 1e-5, (q, q64)
    assert glob < 1e-4 and scaled < 1e-4
    assert ov.forward(pose).item() == q
    assert np.array_equal(ov.backward(pose).cpu().numpy().astype(np.float64), g)


def test_overlap_degenerate_shapes(cuda):
    """No pairs, no container edges, items without poles and items of one pole: every form
    returns what the FP64 oracle does (Q = 0 and a zero gradient where nothing overlaps)."""
    rng = np.random.default_rng(11)
    n_items = 12
    counts = np.array([0, 1, 1, 3, 0, 5, 1, 2, 0, 7, 1, 4], dtype=np.uint32)
    off = np.zeros(n_items + 1, dtype=np.uint32)
    off[1:] = np.cumsum(counts)
    poles = np.stack([rng.uniform(-2, 2, off[-1]), rng.uniform(-2, 2, off[-1]), rng.uniform(0.5, 2.0, off[-1])],
                     axis=1).astype(np.float32)
    pose = np.stack([rng.uniform(0, 12, n_items), rng.uniform(0, 6, n_items), rng.uniform(0, 6.28, n_items)],
                    axis=1).astype(np.float32)
    pairs = np.array([(i, j) for i in range(n_items) for j in range(i + 1, n_items)], dtype=np.uint32)
    box = np.array([[0, 0, 12, 0], [12, 0, 12, 6], [12, 6, 0, 6], [0, 6, 0, 0]], dtype=np.float32)
    pt = torch.from_numpy(pose).to(cuda)
    for prs, ce in [(pairs[:0], box[:0]), (pairs[:0], box), (pairs, box[:0]), (pairs, box)]:
        ov = J.PoleOverlap(poles, off, prs, ce)
        q, g = ov.forward_backward(pt)
        q, g = q.item(), g.cpu().numpy().astype(np.float64)
        q64, g64 = O.overlap64(pose.astype(np.float64), poles, off, prs, ce)
        if q64 == 0.0:
            assert q == 0.0 and not g.any(), (len(prs), len(ce), q, np.abs(g).max())
        else:
            assert abs(q - q64) / q64 < 1e-5, (len(prs), len(ce), q, q64)
            assert np.linalg.norm(g - g64) / np.linalg.norm(g64) < 1e-4
        assert ov.forward(pt).item() == q
        assert np.array_equal(ov.backward(pt).cpu().numpy().astype(np.float64), g)
    empty = J.PoleOverlap(poles[:0], np.zeros(1, dtype=np.uint32), pairs[:0], box)
    assert empty.forward(pt[:0]).item() == 0.0                 # an empty pose tensor has no address
    q0, g0 = empty.forward_backward(pt[:0])
    assert q0.item() == 0.0 and g0.numel() == 0


@pytest.mark.parametrize("h,tol", [(1e-3, 2e-3), (1e-5, 1e-4)])
def test_overlap_backward_vs_central_differences(cuda, h, tol):
    w = S.config4_workload(seed=6, n_items=64, width=250.0, height=100.0)
    ov = J.PoleOverlap(w["poles"], w["off"], w["pairs"], w["cedges"])
    pose = w["pose"].astype(np.float64)
    g = ov.backward(torch.from_numpy(w["pose"]).to(cuda)).cpu().numpy().astype(np.float64)
    rng = np.random.default_rng(0)
    worst = 0.0
    for i in rng.choice(len(pose), 16, replace=False):
        for c in range(3):
            p1, p2 = pose.copy(), pose.copy()
            p1[i, c] += h
            p2[i, c] -= h
            fd = (O.overlap64(p1, w["poles"], w["off"], w["pairs"], w["cedges"], grad=False)[0]
                  - O.overlap64(p2, w["poles"], w["off"], w["pairs"], w["cedges"], grad=False)[0]) / (2 * h)
            worst = max(worst, abs(fd - g[i, c]) / max(1.0, abs(fd)))
    print(f"h={h}: worst |fd - g| / max(1, |fd|) = {worst:.2e}")
    assert worst < tol


def test_overlap_pairs_update(cuda):
    w = S.config4_workload(seed=7, n_items=128, width=350.0, height=140.0)
    ov = J.PoleOverlap(w["poles"], w["off"], w["pairs"][:10], w["cedges"])
    ov.set_pairs(w["pairs"])
    pose = torch.from_numpy(w["pose"]).to(cuda)
    q64, _ = O.overlap64(w["pose"].astype(np.float64), w["poles"], w["off"], w["pairs"], w["cedges"], grad=False)
    assert abs(ov.forward(pose).item() - q64) / q64 < 1e-5


def test_overlap_pairs_from_the_quadtree(cuda):
    """Pair candidates found on the GPU in the engine's quadtree (items as placed hazards):
    deterministic, and Q over them equals Q over every pair of overlapping bounding circles
    (the pairs they leave out contribute nothing); forward and backward against FP64.  Some
    items are moved past the container walls (the signed container term)."""
    w = S.config4_workload(seed=8, n_items=512, width=700.0, height=280.0, jitter=4.0)
    w["pose"][[0, 2], 0] = [-8.0, 710.0]          # centres past the left and right walls
    w["pose"][4, :2] = (350.0, -2.0)              # and past the bottom and top walls
    w["pose"][5, :2] = (351.0, 285.0)
    cde, keys = S.overlap_engine(w)
    grid = J.candidate_pairs(w["pose"], w["rad"], [-20.0, -20.0, 720.0, 300.0], 8)
    ov = J.PoleOverlap(w["poles"], w["off"], np.zeros((0, 2), np.uint32), w["cedges"])
    pose = torch.from_numpy(w["pose"]).to(cuda)
    n = ov.pairs_from_cde(cde, keys, pose, w["rad"])
    pairs = ov.pairs()
    assert n == len(pairs) > 0
    assert (pairs[:, 0] < pairs[:, 1]).all()
    gset = set(map(tuple, grid.tolist()))
    assert set(map(tuple, pairs.tolist())) <= gset
    assert ov.pairs_from_cde(cde, keys, pose, w["rad"]) == n and np.array_equal(ov.pairs(), pairs)
    q64, g64 = O.overlap64(w["pose"].astype(np.float64), w["poles"], w["off"], pairs, w["cedges"])
    q64g, g64g = O.overlap64(w["pose"].astype(np.float64), w["poles"], w["off"], grid, w["cedges"])
    assert abs(q64 - q64g) <= 1e-12 * q64g and np.allclose(g64, g64g, rtol=0, atol=1e-9 * np.abs(g64g).max())
    q = ov.forward(pose).item()
    g = ov.backward(pose).cpu().numpy().astype(np.float64)
    glob, scaled, comp = errors(g, g64)
    print(f"quadtree pairs {n} (grid {len(grid)}); Q rel err {abs(q - q64) / q64:.2e}; grad L2 {glob:.2e}, "
          f"max/scale {scaled:.2e}")
    assert abs(q - q64) / q64 < 1e-5
    assert glob < 1e-4 and scaled < 1e-4
    assert np.array_equal(ov.backward(pose).cpu().numpy().astype(np.float64), g)
