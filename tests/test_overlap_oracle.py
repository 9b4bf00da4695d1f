"""The FP64 overlap oracle (the checker of the GPU overlap kernels) pinned by closed forms
and by central finite differences of itself.  No GPU."""
import math

import numpy as np

from oracle import oracle as O


def test_two_single_pole_items_closed_form():
    poles = np.array([[0, 0, 3.0], [0, 0, 2.0]], np.float32)
    off = np.array([0, 1, 2], np.uint32)
    pose = np.array([[10.0, 10.0, 0.3], [13.0, 14.0, 1.1]])
    q, g = O.overlap64(pose, poles, off, [[0, 1]], np.zeros((0, 4), np.float32))
    d = 5.0
    assert abs(q - (5.0 - d) ** 2) < 1e-12                      # touching: zero penetration
    pose[1, :2] = (12.0, 13.0)
    q, g = O.overlap64(pose, poles, off, [[0, 1]], np.zeros((0, 4), np.float32))
    d = math.hypot(2.0, 3.0)
    assert abs(q - (5.0 - d) ** 2) < 1e-12
    k = -2 * (5.0 - d) / d
    assert np.allclose(g[0, :2], [k * -2.0, k * -3.0]) and np.allclose(g[1, :2], [-k * -2.0, -k * -3.0])
    assert abs(g[0, 2]) < 1e-12 and abs(g[1, 2]) < 1e-12         # poles at the item origin: no torque


def test_container_edge_closed_form():
    """The container term is signed: D is the distance to the container boundary, positive
    inside and negative outside, and a pole is penalised by (r - D)^2 when r > D."""
    poles = np.array([[1.0, 0.0, 2.0]], np.float32)
    off = np.array([0, 1], np.uint32)
    ce = np.array([[0, 0, 10, 0], [10, 0, 10, 10], [10, 10, 0, 10], [0, 10, 0, 0]], np.float32)
    none = np.zeros((0, 2), np.uint32)
    # inside, 1.5 from the wall x = 0: (2 - 1.5)^2, pushed to +x
    q, g = O.overlap64(np.array([[0.5, 5.0, 0.0]]), poles, off, none, ce)
    assert abs(q - 0.25) < 1e-12
    assert abs(g[0, 0] - (-1.0)) < 1e-12 and abs(g[0, 1]) < 1e-12
    assert abs(g[0, 2]) < 1e-12       # dc/dtheta = (-ry, rx) = (0, 1) at theta = 0: no torque
    # the centre 1 past the wall: D = -1, Q = (2 + 1)^2, dQ/dx = -2 (2 - x) = -6 at x = -1
    q, g = O.overlap64(np.array([[-2.0, 5.0, 0.0]]), poles, off, none, ce)
    assert abs(q - 9.0) < 1e-12
    assert abs(g[0, 0] - (-6.0)) < 1e-12 and abs(g[0, 1]) < 1e-12
    # wholly outside (5 past the wall): still penalised, (2 + 5)^2
    q, _ = O.overlap64(np.array([[-6.0, 5.0, 0.0]]), poles, off, none, ce)
    assert abs(q - 49.0) < 1e-12
    # outside past a corner: D is minus the distance to the corner
    q, g = O.overlap64(np.array([[-4.0, -3.0, 0.0]]), poles, off, none, ce)   # centre (-3, -3)
    d = 3 * np.sqrt(2)
    assert abs(q - (2 + d) ** 2) < 1e-9
    k = -2 * (2 + d) * -1 / d
    assert np.allclose(g[0, :2], [k * -3, k * -3])
    # deep inside: no term
    q, _ = O.overlap64(np.array([[4.0, 5.0, 0.0]]), poles, off, none, ce)
    assert q == 0.0


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(0)
    n = 12
    poles = rng.uniform(-4, 4, (n * 6, 3)).astype(np.float32)
    poles[:, 2] = rng.uniform(1.5, 3.5, n * 6)
    off = np.arange(0, n * 6 + 1, 6, dtype=np.uint32)
    pose = np.column_stack([rng.uniform(4, 26, n), rng.uniform(4, 16, n), rng.uniform(0, 6.28, n)])
    pairs = np.array([[i, j] for i in range(n) for j in range(i + 1, n)], np.uint32)
    ce = np.array([[0, 0, 30, 0], [30, 0, 30, 20], [30, 20, 0, 20], [0, 20, 0, 0]], np.float32)
    pose[0, :2] = (-1.0, 8.0)                    # one item with its centre outside the container
    q, g = O.overlap64(pose, poles, off, pairs, ce)
    assert q > 0
    h = 1e-3
    fd = np.zeros_like(g)
    for i in range(n):
        for c in range(3):
            p1, p2 = pose.copy(), pose.copy()
            p1[i, c] += h
            p2[i, c] -= h
            fd[i, c] = (O.overlap64(p1, poles, off, pairs, ce, grad=False)[0]
                        - O.overlap64(p2, poles, off, pairs, ce, grad=False)[0]) / (2 * h)
    err = np.linalg.norm(fd - g) / np.linalg.norm(g)
    assert err < 1e-4, err
