"""One HIP graph (jg_graph_*) holding a batched detection, the batched hazard collection and an
overlap forward + backward pass: every replay equals the eager calls, with eager work and
synchronisations between the replays, and the detection and sets equal the oracle's."""
import numpy as np
import pytest
import torch

import jagua_amd as J
from jagua_amd import synth as S
from oracle import oracle as O

from helpers import engines, items

pytestmark = pytest.mark.gpu


def test_detect_collect_overlap_in_one_graph(cuda):
    sc = S.config2_scene(seed=71, n_items=300)
    cde, keys, oc = engines(sc)
    polys = [S.polygon(710 + i, 8, 90, 8.0, 30.0, 2) for i in range(3)]
    prod, orc = items(polys)
    box = sc.cde_bbox()
    sets, mats_all, first, expect = [], [], 0, []
    for i in range(3):
        mats = J.compose(S.transforms(720 + i, 3000, box, 0))
        sets.append((cde, prod[i], first, len(mats), None))
        expect.append(O.query(oc, orc[i], mats=mats, mode=3, keys_cap=6, threads=8))
        mats_all.append(mats)
        first += len(mats)
    plan = J.QueryPlan(sets)
    m = torch.from_numpy(np.concatenate(mats_all)).to(cuda)
    n = len(m)
    w = S.config4_workload(seed=9, n_items=300, width=600.0, height=240.0)
    ocde, okeys = S.overlap_engine(w)
    ov = J.PoleOverlap(w["poles"], w["off"], np.zeros((0, 2), np.uint32), w["cedges"])
    pose = torch.from_numpy(w["pose"]).to(cuda)
    ov.pairs_from_cde(ocde, okeys, pose, w["rad"])

    side = torch.cuda.Stream()
    st, cnt, hsh = (torch.empty(n, dtype=torch.uint8, device=cuda), torch.empty(n, dtype=torch.int32, device=cuda),
                    torch.empty(n, dtype=torch.int64, device=cuda))
    keys_out = torch.empty((n, 6), dtype=torch.int32, device=cuda)
    q, g = torch.empty(1, dtype=torch.float32, device=cuda), torch.empty((len(pose), 3), dtype=torch.float32, device=cuda)
    q1, g1 = torch.empty_like(q), torch.empty_like(g)

    def calls():
        plan.detect(m, st, stream=side)
        plan.collect(m, capacity=6, count=cnt, hash=hsh, keys=keys_out, stream=side)
        ov.forward(pose, q, stream=side)
        ov.backward(pose, g, stream=side)
        ov.forward_backward(pose, q1, g1, stream=side)

    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        calls()                       # builds every table and buffer before the capture
        side.synchronize()
        eager = [t.clone() for t in (st, cnt, hsh, keys_out, q, g, q1, g1)]
        with J.Graph.capture(side) as graph:
            calls()
    np.testing.assert_array_equal(eager[0].cpu().numpy(), np.concatenate([e["status"] for e in expect]))
    np.testing.assert_array_equal(eager[1].cpu().numpy().astype(np.uint32), np.concatenate([e["count"] for e in expect]))
    np.testing.assert_array_equal(eager[2].cpu().numpy().view(np.uint64), np.concatenate([e["hash"] for e in expect]))
    np.testing.assert_array_equal(eager[3].cpu().numpy().view(np.uint32), np.concatenate([e["keys"] for e in expect]))
    assert torch.equal(eager[6], eager[4]) and torch.equal(eager[7], eager[5])   # one pass = two passes
    for r in range(4):
        for t in (st, cnt, hsh, keys_out, q, g, q1, g1):
            t.fill_(0 if t.dtype == torch.uint8 else -1)
        x = torch.randn(1000, device=cuda).sum()      # eager work between replays
        torch.cuda.synchronize()
        graph.launch(side)
        side.synchronize()
        for t, e in zip((st, cnt, hsh, keys_out, q, g, q1, g1), eager):
            assert torch.equal(t, e), f"replay {r} differs"
        del x
