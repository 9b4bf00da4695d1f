"""jg_detect_sets_noalloc (the batched detection that allocates nothing: its tables live in a
fixed per-GPU scratch array) against jg_plan_detect, the entry point whose kernels it shares,
through the C ABI: two calls in a row, the second with a smaller input, so that the second
call's tables overwrite the first's in the shared scratch."""
import ctypes as C

import numpy as np
import pytest
import torch

import jagua_amd as J
from jagua_amd import _native as N
from jagua_amd import synth as S

from helpers import engines, items

pytestmark = pytest.mark.gpu


def _workload(n_per_set):
    """Two layouts x three items, a key filter and a hashed filter among the sets."""
    scenes = [S.config1_scene(40), S.config2_scene(seed=41, n_items=200)]
    built = [engines(s) for s in scenes]
    polys = [S.polygon(400 + i, 8, 120, 8.0, 30.0, 2) for i in range(2)] + [S.config1_queries(41)[0][0]]
    prod, _ = items(polys)
    sets, mats, first = [], [], 0
    for li, (cde, keys, _oc) in enumerate(built):
        box = scenes[li].cde_bbox()
        for ii in range(len(polys)):
            n = n_per_set + 37 * ii + 11 * li
            f = [None, J.HazKeyFilter(keys[::5]), J.HashedFilter(9 + li, 0.1)][(li + ii) % 3]
            mats.append(J.compose(S.transforms(500 * li + ii, n, box, 0)))
            sets.append((cde, prod[ii], first, n, f))
            first += n
    return built, prod, sets, np.concatenate(mats)


def _noalloc(sets, m, out, stages=0):
    arr, keep = J._query_sets(sets)
    cfg = N.PlanConfigC(0, 0, 0, 0, stages)
    st = N.lib().jg_detect_sets_noalloc(arr, len(sets), C.byref(cfg), C.c_void_p(m.data_ptr()),
                                        C.c_void_p(out.data_ptr()), None)
    return st


@pytest.mark.parametrize("stages", [2, 1])
def test_noalloc_matches_plan_detect(cuda, stages):
    built, prod, sets, mats = _workload(2500)
    for cde, _, _ in built:
        cde.sync()
    m = torch.from_numpy(mats).to(cuda)
    ref = torch.full((len(mats),), 255, dtype=torch.uint8, device=cuda)
    J.QueryPlan(sets, stages=stages).detect(m, ref)
    out = torch.full((len(mats),), 255, dtype=torch.uint8, device=cuda)
    N.check(_noalloc(sets, m, out, stages))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), ref.cpu().numpy())
    # a second call with a smaller input: fewer sets and queries, other items first
    small = [(c, it, f0 - sets[2][2], n // 3, flt) for (c, it, f0, n, flt) in sets[2:5]]
    m2 = m[sets[2][2]:].contiguous()
    n2 = max(s[2] + s[3] for s in small)
    ref2 = torch.full((n2,), 255, dtype=torch.uint8, device=cuda)
    J.QueryPlan(small, stages=stages).detect(m2, ref2)
    out2 = torch.full((n2,), 255, dtype=torch.uint8, device=cuda)
    N.check(_noalloc(small, m2, out2, stages))
    torch.cuda.synchronize()
    r2 = ref2.cpu().numpy()
    o2 = out2.cpu().numpy()
    covered = np.zeros(n2, dtype=bool)
    for _, _, f0, n, _ in small:
        covered[f0:f0 + n] = True
    np.testing.assert_array_equal(o2[covered], r2[covered])
    assert (o2[~covered] == 255).all(), "outputs outside the sets were written"


def test_noalloc_rejects_what_it_cannot_do_without_allocating(cuda):
    sc = S.config1_scene(42)
    cde, keys, _ = engines(sc)
    cde.sync()
    prod, _ = items([S.config1_queries(42)[0][0]])
    m = torch.from_numpy(J.compose(S.config1_queries(42, n=64)[1])).to(cuda)
    out = torch.zeros(64, dtype=torch.uint8, device=cuda)
    # a table larger than the fixed scratch: rejected on the host, before any launch
    st = _noalloc([(cde, prod[0], 0, 1 << 24, None)], m, out)
    assert st == -8, st                                          # JG_ERR_CAPACITY
    assert "scratch" in N.lib().jg_last_error().decode()
    # an engine changed since its last upload: uploading would allocate, so the call refuses
    cde.deregister_hazard_by_key(keys[0])
    st = _noalloc([(cde, prod[0], 0, 64, None)], m, out)
    assert st == -1, st                                          # JG_ERR_INVALID_ARG
    assert "jg_cde_sync" in N.lib().jg_last_error().decode()
    cde.sync()
    N.check(_noalloc([(cde, prod[0], 0, 64, None)], m, out))
    torch.cuda.synchronize()
