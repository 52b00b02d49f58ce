"""Voxel-hashing workload: per frame, the GPU pipeline (bulk insert, slot pop + assign, find, erase of
far blocks, slot push-back) must hold the same block-key set as the reference caller pattern run on
the oracle (pop a slot, insert, push back when present), and both free lists must satisfy
free-list + assigned slots = {0, ..., capacity-1}."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import stdgpu_amd as sa  # noqa: E402
from stdgpu_amd import voxel as vx  # noqa: E402
from oracle import pyoracle as po  # noqa: E402


def test_voxel_frames_match_reference_caller_pattern():
    # sparse frames: nearly every point is a block of its own, so 6 frames need up to 6 x 2^16 slots
    frames, npts, cap, fl_cap = 6, 1 << 16, 1 << 19, 1 << 19
    scene = vx.make_scene(frames=frames, seed=3)
    tab = sa.unordered_map_u64_u32.createDeviceObject(cap)
    fl = sa.deque_u32(fl_cap)
    fl.fill_iota()
    rtab = po.Table("u64", "u32", cap)
    rq = po.Deque(fl_cap)
    rq.push_back(np.arange(fl_cap, dtype=np.uint32))
    for f in range(frames):
        keys = vx.frame_keys_np(scene, f, npts)
        out = vx.run_frame(tab, fl, torch.from_numpy(keys).cuda(), scene.cams[f])
        assert out["popped_all"] and out["all_found"] and out["erased_all"], out
        # oracle: the reference caller pattern, then find, then erase far blocks + push back slots
        added = po.voxel_insert(rtab, rq, keys.view(np.uint64))
        assert added == out["new_blocks"]
        h, _ = rtab.find(keys.view(np.uint64))
        assert h.all()
        rk, rv = rtab.contents()
        far = vx.far_mask(rk.view(np.int64), scene.cams[f])
        ok = rtab.erase(rk[far])
        assert ok.all()
        rq.push_back(rv[far])
        # same key set, same sizes, both free lists consistent
        gk, _ = tab.contents()
        np.testing.assert_array_equal(np.sort(gk.cpu().numpy().view(np.uint64)), np.sort(rtab.contents()[0]))
        assert tab.size() == rtab.size()
        assert fl.size() == rq.size() == fl_cap - tab.size()
        assert vx.free_list_invariant(tab, fl, fl_cap)
        rk2, rv2 = rtab.contents()
        assert sorted(np.concatenate([rv2, rq.contents()]).tolist()) == list(range(fl_cap))
    assert tab.valid()
    fl.destroy()
