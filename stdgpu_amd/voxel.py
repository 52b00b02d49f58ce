"""Voxel-hashing-shaped workload: surface points -> block keys -> hash map of block slots.

Scene (seeded): 200 spheres (radius uniform 0.5-5 m) and 50 square plane patches (half extent
uniform 2-8 m) inside a 50 m cube; a camera random-walks through it (steps of 2 m per axis, clipped
to the cube).  Each frame samples `npoints` surface points, area-weighted, on the primitives whose
centre lies within the 16 m sensor range of the camera, quantises them at 8 mm voxels into 8^3-voxel blocks (64 mm)
and packs the signed block coordinates into 63 bits (21 bits per axis, offset 2^20).

Frame pipeline (GPU, bulk calls, no host synchronisation): insert the frame's block keys (duplicates
included, per-element flags) -> the keys that were new receive a slot from the front of the free-list
deque<uint32_t> (htab_assign_from_deque) -> find every point's block -> erase the blocks whose centre
is more than 20 m from the camera and push their slots onto the back of the free list in the same
pass over the table (htab_erase_far_blocks).  The reference's caller does the same per point: pop a slot, insert
(key, slot), push the slot back when the key existed.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

VOXEL = 0.008
BLOCK = 8 * VOXEL
OFFSET = 1 << 20
RADIUS = 20.0      # blocks farther than this from the camera are erased
SENSOR = 16.0      # primitives whose centre is within this range are observed in a frame


@dataclass
class Scene:
    centers: np.ndarray   # (250, 3)
    radius: np.ndarray    # spheres: radius; planes: half extent
    is_plane: np.ndarray  # bool
    axes: np.ndarray      # (250, 2, 3) plane axes (unused for spheres)
    area: np.ndarray
    cams: np.ndarray      # (frames, 3)


def make_scene(frames: int = 30, seed: int = 7, nsph: int = 200, npl: int = 50, cube: float = 50.0) -> Scene:
    rng = np.random.default_rng(seed)
    n = nsph + npl
    centers = rng.uniform(0, cube, size=(n, 3))
    radius = np.concatenate([rng.uniform(0.5, 5.0, nsph), rng.uniform(2.0, 8.0, npl)])
    is_plane = np.concatenate([np.zeros(nsph, bool), np.ones(npl, bool)])
    normals = rng.normal(size=(n, 3))
    normals /= np.linalg.norm(normals, axis=1, keepdims=True)
    a = np.cross(normals, np.array([1.0, 0.0, 0.0]))
    bad = np.linalg.norm(a, axis=1) < 1e-6
    a[bad] = np.cross(normals[bad], np.array([0.0, 1.0, 0.0]))
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    b = np.cross(normals, a)
    axes = np.stack([a, b], axis=1)
    area = np.where(is_plane, 4 * radius**2, 4 * np.pi * radius**2)
    cams = np.empty((frames, 3))
    c = np.full(3, cube / 2)
    for f in range(frames):
        cams[f] = c
        c = np.clip(c + rng.normal(0, 2.0, size=3), 0, cube)
    return Scene(centers, radius, is_plane, axes, area, cams)


def pack(bx, by, bz):
    """(x, y, z) signed block coordinates -> 63-bit key (21 bits per axis, offset 2^20)."""
    return ((bx + OFFSET) << 42) | ((by + OFFSET) << 21) | (bz + OFFSET)


def unpack_centers(keys):
    """block keys (int64 array, numpy or torch) -> block centre coordinates in metres (x, y, z)."""
    m = (1 << 21) - 1
    bx = ((keys >> 42) & m) - OFFSET
    by = ((keys >> 21) & m) - OFFSET
    bz = (keys & m) - OFFSET
    return (bx + 0.5) * BLOCK, (by + 0.5) * BLOCK, (bz + 0.5) * BLOCK


def frame_keys_np(scene: Scene, frame: int, npoints: int, seed: int = 11) -> np.ndarray:
    """Block keys of one frame's points (numpy, float64): the parity-test generator."""
    rng = np.random.default_rng(seed * 1000003 + frame)
    cam = scene.cams[frame]
    vis = np.linalg.norm(scene.centers - cam, axis=1) <= SENSOR
    w = np.where(vis, scene.area, 0.0)
    if w.sum() == 0:
        w = scene.area.copy()
    prim = rng.choice(len(w), size=npoints, p=w / w.sum())
    c = scene.centers[prim]
    r = scene.radius[prim]
    d = rng.normal(size=(npoints, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    uv = rng.uniform(-1, 1, size=(npoints, 2)) * r[:, None]
    ax = scene.axes[prim]
    plane_pts = c + uv[:, :1] * ax[:, 0] + uv[:, 1:] * ax[:, 1]
    sph_pts = c + d * r[:, None]
    p = np.where(scene.is_plane[prim][:, None], plane_pts, sph_pts)
    blk = np.floor(np.floor(p / VOXEL) / 8).astype(np.int64)
    return pack(blk[:, 0], blk[:, 1], blk[:, 2])


def frame_keys_device(scene: Scene, frame: int, npoints: int, device, seed: int = 11):
    """Same distribution as frame_keys_np, generated in device memory with torch (float64)."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed * 1000003 + frame)
    cam = scene.cams[frame]
    vis = np.linalg.norm(scene.centers - cam, axis=1) <= SENSOR
    w = np.where(vis, scene.area, 0.0)
    if w.sum() == 0:
        w = scene.area.copy()
    wt = torch.tensor(w / w.sum(), dtype=torch.float64, device=device)
    prim = torch.multinomial(wt, npoints, replacement=True, generator=g)
    C = torch.tensor(scene.centers, dtype=torch.float64, device=device)[prim]
    R = torch.tensor(scene.radius, dtype=torch.float64, device=device)[prim]
    A = torch.tensor(scene.axes, dtype=torch.float64, device=device)[prim]
    plane = torch.tensor(scene.is_plane, device=device)[prim]
    d = torch.randn(npoints, 3, dtype=torch.float64, device=device, generator=g)
    d = d / d.norm(dim=1, keepdim=True)
    uv = (torch.rand(npoints, 2, dtype=torch.float64, device=device, generator=g) * 2 - 1) * R[:, None]
    p = torch.where(plane[:, None], C + uv[:, :1] * A[:, 0] + uv[:, 1:] * A[:, 1], C + d * R[:, None])
    blk = torch.floor(torch.floor(p / VOXEL) / 8).to(torch.int64)
    return pack(blk[:, 0], blk[:, 1], blk[:, 2])


def far_mask(keys, cam):
    """blocks whose centre is more than 20 m from the camera"""
    x, y, z = unpack_centers(keys)
    return (x - cam[0]) ** 2 + (y - cam[1]) ** 2 + (z - cam[2]) ** 2 > RADIUS * RADIUS


def frame_buffers(n: int, device) -> dict:
    """The device arrays one frame of n points writes: reused from frame to frame, so a frame
    allocates and fills nothing before its library calls."""
    import torch

    return {
        "vals": torch.full((n,), -1, dtype=torch.int32, device=device),  # the value a new block holds until its slot
        "ins": torch.empty(n, dtype=torch.uint8, device=device),
        "hits": torch.empty(n, dtype=torch.uint8, device=device),
        "found": torch.empty(n, dtype=torch.int32, device=device),
        "counts": torch.zeros(2, dtype=torch.int64, device=device),  # slots popped, blocks erased
    }


def run_frame(tab, freelist, keys, cam, bufs: dict | None = None, check: bool = True, counts: bool = True):
    """One frame on the GPU, all library calls on the current stream with no host synchronisation:
    bulk insert of the frame's block keys (per-element flags), a slot from the front of the free list
    for every key the insert added (htab_assign_from_deque), find of every point's block, erase of
    the blocks whose centre lies more than 20 m from the camera with their slots pushed onto the back
    of the free list (htab_erase_far_blocks).  keys: int64 device tensor.  Returns the frame's counts
    (`counts`: read back, which waits for the frame; otherwise the device tensor "counts" of
    [new blocks, erased blocks]); `check` adds the pipeline's own checks (also waiting)."""
    n = keys.numel()
    b = bufs if bufs is not None else frame_buffers(n, keys.device)
    ins, c = b["ins"][:n], b["counts"]
    far_before = None
    tab.insert(keys, b["vals"][:n], inserted=ins)
    tab.assign_from_deque(keys, ins, freelist, count=c[0:1])
    hits, _ = tab.find(keys, hits=b["hits"][:n], values_out=b["found"][:n])
    if check:
        ck, _ = tab.contents()
        far_before = int(far_mask(ck, cam).sum())
    tab.erase_far_blocks(cam, RADIUS, BLOCK, freelist, bits=21, offset=OFFSET, count=c[1:2])
    out = {"points": n, "counts": c}
    if counts or check:
        new_blocks, erased = (int(x) for x in c.cpu())
        out.update(new_blocks=new_blocks, erased=erased)
    if check:
        out["popped_all"] = bool((ins <= 1).all())
        out["all_found"] = bool(hits.all())
        out["erased_all"] = out["erased"] == far_before
    return out


def free_list_invariant(tab, freelist, capacity: int) -> bool:
    """the slots held by the table and the slots on the free list are 0..capacity-1, each once"""
    import torch

    _, v = tab.contents()
    f = freelist.contents()
    allv = torch.cat([v, f]).to(torch.int64) & 0xFFFFFFFF
    if allv.numel() != capacity:
        return False
    return bool(torch.equal(torch.sort(allv)[0], torch.arange(capacity, device=allv.device)))


def measure(frames: int = 30, log2_points: int = 24, capacity_log2: int = 23) -> dict:
    """Times the frame pipeline on cuda:0: `frames` frames of 2^log2_points points, a map and a
    free-list deque<uint32_t> of 2^capacity_log2 slots; HIP events around each frame (key generation
    outside).  Returns the figures the benchmark reports."""
    import torch

    import stdgpu_amd as sa

    npts, cap = 1 << log2_points, 1 << capacity_log2
    scene = make_scene(frames=frames)
    # one untimed frame on throwaway objects: the allocator's first allocations of the frame's
    # temporaries are not part of a frame
    wtab, wfl = sa.unordered_map_u64_u32.createDeviceObject(cap), sa.deque_u32(cap)
    wfl.fill_iota()
    bufs = frame_buffers(npts, "cuda")
    run_frame(wtab, wfl, frame_keys_device(scene, 0, npts, "cuda"), scene.cams[0], bufs, check=False)
    torch.cuda.synchronize()
    wfl.destroy()
    sa.unordered_map_u64_u32.destroyDeviceObject(wtab)
    tab = sa.unordered_map_u64_u32.createDeviceObject(cap)
    fl = sa.deque_u32(cap)
    fl.fill_iota()
    stream = torch.cuda.current_stream()
    per_frame_ms, new_blocks, erased = [], [], []
    for f in range(frames):
        keys = frame_keys_device(scene, f, npts, "cuda")
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        out = run_frame(tab, fl, keys, scene.cams[f], bufs, check=False, counts=False)
        b.record(stream)
        b.synchronize()
        per_frame_ms.append(a.elapsed_time(b))
        nb, er = (int(x) for x in out["counts"].cpu())
        new_blocks.append(nb)
        erased.append(er)
    ok = free_list_invariant(tab, fl, cap) and tab.valid()
    ops = frames * 2 * npts + sum(erased)
    res = {
        "workload": f"{frames} frames x 2^{log2_points} points on 200 spheres + 50 planes in a 50 m cube, "
                    "8 mm voxels in 8^3 blocks; insert+slot pop, find, erase > 20 m + slot push",
        "frames": frames,
        "gops": round(ops / (sum(per_frame_ms) / 1e3) / 1e9, 3),
        "frame_ms_mean": round(float(np.mean(per_frame_ms)), 3),
        "frame_ms_p99": round(float(np.percentile(per_frame_ms, 99)), 3),
        "new_blocks_per_frame_mean": int(np.mean(new_blocks)),
        "erased_per_frame_mean": int(np.mean(erased)),
        "blocks_resident_end": tab.size(),
        "free_list_and_table_consistent": bool(ok),
    }
    fl.destroy()
    sa.unordered_map_u64_u32.destroyDeviceObject(tab)
    return res


def main(argv=None) -> None:
    """python -m stdgpu_amd.voxel [frames] [log2 points]: the timed frame pipeline alone (for
    profiling), one JSON line."""
    import json
    import sys

    argv = sys.argv[1:] if argv is None else argv
    frames = int(argv[0]) if argv else 30
    log2_points = int(argv[1]) if len(argv) > 1 else 24
    print(json.dumps(measure(frames=frames, log2_points=log2_points)), flush=True)


if __name__ == "__main__":
    main()
