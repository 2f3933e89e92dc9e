"""Closest-hit and any-hit queries on rays that stress the BVH box test,
against the oracle's independent BVH, bit for bit.

The device's slab test computes each child's entry and exit distances from
quantised bounds with widened offsets; it must never cull a box that the ray
meets.  These rays sit where a careless test loses hits: directions with
exactly zero (and negative-zero) components, origins exactly on box faces
and rays running inside a face's plane, adjacent boxes that share faces,
slivers, and the same geometry moved 3e4 units from the world origin, where
binary32 spacing is ~2e-3.  Any-hit queries use tmax at and just below the
closest hit's distance.
"""
import numpy as np
import pytest

from tests.parity_util import make_rays
from tests.scene_util import SceneBuilder

pytestmark = pytest.mark.gpu

FAR = (3.0e4, -2.0e4, 1.5e4)


def edge_scene(rp, offset):
    """Adjacent axis-aligned boxes (shared faces), a fine grid of small
    triangles and long slivers, every vertex a multiple of 1/8 plus
    `offset` (float32-exact)."""
    b = SceneBuilder(4, 4, 3)
    o = b.object(b.diffuse(b.constant(0.5)))
    off = np.asarray(offset, np.float64)
    # 4 x 4 x 2 boxes of 1 x 0.75 x 0.5, touching
    for i in range(4):
        for j in range(4):
            for k in range(2):
                lo = off + (i * 1.0, k * 0.75, j * 0.5)
                b.box(lo, lo + (1.0, 0.75, 0.5), o, inward=False)
    # grid of 24 x 24 cells of 0.125 above the boxes, alternating diagonals
    n = 24
    pts = [off + (x * 0.125, 2.0 + 0.125 * ((x * 7 + z * 3) % 5), z * 0.125) for z in range(n + 1)
           for x in range(n + 1)]
    tris = []
    for z in range(n):
        for x in range(n):
            a, c = z * (n + 1) + x, (z + 1) * (n + 1) + x
            if (x + z) % 2:
                tris += [(a, a + 1, c + 1), (a, c + 1, c)]
            else:
                tris += [(a, a + 1, c), (a + 1, c + 1, c)]
    b.mesh(pts, tris, o)
    # slivers: long in x, 1/8 wide in y, at fixed z planes
    for s in range(6):
        z = 2.5 + 0.125 * s
        b.mesh([off + (-1.0, 1.0, z), off + (5.0, 1.0 + 0.125, z), off + (5.0, 1.0, z)], [(0, 1, 2)], o)
    b.environment(b.constant(1.0))
    b.look_at(off + (2.0, 1.0, -6.0), off + (2.0, 1.0, 1.0), (0, 1, 0), 40)
    desc = b.build()
    return b, desc


def edge_rays(offset, rng):
    """Rays of four kinds, 196,608 in all (float32 origins and directions)."""
    off = np.asarray(offset, np.float64)
    lo, hi = off + (-1.5, -0.5, -1.0), off + (5.5, 3.0, 3.5)
    n = 49152
    rays_o, rays_d = [], []
    # 1: axis directions (two exact zeros, signed zeros included) from random
    #    origins and from origins on the boxes' face planes
    axes = np.array([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
                     (-0.0, -0.0, 1), (1, -0.0, 0.0)], np.float64)
    o = rng.uniform(lo, hi, size=(n, 3))
    on_plane = rng.random(n) < 0.5
    axis = rng.integers(0, 3, n)
    steps = (1.0, 0.75, 0.5)
    for a in range(3):
        m = on_plane & (axis == a)
        o[m, a] = off[a] + steps[a] * rng.integers(0, 5, m.sum())
    rays_o.append(o)
    rays_d.append(axes[rng.integers(0, len(axes), n)])
    # 2: one zero component, origins on face planes of that axis (rays
    #    running inside the plane)
    o = rng.uniform(lo, hi, size=(n, 3))
    d = rng.normal(size=(n, 3))
    zero = rng.integers(0, 3, n)
    for a in range(3):
        m = zero == a
        d[m, a] = 0.0
        o[m, a] = off[a] + steps[a] * rng.integers(0, 5, m.sum())
    rays_o.append(o)
    rays_d.append(d / np.linalg.norm(d, axis=1, keepdims=True))
    # 3: random directions from inside and around the geometry
    o = rng.uniform(lo, hi, size=(n, 3))
    d = rng.normal(size=(n, 3))
    rays_o.append(o)
    rays_d.append(d / np.linalg.norm(d, axis=1, keepdims=True))
    # 4: grazing the grid: nearly horizontal directions just above and below it
    o = rng.uniform(lo, hi, size=(n, 3))
    o[:, 1] = off[1] + 2.0 + rng.uniform(-0.05, 0.7, n)
    d = np.stack([rng.normal(size=n), rng.normal(0, 0.02, n), rng.normal(size=n)], axis=1)
    rays_o.append(o)
    rays_d.append(d / np.linalg.norm(d, axis=1, keepdims=True))
    return (np.concatenate(rays_o).astype(np.float32), np.concatenate(rays_d).astype(np.float32))


@pytest.fixture(scope="module", params=[(0.0, 0.0, 0.0), FAR], ids=["origin", "far"])
def edge(request, rp, orc, gpu_available):
    b, desc = edge_scene(rp, request.param)
    return request.param, desc, rp.Scene(desc), orc.scene(desc)


def test_edge_closest_hits(rp, edge):
    offset, desc, scene, oscene = edge
    rng = np.random.default_rng(23)
    o, d = edge_rays(offset, rng)
    assert np.count_nonzero(d == 0) > 100000 and np.count_nonzero(np.signbit(d) & (d == 0)) > 1000
    rays = make_rays(rp, o, d)
    print(f"[edge {offset}] {len(rays)} rays", flush=True)
    hits, _ = scene.trace_closest(rays)
    ref = oscene.trace_closest(rays)
    hit = ref["triangle"] != rp.NONE
    print(f"[edge {offset}] hit fraction {hit.mean():.3f}", flush=True)
    assert 0.2 < hit.mean() < 0.95
    assert np.array_equal(hits["triangle"], ref["triangle"])
    assert np.array_equal(hits["t"].view(np.uint32), ref["t"].view(np.uint32))


def test_edge_any_hits(rp, edge):
    """Occlusion with tmax = inf, at the closest hit and one ulp below it
    agrees with the oracle's closest hit under the same tmax."""
    offset, desc, scene, oscene = edge
    rng = np.random.default_rng(29)
    o, d = edge_rays(offset, rng)
    first = oscene.trace_closest(make_rays(rp, o, d))
    t = first["t"].astype(np.float32)
    hit = first["triangle"] != rp.NONE
    at = np.where(hit, t, np.float32(np.inf)).astype(np.float32)
    below = np.where(hit, np.nextafter(t, np.float32(0)), np.float32(np.inf)).astype(np.float32)
    for tmax in (np.float32(np.inf), at, below):
        rays = make_rays(rp, o, d, tmax)
        occ, _ = scene.trace_any(rays)
        ref = oscene.trace_closest(rays)["triangle"] != rp.NONE
        assert np.array_equal(occ.astype(bool), ref)
