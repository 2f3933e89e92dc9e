"""Known answers for closest-hit ties (oracle, CPU).

Rays along +-x (with +0 and -0 in the other components) start at random
distances from the x = 0 faces of adjacent boxes, on the edges those faces
share (z = 0.5, 1.0, 1.5), so the point they meet lies in two triangles.
The expected answer is brute force: the same rays against a scene that
holds only the triangles containing that point (found with an exact
rational point-in-triangle test), which fits one BVH leaf, so every
triangle is tested and nothing is culled.  The full scene's closest hit
must equal it: same t, and on a tie the smaller scene triangle id.
Without the cull growth of the oracle's traversal this fails (a box whose
rounded entry distance exceeds the tied hit's rounded t is culled).  The
device is held to the oracle on such rays in tests/test_gpu_traversal.py.
"""
from collections import defaultdict
from fractions import Fraction

import numpy as np

from tests.parity_util import make_rays
from tests.scene_util import SceneBuilder
from tests.test_gpu_traversal import edge_scene


def _contains(tri, y, z):
    (y0, z0), (y1, z1), (y2, z2) = tri
    e = [(y1 - y0) * (z - z0) - (z1 - z0) * (y - y0), (y2 - y1) * (z - z1) - (z2 - z1) * (y - y1),
         (y0 - y2) * (z - z2) - (z0 - z2) * (y - y2)]
    return all(v >= 0 for v in e) or all(v <= 0 for v in e)


def _only(rp, points):
    """A scene of the given triangles (in order) and nothing else."""
    b = SceneBuilder(4, 4, 3)
    o = b.object(b.diffuse(b.constant(0.5)))
    for tri in points:
        b.mesh(tri, [(0, 1, 2)], o)
    b.environment(b.constant(1.0))
    b.look_at((2, 1, -6), (2, 1, 1), (0, 1, 0), 40)
    return b.build()


def test_shared_edge_ties_against_brute_force(rp, orc):
    b, desc = edge_scene(rp, (0.0, 0.0, 0.0))
    p = np.asarray(b.positions, np.float64)
    idx = np.asarray(b.indices)
    flat = [t for t in range(len(idx)) if np.all(p[idx[t], 0] == 0.0)]
    tris = {t: [(Fraction(float(p[v, 1])), Fraction(float(p[v, 2]))) for v in idx[t]] for t in flat}
    rng = np.random.default_rng(7)
    n = 4096
    y = rng.uniform(0.01, 1.49, n).astype(np.float32)
    keep = np.abs(y - 0.75) > 1e-3
    y = y[keep]
    n = len(y)
    z = rng.choice([0.5, 1.0, 1.5], n)
    sign = rng.choice([1.0, -1.0], n)
    a = rng.uniform(0.01, 0.9, n).astype(np.float32)
    o = np.stack([-a * sign, y, z], axis=1).astype(np.float32)
    d = np.stack([sign, np.where(rng.random(n) < 0.5, 0.0, -0.0), np.zeros(n)], axis=1).astype(np.float32)
    groups = defaultdict(list)
    for i in range(n):
        hold = tuple(t for t, v in tris.items() if _contains(v, Fraction(float(y[i])), Fraction(float(z[i]))))
        assert len(hold) >= 2  # a point of a shared edge
        groups[hold].append(i)
    expected_tri = np.zeros(n, np.uint32)
    expected_t = np.zeros(n, np.float32)
    for hold, rows in groups.items():
        rows = np.asarray(rows)
        small = _only(rp, [p[idx[t]] for t in hold])
        r = orc.scene(small).trace_closest(make_rays(rp, o[rows], d[rows]))
        assert np.all(r["triangle"] != rp.NONE)
        expected_tri[rows] = np.asarray(hold, np.uint32)[r["triangle"]]
        expected_t[rows] = r["t"]
    hits = orc.scene(desc).trace_closest(make_rays(rp, o, d))
    assert np.array_equal(hits["t"].view(np.uint32), expected_t.view(np.uint32))
    assert np.array_equal(hits["triangle"], expected_tri)
