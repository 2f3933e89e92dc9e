"""Light sampling of the oracle (CPU): the light BVH, the power alias table
and the uniform sampler, and the spherical-rectangle proposal of quad
emitters, pinned by what the mathematics fixes.

* atan2 / asin (series with octant reduction) against numpy's float64.
* Selection: for every sampler and random receiving points and normals,
  the pmf sums to one over the lights, the pmf returned by selection equals
  the pmf evaluated for that light, and the fraction of a fine stratified
  grid of u that selects each light equals its pmf.
* Quad proposal: E[1/pdf] over a stratified grid of samples equals the
  rectangle's solid angle from the Van Oosterom-Strackee formula in float64
  (with and without the cosine warp); the proposal density integrates to one
  over the quad (area quadrature of pdf cos / d^2); E[f / pdf] reproduces a
  cosine-weighted integral computed by quadrature.
"""
import numpy as np
import pytest

import rypath_amd as rp
from tests.scene_util import SceneBuilder


def test_atan2_asin_match_float64(orc):
    rng = np.random.default_rng(3)
    ys = np.concatenate([rng.normal(size=400) * 10 ** rng.uniform(-6, 3, 400), [0.0, 1.0, -1.0, 0.0]])
    xs = np.concatenate([rng.normal(size=400) * 10 ** rng.uniform(-6, 3, 400), [1.0, 0.0, 0.0, -1.0]])
    got = np.array([orc.L.or_atan2(float(y), float(x)) for y, x in zip(ys.astype(np.float32), xs.astype(np.float32))])
    want = np.arctan2(ys.astype(np.float32).astype(np.float64), xs.astype(np.float32).astype(np.float64))
    assert np.max(np.abs(got - want)) < 4.0e-7
    v = np.linspace(-1, 1, 2001).astype(np.float32)
    got = np.array([orc.L.or_asin(float(x)) for x in v])
    assert np.max(np.abs(got - np.arcsin(v.astype(np.float64)))) < 5.0e-7


def many_lights_scene(sampler):
    """Quads of several sizes, orientations and sidedness, loose triangles,
    two spot lights and an environment around a small floor."""
    b = SceneBuilder(8, 8, 3)
    floor = b.object(b.diffuse(b.constant(0.5)))
    b.quad((-4, 0, -4), (4, 0, -4), (4, 0, 4), (-4, 0, 4), floor, facing=(0, 1, 0))
    rng = np.random.default_rng(11)
    for i in range(9):
        c = rng.uniform(-3, 3, 3) + np.array([0, 4, 0])
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        u = np.cross(ax, [0.3, 1, 0.2])
        u /= np.linalg.norm(u)
        v = np.cross(ax, u)
        w, h = rng.uniform(0.1, 1.2, 2)
        e = b.emission(b.linear(rng.uniform(0.5, 5), rng.uniform(0.5, 5)), two_sided=bool(i % 3 == 0))
        o = b.object(b.diffuse(b.constant(0.0)), e)
        p = [c - w * u - h * v, c + w * u - h * v, c + w * u + h * v, c - w * u + h * v]
        b.quad(*p, o, facing=-ax)
    for i in range(3):  # loose triangles
        e = b.emission(b.constant(2.0 + i))
        o = b.object(b.diffuse(b.constant(0.0)), e)
        base = len(b.positions)
        q = rng.uniform(-3, 3, (3, 3)) + np.array([0, 3, 0])
        b.positions.extend(q.tolist())
        b.indices.append([base, base + 1, base + 2])
        b.tri_object.append(o)
    b.spot((0.5, 5, 0.2), (0.1, -1, 0), np.cos(np.radians(20)), np.cos(np.radians(35)), b.constant(30.0))
    b.spot((-2, 3, 1), (0.5, -1, 0.3), np.cos(np.radians(10)), np.cos(np.radians(15)), b.linear(5, 50))
    b.environment(b.constant(0.3))
    b.look_at((0, 8, -8), (0, 0, 0), (0, 1, 0), 40)
    d = b.build()
    d.light_sampler = sampler
    return b, d


@pytest.mark.parametrize("sampler", [rp.LIGHT_SAMPLER_BVH, rp.LIGHT_SAMPLER_POWER, rp.LIGHT_SAMPLER_UNIFORM])
def test_selection_pmf_is_a_distribution_that_selection_follows(orc, sampler):
    b, d = many_lights_scene(sampler)
    sc = orc.scene(d)
    tab = sc.light_table()
    n_lights = len(tab["kind"])
    assert list(tab["kind"]).count(3) == 9 and list(tab["kind"]).count(0) == 3
    if sampler == rp.LIGHT_SAMPLER_BVH:
        assert sc.light_node_count() == 2 * 14 - 1  # 9 quads, 3 triangles, 2 spots
    rng = np.random.default_rng(5)
    grid = (np.arange(20000) + 0.5) / 20000
    for trial in range(12):
        p = rng.uniform(-3.5, 3.5, 3) * np.array([1, 0, 1]) + np.array([0, rng.uniform(0, 2), 0])
        n = np.zeros(3) if trial % 4 == 0 else rng.normal(size=3)
        if trial % 4:
            n /= np.linalg.norm(n)
        pmf = np.array([sc.light_pmf(p, n, i) for i in range(n_lights)], np.float64)
        if sampler != rp.LIGHT_SAMPLER_BVH:
            assert abs(pmf.sum() - 1) < 2e-6, pmf.sum()
        else:
            # the tree returns no light where the importance vanishes: only
            # for lights that cannot reach p (one-sided emitters facing away,
            # points outside a spot cone)
            assert pmf.sum() <= 1 + 2e-6
            for i in np.nonzero(pmf == 0)[0]:
                assert not can_reach(b, d, tab, i, p), i
        counts = np.zeros(n_lights + 1)
        for u in grid:
            i, q = sc.light_select(p, n, float(u))
            if i == rp.NONE:
                counts[n_lights] += 1
                continue
            assert i < n_lights
            assert q == pytest.approx(pmf[i], rel=2e-6)
            counts[i] += 1
        assert np.max(np.abs(counts[:n_lights] / len(grid) - pmf)) < 2.5e-4
        assert abs(counts[n_lights] / len(grid) - (1 - pmf.sum())) < 2.5e-4


def can_reach(b, d, tab, i, p):
    """Whether light i can send radiance to point p (float64 geometry)."""
    kind, index = int(tab["kind"][i]), int(tab["index"][i])
    P = np.asarray(b.positions, np.float64)
    if kind in (0, 3):
        tri = b.indices[index]
        a, bb, c = P[tri[0]], P[tri[1]], P[tri[2]]
        ng = np.cross(bb - a, c - a)
        e = d.emissions[b.objects[b.tri_object[index]].emission]
        corners = [a, bb, c] + ([P[b.indices[index + 1][2]]] if kind == 3 else [])
        return bool(e.two_sided) or any(ng @ (p - q) > 0 for q in corners)
    if kind == 1:
        sp = d.spots[index]
        pos = np.array(sp.position[:], np.float64)
        axis = np.array(sp.direction[:], np.float64)
        w = (p - pos) / np.linalg.norm(p - pos)
        return axis @ w > sp.cos_falloff_end
    return True


def single_quad(two_sided=False, c=(0.2, 1.5, 0.1), w=0.6, h=0.35, tilt=0.3):
    b = SceneBuilder(4, 4, 1)
    floor = b.object(b.diffuse(b.constant(0.5)))
    b.quad((-2, 0, -2), (2, 0, -2), (2, 0, 2), (-2, 0, 2), floor, facing=(0, 1, 0))
    ax = np.array([np.sin(tilt), -np.cos(tilt), 0.1])
    ax /= np.linalg.norm(ax)
    u = np.cross(ax, [0, 0, 1])
    u /= np.linalg.norm(u)
    v = np.cross(ax, u)
    c = np.array(c)
    p = [c - w * u - h * v, c + w * u - h * v, c + w * u + h * v, c - w * u + h * v]
    e = b.emission(b.constant(1.0), two_sided=two_sided)
    b.quad(*p, b.object(b.diffuse(b.constant(0.0)), e), facing=ax)  # emits downward, toward the floor
    b.look_at((0, 3, -3), (0, 0, 0), (0, 1, 0), 40)
    d = b.build()
    d.light_sampler = rp.LIGHT_SAMPLER_UNIFORM
    return b, d


def quad_corners(b):
    # the emitter is the last two triangles (a, b, c), (a, c, d)
    P = np.asarray(b.positions, np.float64)
    t1, t2 = b.indices[-2], b.indices[-1]
    return P[t1[0]], P[t1[1]], P[t2[2]], P[t1[2]]  # p00, p10, p01, p11


def solid_angle_f64(p00, p10, p01, p11, r):
    def tri(a, b, c):
        a, b, c = a - r, b - r, c - r
        la, lb, lc = np.linalg.norm(a), np.linalg.norm(b), np.linalg.norm(c)
        num = abs(np.dot(a, np.cross(b, c)))
        den = la * lb * lc + np.dot(a, b) * lc + np.dot(a, c) * lb + np.dot(b, c) * la
        return 2 * np.arctan2(num, den)
    return tri(p00, p10, p11) + tri(p00, p11, p01)


@pytest.mark.parametrize("with_normal", [False, True])
def test_rectangle_proposal_solid_angle_and_normalisation(orc, with_normal):
    b, d = single_quad()
    sc = orc.scene(d)
    p00, p10, p01, p11 = quad_corners(b)
    ng = np.cross(p10 - p00, p01 - p00)
    ng /= np.linalg.norm(ng)
    for r in (np.array([0.1, 0.0, 0.2]), np.array([-1.2, 0.3, 0.8]), np.array([0.9, 0.05, -1.1])):
        n = np.array([0.0, 1.0, 0.0]) if with_normal else np.zeros(3)
        omega = solid_angle_f64(p00, p10, p01, p11, r)
        # E[1 / pdf] over a stratified grid = the solid angle
        g = (np.arange(96) + 0.5) / 96
        inv = []
        for u0 in g:
            for u1 in g:
                i, q, pdf, _ = sc.light_sample(r, n, 0.5, u0, u1)
                assert i == 0 and pdf > 0
                inv.append(1.0 / pdf)
        assert np.mean(inv) == pytest.approx(omega, rel=2e-3)
        # the density integrates to one over the quad: sum pdf * cos / d^2 dA
        m = 160
        s = (np.arange(m) + 0.5) / m
        ex, ey = p10 - p00, p01 - p00
        dA = np.linalg.norm(ex) * np.linalg.norm(ey) / m / m
        total = 0.0
        for a in s:
            for c in s:
                q = p00 + a * ex + c * ey
                dv = q - r
                d2 = dv @ dv
                cos_l = abs(ng @ dv) / np.sqrt(d2)
                total += sc.area_light_pdf(0, r, n, q, ng) * cos_l / d2 * dA
        assert total == pytest.approx(1.0, rel=2e-3)


def test_rectangle_proposal_integrates_a_cosine_weighted_function(orc):
    """E[f / pdf] with f = max(0, n . w) over the warped proposal equals the
    quadrature of f over the quad's solid angle (the irradiance form factor)."""
    b, d = single_quad(c=(0.4, 0.8, -0.3), w=0.9, h=0.5, tilt=0.8)
    sc = orc.scene(d)
    p00, p10, p01, p11 = quad_corners(b)
    ng = np.cross(p10 - p00, p01 - p00)
    ng /= np.linalg.norm(ng)
    r = np.array([0.5, 0.0, -0.2])  # in front of the one-sided emitter
    n = np.array([0.0, 1.0, 0.0])
    g = (np.arange(128) + 0.5) / 128
    est = []
    for u0 in g:
        for u1 in g:
            i, q, pdf, _ = sc.light_sample(r, n, 0.5, u0, u1)
            assert i == 0 and pdf > 0
            w = (q - r) / np.linalg.norm(q - r)
            est.append(max(0.0, n @ w) / pdf)
    m = 400
    s = (np.arange(m) + 0.5) / m
    ex, ey = p10 - p00, p01 - p00
    dA = np.linalg.norm(ex) * np.linalg.norm(ey) / m / m
    q = p00 + s[:, None, None] * ex + s[None, :, None] * ey
    dv = q - r
    d2 = (dv * dv).sum(-1)
    w = dv / np.sqrt(d2)[..., None]
    want = (np.maximum(0, w @ n) * np.abs(dv @ ng) / np.sqrt(d2) / d2 * dA).sum()
    assert np.mean(est) == pytest.approx(want, rel=1e-3)
