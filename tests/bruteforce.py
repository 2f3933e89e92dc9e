"""Independent float64 BSDF-only path estimator for diffuse scenes lit by a
constant environment, vectorised over paths with numpy.

It restates the reference's test-only CPU estimator
(tests/render_reference.cpp:142-198 intersect / sample_cosine / trace): a
world-space triangle scan in double precision, cosine-weighted continuation,
no next-event estimation, the environment as the only emitter, and a fixed
3/4 roulette from the third scattering on.  It shares no code, random
numbers or sampling scheme with the oracle or the device: pixel positions,
wavelengths and directions come from numpy's PCG64 stream, one wavelength
per path, uniform on [360, 830] nm, with the sensor weight
470 CIE(round lambda) -- the same nearest-nm sensor the hero estimator
integrates -- or, with sensor="demo", the reference demo model's sensor:
CIE linearly interpolated at lambda times 470 / Y, Y the trapezoid integral
of y-bar (tests/render_reference.cpp:18-33, 266-272).
"""
import numpy as np

from tests import transport_cases as tc


class Mesh:
    def __init__(self, builder, desc):
        p = np.asarray(builder.positions, np.float64)
        idx = np.asarray(builder.indices, np.int64)
        self.p0, self.p1, self.p2 = p[idx[:, 0]], p[idx[:, 1]], p[idx[:, 2]]
        n = np.cross(self.p1 - self.p0, self.p2 - self.p0)
        self.n = n / np.linalg.norm(n, axis=1, keepdims=True)
        mats = [desc.materials[builder.objects[o].material] for o in builder.tri_object]
        self.refl = [m.reflectance for m in mats]
        self.pool = builder.pool

    def intersect(self, o, d):
        """Closest hit per ray: (triangle index or -1, distance)."""
        n_rays = len(o)
        best = np.full(n_rays, np.inf)
        tri = np.full(n_rays, -1, np.int64)
        for i in range(len(self.p0)):
            e1, e2 = self.p1[i] - self.p0[i], self.p2[i] - self.p0[i]
            pv = np.cross(d, e2)
            det = pv @ e1
            with np.errstate(divide="ignore", invalid="ignore"):
                inv = 1.0 / det
                tv = o - self.p0[i]
                u = (tv * pv).sum(1) * inv
                qv = np.cross(tv, e1)
                v = (d * qv).sum(1) * inv
                t = (qv @ e2) * inv
            hit = (det != 0) & (u >= 0) & (v >= 0) & (u + v <= 1) & (t > 1e-9) & (t < best)
            best = np.where(hit, t, best)
            tri = np.where(hit, i, tri)
        return tri, best

    def reflectance(self, tri, lam):
        out = np.empty(len(tri))
        for i in np.unique(tri):
            sel = tri == i
            out[sel] = tc.spectrum_value(self.refl[i], self.pool, lam[sel])
        return out


def _cosine(n, rng):
    m = len(n)
    a = np.where(np.abs(n[:, 2:3]) < 0.9, np.array([[0.0, 0.0, 1.0]]), np.array([[0.0, 1.0, 0.0]]))
    t = np.cross(n, a)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(n, t)
    phi = 2 * np.pi * rng.random(m)
    c = np.sqrt(rng.random(m))
    s = np.sqrt(1 - c * c)
    return t * (s * np.cos(phi))[:, None] + b * (s * np.sin(phi))[:, None] + n * c[:, None]


def demo_sensor(lam):
    """CIE at lambda by linear interpolation between 1 nm rows, times 470 / Y."""
    y_norm = ((tc.CIE[1:, 1] + tc.CIE[:-1, 1]) * 0.5).sum()
    off = np.asarray(lam, np.float64) - 360.0
    row = np.minimum(off.astype(np.int64), 469)
    f = (off - row)[:, None]
    return (470.0 / y_norm) * ((1 - f) * tc.CIE[row] + f * tc.CIE[row + 1])


def render(builder, desc, paths_per_pixel, seed, max_scatterings=1024, sensor="nearest"):
    """Per-path raw XYZ (float64) and pixel index for every pixel of the
    builder's camera, and the largest number of scatterings any path made."""
    rng = np.random.Generator(np.random.PCG64(seed))
    mesh = Mesh(builder, desc)
    w, h = builder.width, builder.height
    o_cam, corner, dx, dy = builder.camera
    pix = np.repeat(np.arange(w * h), paths_per_pixel)
    n = len(pix)
    fx = ((pix % w) + rng.random(n)) / w
    fy = ((pix // w) + rng.random(n)) / h
    d = corner[None, :] + fx[:, None] * dx[None, :] + fy[:, None] * dy[None, :]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = np.repeat(o_cam[None, :], n, axis=0)
    lam = 360.0 + 470.0 * rng.random(n)
    beta = np.ones(n)
    radiance = np.zeros(n)
    env = tc.spectrum_value(desc.environment, builder.pool, lam) if desc.has_environment else np.zeros(n)
    alive = np.arange(n)
    deepest = 0
    for scattered in range(max_scatterings):
        if alive.size == 0:
            break
        tri, dist = mesh.intersect(o[alive], d[alive])
        miss = tri < 0
        radiance[alive[miss]] = beta[alive[miss]] * env[alive[miss]]
        alive, tri, dist = alive[~miss], tri[~miss], dist[~miss]
        if alive.size == 0:
            break
        deepest = scattered + 1
        nrm = mesh.n[tri]
        flip = (nrm * d[alive]).sum(1) > 0
        nrm = np.where(flip[:, None], -nrm, nrm)
        beta[alive] *= mesh.reflectance(tri, lam[alive])
        if scattered >= 2:
            keep = rng.random(alive.size) < 0.75
            beta[alive[keep]] /= 0.75
            alive, nrm, dist = alive[keep], nrm[keep], dist[keep]
        o[alive] = o[alive] + d[alive] * dist[:, None] + nrm * 1e-10
        d[alive] = _cosine(nrm, rng)
    else:
        raise RuntimeError("brute-force estimator exceeded its scattering limit")
    if sensor == "demo":
        return demo_sensor(lam) * radiance[:, None], pix, deepest
    row = np.clip(np.floor(lam + 0.5).astype(np.int64) - 360, 0, 470)
    xyz = 470.0 * tc.CIE[row] * radiance[:, None]
    return xyz, pix, deepest
