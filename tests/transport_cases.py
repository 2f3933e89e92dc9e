"""Known-answer scenes for the composed hero transport and their expected
values, computed in float64 from closed forms and quadrature that share no
code with the oracle or the device.

Each expectation is the mean raw XYZ of one path, E[sum_k CIE(round l_k)
L(l_k) / pdf_k / 4], which for uniformly sampled hero wavelengths is
int_360^830 CIE(round l) L(l) dl.  The wavelength integral is split at the
nearest-nm boundaries (nm i covers [i-0.5, i+0.5), clipped to the range) and
each piece is integrated by Gauss-Legendre quadrature.

Cases (what each one would expose):
  furnace       closed box, emission Le and reflectance rho on every wall:
                L = Le sum_{k=0..D} rho^k.  A dropped or doubled emission
                MIS term, a wrong light pmf or area pdf, or a missing
                roulette compensation each move the mean.
  env_floor     a diffuse floor under a constant environment: L = rho E
                (both MIS arms, the 1/(4 pi) environment pdf).
  spot_plane    a spot light (and optionally the environment) over a
                diffuse floor at depth 1: rho/pi I smoothstep(cos) cos/d^2,
                per pixel (delta-light NEE, light selection pmf, falloff).
  slab          a purely absorbing slab in front of an emitter: Le exp(-sigma_a d)
                (hero free flight with absorption as termination).
  single_scatter camera inside a homogeneous medium lit by a spot light, depth 1:
                the single-scattering integral with the Henyey-Greenstein phase
                (sign of g, transmittance on both legs, sigma_s weighting).
"""
from pathlib import Path

import numpy as np

import rypath_amd as rp
from tests.scene_util import SceneBuilder

ROOT = Path(__file__).resolve().parents[1]
CIE = np.loadtxt(ROOT / "rypath_amd" / "data" / "cie_xyz_1931_2deg.csv", delimiter=",")[:, 1:]  # (471, 3)
assert CIE.shape == (471, 3)

# Gauss-Legendre nodes on each nm piece
_GL_X, _GL_W = np.polynomial.legendre.leggauss(8)
_LO = np.maximum(np.arange(360, 831) - 0.5, 360.0)
_HI = np.minimum(np.arange(360, 831) + 0.5, 830.0)
LAMBDA_NODES = (_LO[:, None] + (_HI - _LO)[:, None] * (_GL_X[None, :] + 1) / 2)      # (471, 8)
LAMBDA_WEIGHTS = (_HI - _LO)[:, None] * _GL_W[None, :] / 2                           # (471, 8)


def spectrum_value(s, pool, lam):
    """float64 value of an rp_spectrum at wavelengths `lam` (nm)."""
    lam = np.asarray(lam, np.float64)
    if s.kind == rp.SPECTRUM_CONSTANT:
        v = np.full(lam.shape, s.c[0])
    elif s.kind == rp.SPECTRUM_KNOTS8:
        knots = np.array([360, 420, 470, 520, 570, 620, 700, 830], np.float64)
        v = np.interp(lam, knots, np.asarray(pool, np.float64)[s.table:s.table + 8])
    elif s.kind == rp.SPECTRUM_LINEAR:
        t = (lam - 360.0) / 470.0
        v = (1 - t) * s.c[0] + t * s.c[1]
    elif s.kind == rp.SPECTRUM_TABLE:
        i = np.clip(np.floor(lam + 0.5).astype(int) - 360, 0, 470)
        v = np.asarray(pool, np.float64)[s.table + i]
    elif s.kind == rp.SPECTRUM_SIGMOID:
        x = (s.c[0] * lam + s.c[1]) * lam + s.c[2]
        v = 0.5 + x / (2 * np.sqrt(1 + x * x))
    else:
        raise ValueError("unsupported spectrum kind")
    return v * s.scale


def xyz_integral(f):
    """int_360^830 CIE(round l) f(l) dl; f maps an array of wavelengths to values."""
    vals = f(LAMBDA_NODES)                                           # (471, 8)
    per_nm = (vals * LAMBDA_WEIGHTS).sum(axis=1)                     # (471,)
    return (CIE * per_nm[:, None]).sum(axis=0)


def per_nm_xyz_weights():
    """Weight of nm i in the wavelength integral, times CIE: (471, 3)."""
    return CIE * (_HI - _LO)[:, None]


# ------------------------------------------------------------------ cases

class Case:
    def __init__(self, builder, desc, expected, per_pixel=False, label=""):
        self.builder, self.desc, self.expected, self.per_pixel, self.label = builder, desc, expected, per_pixel, label


def furnace(max_depth, rho=(0.35, 0.75), le=(0.6, 1.4), width=8, height=8):
    """Closed cube [-1,1]^3, every wall one-sided emitting inward with a
    linear spectrum `le` and diffuse with linear reflectance `rho`."""
    b = SceneBuilder(width, height, max_depth)
    m = b.diffuse(b.linear(*rho))
    e = b.emission(b.linear(*le))
    b.box((-1, -1, -1), (1, 1, 1), b.object(m, e), inward=True)
    b.look_at((0.1, -0.2, 0.05), (0.3, 0.1, 1.0), (0, 1, 0), 70)
    desc = b.build()
    rs, ls = desc.materials[0].reflectance, desc.emissions[0].radiance

    def radiance(lam):
        r = spectrum_value(rs, b.pool, lam)
        return spectrum_value(ls, b.pool, lam) * sum(r ** k for k in range(max_depth + 1))

    return Case(b, desc, xyz_integral(radiance), label=f"furnace depth {max_depth}")


def furnace_tunnel(max_depth, rho=(0.35, 0.75), le=(0.6, 1.4), width=8, height=8, half=(1.0, 0.5, 6.0)):
    """Closed box |x| <= hx, |y| <= hy, |z| <= hz whose walls are tiled into
    unit-sized inward-emitting quads (one light each), so that far tiles are
    seen at grazing angles; the furnace expectation is shape-independent."""
    b = SceneBuilder(width, height, max_depth)
    m = b.diffuse(b.linear(*rho))
    e = b.emission(b.linear(*le))
    obj = b.object(m, e)
    h = np.asarray(half, np.float64)
    for axis in range(3):
        u, v = (axis + 1) % 3, (axis + 2) % 3
        nu, nv = max(1, int(round(2 * h[u]))), max(1, int(round(2 * h[v])))
        for side in (-1.0, 1.0):
            for i in range(nu):
                for j in range(nv):
                    u0, u1 = -h[u] + 2 * h[u] * i / nu, -h[u] + 2 * h[u] * (i + 1) / nu
                    v0, v1 = -h[v] + 2 * h[v] * j / nv, -h[v] + 2 * h[v] * (j + 1) / nv
                    corners = []
                    for a_, b_ in ((u0, v0), (u1, v0), (u1, v1), (u0, v1)):
                        q = np.zeros(3)
                        q[axis], q[u], q[v] = side * h[axis], a_, b_
                        corners.append(q)
                    facing = np.zeros(3)
                    facing[axis] = -side
                    b.quad(*corners, obj, facing=facing)
    b.look_at((0.2, -0.1, -5.0), (0.0, 0.0, 6.0), (0, 1, 0), 40)
    desc = b.build()
    rs, ls = desc.materials[0].reflectance, desc.emissions[0].radiance

    def radiance(lam):
        r = spectrum_value(rs, b.pool, lam)
        return spectrum_value(ls, b.pool, lam) * sum(r ** k for k in range(max_depth + 1))

    return Case(b, desc, xyz_integral(radiance), label=f"tunnel furnace depth {max_depth}")


def env_floor(max_depth=2, rho=(0.2, 0.8), env=(1.3, 0.5), width=8, height=8):
    """Diffuse floor y = 0 (normal +y) of 200 x 200 under a constant
    environment; the camera looks down and sees only the floor: L = rho E."""
    b = SceneBuilder(width, height, max_depth)
    m = b.diffuse(b.linear(*rho))
    b.quad((-100, 0, -100), (100, 0, -100), (100, 0, 100), (-100, 0, 100), b.object(m), facing=(0, 1, 0))
    b.environment(b.linear(*env))
    b.look_at((0.3, 1.0, -0.2), (0.0, 0.0, 0.1), (0, 0, 1), 50)
    desc = b.build()
    rs, es = desc.materials[0].reflectance, desc.environment
    exp = xyz_integral(lambda lam: spectrum_value(rs, b.pool, lam) * spectrum_value(es, b.pool, lam))
    return Case(b, desc, exp, label="environment-lit floor")


def _smoothstep(cos_t, cos_start, cos_end):
    t = np.clip((cos_t - cos_end) / (cos_start - cos_end), 0.0, 1.0)
    return t * t * (3 - 2 * t)


def spot_plane(with_env=False, width=12, height=12, sub=16):
    """Spot light 1 unit above a 6 x 6 diffuse floor, pointing down, cone
    20-30 degrees; depth 1 (direct light only).  Per-pixel expectation: the
    mean over the pixel footprint of rho/pi * I * smoothstep * cos_s / d^2
    (+ rho E with the environment, which the floor's points see as a full
    hemisphere).  The floor is small so that the environment's share of the
    power light table leaves the spot a usable selection probability."""
    b = SceneBuilder(width, height, 1)
    m = b.diffuse(b.linear(0.3, 0.7))
    b.quad((-3, 0, -3), (3, 0, -3), (3, 0, 3), (-3, 0, 3), b.object(m), facing=(0, 1, 0))
    cs, ce = np.cos(np.radians(20)), np.cos(np.radians(30))
    b.spot((0.1, 1.0, -0.05), (0, -1, 0), cs, ce, b.linear(2.0, 5.0))
    if with_env:
        b.environment(b.constant(0.25))
    b.look_at((0.0, 2.2, 0.0), (0.0, 0.0, 0.0), (0, 0, 1), 50)
    desc = b.build()
    rs, isp = desc.materials[0].reflectance, desc.spots[0].intensity
    o = b.camera[0]
    d = b.camera_rays(sub)                                     # (H, W, S, 3)
    t = -o[1] / d[..., 1]
    p = o + t[..., None] * d                                   # floor points
    lp = np.array([desc.spots[0].position[i] for i in range(3)], np.float64)
    v = lp - p
    dist2 = (v * v).sum(-1)
    w = v / np.sqrt(dist2)[..., None]
    axis = np.array([desc.spots[0].direction[i] for i in range(3)], np.float64)
    fall = _smoothstep((-w * axis).sum(-1), cs, ce)
    geo = fall * w[..., 1] / dist2 / np.pi                     # (H, W, S), cos_s = w . (0,1,0)
    geo = geo.mean(-1)                                         # footprint mean
    spot_xyz = xyz_integral(lambda lam: spectrum_value(rs, b.pool, lam) * spectrum_value(isp, b.pool, lam))
    exp = geo[..., None] * spot_xyz
    if with_env:
        exp = exp + xyz_integral(lambda lam: spectrum_value(rs, b.pool, lam) * 0.25)
    return Case(b, desc, exp, per_pixel=True, label=f"spot over a floor{' + environment' if with_env else ''}")


def slab(width=8, height=8, sub=6):
    """Camera at the origin looking +z through an absorbing slab
    (z in [1, 2], sigma_a in 8 bins, sigma_s = 0, null interfaces) at a
    one-sided emitter z = 3 facing the camera: L = Le exp(-sigma_a / cos)."""
    b = SceneBuilder(width, height, 1)
    sa = [0.1, 0.4, 0.9, 1.6, 0.3, 0.05, 1.2, 0.7]
    med = b.medium(b.binned(sa), b.constant(0.0), 0.0)
    iface = b.interface()
    b.box((-10, -10, 1), (10, 10, 2), b.object(iface, inside=med), inward=False)
    dark = b.diffuse(b.constant(0.0))
    b.quad((-10, -10, 3), (10, -10, 3), (10, 10, 3), (-10, 10, 3), b.object(dark, b.emission(b.linear(1.0, 3.0))),
           facing=(0, 0, -1))
    b.look_at((0, 0, 0), (0.05, -0.03, 1), (0, 1, 0), 40)
    desc = b.build()
    les = desc.emissions[0].radiance
    cosz = b.camera_rays(sub)[..., 2]                           # (H, W, S)
    sa_nm = np.array([sa[i * 8 // 471] for i in range(471)], np.float64)
    # per nm: Le integrated over the nm piece times the footprint mean of exp(-sigma_a / cos)
    trans = np.exp(-sa_nm[None, None, None, :] / cosz[..., None]).mean(axis=2)   # (H, W, 471)
    le_nm = (spectrum_value(les, b.pool, LAMBDA_NODES) * LAMBDA_WEIGHTS).sum(axis=1)  # (471,)
    exp = (trans * le_nm) @ CIE
    return Case(b, desc, exp, per_pixel=True, label="absorbing slab")


def single_scatter(g, width=6, height=6, sub=3, nodes=96):
    """Camera inside a homogeneous medium box [-5,5]^3 (8-bin sigma_a,
    sigma_s, HG g) lit by a spot light, depth 1, nothing outside the box:
    L(l) = int_0^exit sigma_s e^{-sigma_t t} p_HG(cos) I fall / r^2 e^{-sigma_t r} dt,
    with cos the angle between the light's and the scattered propagation
    directions and p_HG = (1 - g^2) / (4 pi (1 + g^2 - 2 g cos)^1.5)."""
    b = SceneBuilder(width, height, 1)
    sa = [0.05, 0.2, 0.1, 0.3, 0.15, 0.08, 0.25, 0.12]
    ss = [0.4, 0.3, 0.6, 0.2, 0.5, 0.35, 0.45, 0.25]
    med = b.medium(b.binned(sa), b.binned(ss), g)
    b.box((-5, -5, -5), (5, 5, 5), b.object(b.interface(), inside=med), inward=False)
    b.camera_medium = med
    cs, ce = np.cos(np.radians(50)), np.cos(np.radians(75))
    lp = np.array([0.3, 1.6, 1.5])
    b.spot(lp, (-0.1, -1, 0.2), cs, ce, b.binned([3.0, 4.0, 2.5, 5.0, 3.5, 2.0, 4.5, 3.0]))
    b.look_at((0, 0, -2), (0, 0.1, 1), (0, 1, 0), 35)
    desc = b.build()
    sp = desc.spots[0]
    axis = np.array([sp.direction[i] for i in range(3)], np.float64)
    inten = np.array([3.0, 4.0, 2.5, 5.0, 3.5, 2.0, 4.5, 3.0])
    sa8, ss8 = np.array(sa), np.array(ss)
    st8 = sa8 + ss8
    o = b.camera[0]
    d = b.camera_rays(sub)                                      # (H, W, S, 3)
    # exit distance from the box along d
    with np.errstate(divide="ignore"):
        tx = np.where(d > 0, (5 - o) / d, np.where(d < 0, (-5 - o) / d, np.inf))
    texit = tx.min(-1)                                          # (H, W, S)
    gx, gw = np.polynomial.legendre.leggauss(nodes)
    segs = 16
    acc = np.zeros(d.shape[:3] + (8,))
    for sgi in range(segs):
        a0, a1 = texit * sgi / segs, texit * (sgi + 1) / segs
        t = a0[..., None] + (a1 - a0)[..., None] * (gx + 1) / 2       # (H, W, S, n)
        wt = (a1 - a0)[..., None] * gw / 2
        p = o + t[..., None] * d[..., None, :]                       # (H, W, S, n, 3)
        v = lp - p
        r2 = (v * v).sum(-1)
        r = np.sqrt(r2)
        wi = v / r[..., None]                                        # toward the light
        fall = _smoothstep((-wi * axis).sum(-1), cs, ce)
        # propagation before scattering: -wi (from the light); after: -d (toward the camera)
        cos_theta = ((-wi) * (-d[..., None, :])).sum(-1)
        ph = (1 - g * g) / (4 * np.pi * (1 + g * g - 2 * g * cos_theta) ** 1.5)
        geo = fall * ph / r2                                         # (H, W, S, n)
        for k in range(8):
            f = ss8[k] * np.exp(-st8[k] * (t + r)) * geo * inten[k]
            acc[..., k] += (f * wt).sum(-1)
    lbin = acc.mean(axis=2)                                           # (H, W, 8)
    wnm = per_nm_xyz_weights()
    bin_of = np.array([i * 8 // 471 for i in range(471)])
    bin_xyz = np.stack([wnm[bin_of == k].sum(0) for k in range(8)])   # (8, 3)
    exp = lbin @ bin_xyz
    return Case(b, desc, exp, per_pixel=True, label=f"single scattering g={g}")


# ------------------------------------------------------------------ checks

def path_xyz(records):
    """Per-path raw XYZ of successful paths (failed paths contribute zero, as in the film)."""
    x = records["xyz"].astype(np.float64)
    x[records["status"] != 0] = 0
    return x


def mean_check(samples, expected, nsig=6.0, label=""):
    """|mean - expected| <= nsig standard errors + 2e-5 max(1, |expected|), per channel
    (the gate of the reference's analytic transport checks)."""
    samples = np.asarray(samples, np.float64)
    n = samples.shape[0]
    mean = samples.mean(0)
    se = samples.std(0, ddof=1) / np.sqrt(n)
    tol = nsig * se + 2e-5 * np.maximum(1.0, np.abs(expected))
    ok = np.all(np.abs(mean - expected) <= tol)
    return ok, f"{label}: mean {mean} expected {np.asarray(expected)} se {se} (|d|/se {np.abs(mean - expected) / np.maximum(se, 1e-30)})"


def pixel_check(records, case, nsig=6.0, rel=2e-3):
    """Per-pixel means against the per-pixel expectation: each pixel within
    nsig standard errors plus `rel` relative (footprint quadrature error)."""
    h, w = case.builder.height, case.builder.width
    x = path_xyz(records)
    pix = records["pixel"].astype(np.int64)
    n = np.bincount(pix, minlength=h * w).astype(np.float64)
    s1 = np.stack([np.bincount(pix, x[:, c], minlength=h * w) for c in range(3)], 1)
    s2 = np.stack([np.bincount(pix, x[:, c] ** 2, minlength=h * w) for c in range(3)], 1)
    mean = s1 / n[:, None]
    var = np.maximum(s2 / n[:, None] - mean ** 2, 0) * n[:, None] / (n[:, None] - 1)
    se = np.sqrt(var / n[:, None])
    exp = case.expected.reshape(h * w, 3)
    tol = nsig * se + rel * np.abs(exp) + 2e-4 * np.abs(exp).max()
    bad = np.abs(mean - exp) > tol
    worst = np.argmax((np.abs(mean - exp) / np.maximum(tol, 1e-30)).max(1))
    return (not bad.any(), f"{case.label}: {int(bad.any(1).sum())} of {h * w} pixels off; worst pixel {worst}: "
                           f"mean {mean[worst]} expected {exp[worst]} se {se[worst]}")


def open_box(width=16, height=16, max_depth=64):
    """Diffuse box open at the top (floor y = 0, four walls up to y = 1.2)
    with a diffuse block inside, lit only by a constant environment and
    seen from above.  No closed form: the expectation is the independent
    float64 BSDF-only estimator (tests/bruteforce.py); multi-bounce paths
    (a fourth hit and beyond) carry a large share of the radiance."""
    b = SceneBuilder(width, height, max_depth)
    floor = b.object(b.diffuse(b.linear(0.7, 0.5)))
    walls = [b.object(b.diffuse(b.linear(*r))) for r in ((0.6, 0.2), (0.15, 0.55), (0.5, 0.5), (0.4, 0.7))]
    block = b.object(b.diffuse(b.linear(0.8, 0.3)))
    b.quad((-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1), floor, facing=(0, 1, 0))
    top = 1.2
    b.quad((-1, 0, -1), (-1, top, -1), (-1, top, 1), (-1, 0, 1), walls[0], facing=(1, 0, 0))
    b.quad((1, 0, -1), (1, top, -1), (1, top, 1), (1, 0, 1), walls[1], facing=(-1, 0, 0))
    b.quad((-1, 0, -1), (1, 0, -1), (1, top, -1), (-1, top, -1), walls[2], facing=(0, 0, 1))
    b.quad((-1, 0, 1), (1, 0, 1), (1, top, 1), (-1, top, 1), walls[3], facing=(0, 0, -1))
    b.box((-0.45, 0.0, -0.3), (0.1, 0.55, 0.25), block, inward=False)
    b.environment(b.linear(1.1, 0.7))
    b.look_at((0.15, 2.4, 0.35), (0.0, 0.2, 0.0), (0, 0, 1), 48)
    desc = b.build()
    return Case(b, desc, None, label="open box under an environment")
