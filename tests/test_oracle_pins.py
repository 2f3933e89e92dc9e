"""Oracle pins from mathematics: closed forms, normalisation integrals,
sample/evaluate consistency and a double-precision brute-force triangle
test.  Each check is chosen so that a plausible slip (a dropped term, a
wrong sign, a swapped operand) fails it.
"""
import ctypes as C
import math

import numpy as np
import pytest

F = C.c_float


def fa(*v):
    return (F * len(v))(*v)


# ------------------------------------------------------------ transcendentals
def test_sincos_turns_accuracy(orc):
    t = np.linspace(0, 1, 20001, endpoint=False, dtype=np.float32)
    s, c = F(), F()
    err = 0.0
    for x in t[::7]:
        orc.L.or_sincos_turns(float(x), C.byref(s), C.byref(c))
        ang = 2 * math.pi * float(np.float32(x))
        err = max(err, abs(s.value - math.sin(ang)), abs(c.value - math.cos(ang)))
    assert err < 4e-7


def test_exp_log_accuracy(orc):
    xs = np.linspace(-80, 80, 4001, dtype=np.float32)
    rel = max(abs(orc.L.or_exp(float(x)) / math.exp(float(x)) - 1) for x in xs)
    assert rel < 4e-7
    ys = np.exp(np.linspace(-60, 60, 4001)).astype(np.float32)
    ab = max(abs(orc.L.or_log(float(y)) - math.log(float(y))) for y in ys)
    assert ab < 8e-6  # |log| up to 60: relative 1.3e-7
    assert orc.L.or_exp(0.0) == 1.0 and orc.L.or_log(1.0) == 0.0
    assert orc.L.or_exp(-200.0) == 0.0 and math.isinf(orc.L.or_log(0.0))


# ------------------------------------------------------------------- Fresnel
@pytest.mark.parametrize("eta", [1.33, 1.5, 2.4])
def test_fresnel_dielectric_closed_forms(orc, eta):
    fd = orc.L.or_fresnel_dielectric
    assert abs(fd(1.0, eta) - ((eta - 1) / (eta + 1)) ** 2) < 1e-6
    # Brewster angle: parallel reflectance vanishes, F = r_perp^2 / 2
    tb = math.atan(eta)
    ci, ct = math.cos(tb), math.cos(math.asin(math.sin(tb) / eta))
    rperp = (ci - eta * ct) / (ci + eta * ct)
    assert abs(fd(ci, eta) - rperp * rperp / 2) < 2e-6
    # total internal reflection from inside beyond the critical angle
    crit = math.asin(1 / eta)
    assert fd(-math.cos(crit + 0.05), eta) == 1.0
    # reciprocity of the interface at matched angles (Snell pair)
    ti = 0.6
    tt = math.asin(math.sin(ti) / eta)
    assert abs(fd(math.cos(ti), eta) - fd(-math.cos(tt), eta)) < 5e-6


@pytest.mark.parametrize("eta,k", [(0.2, 3.0), (1.1, 2.2), (2.0, 0.5)])
def test_fresnel_conductor_closed_forms(orc, eta, k):
    fc = orc.L.or_fresnel_complex
    f0 = ((eta - 1) ** 2 + k * k) / ((eta + 1) ** 2 + k * k)
    assert abs(fc(1.0, eta, k) - f0) < 2e-6
    assert abs(fc(0.0, eta, k) - 1.0) < 1e-6
    # k = 0 reduces to the dielectric formula
    assert abs(fc(0.7, eta + 1.0, 0.0) - orc.L.or_fresnel_dielectric(0.7, eta + 1.0)) < 2e-6


# ------------------------------------------------------------------- BSDFs
def material(rp, kind, rough=0.0, eta=1.5, k=0.0, refl=0.5, remap=0):
    m = rp.Material()
    m.kind, m.remap_roughness, m.roughness_u, m.roughness_v = kind, remap, rough, rough
    for s, v in ((m.reflectance, refl), (m.eta, eta), (m.k, k)):
        s.kind, s.c[0], s.scale = rp.SPECTRUM_CONSTANT, v, 1.0
    return m


def bsdf_eval(orc, m, wo, wi, lam=550.0):
    f, p = F(), F()
    orc.L.or_bsdf_eval(C.byref(m), None, lam, fa(*wo), fa(*wi), C.byref(f), C.byref(p))
    return f.value, p.value


def bsdf_sample(orc, m, wo, uc, u0, u1, lam=550.0):
    wi, pdf, delta = (F * 3)(), F(), C.c_int()
    ok = orc.L.or_bsdf_sample(C.byref(m), None, lam, fa(*wo), uc, u0, u1, wi, C.byref(pdf), C.byref(delta))
    return ok, np.array(list(wi)), pdf.value, delta.value


def sph(theta, phi):
    return (math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta))


def hemisphere_integral(fn, n_t=160, n_p=160, lower=False):
    """Midpoint quadrature of fn(w) over the upper (or lower) hemisphere."""
    total = 0.0
    for i in range(n_t):
        th = (i + 0.5) * (math.pi / 2) / n_t
        if lower:
            th = math.pi - th
        for j in range(n_p):
            ph = (j + 0.5) * 2 * math.pi / n_p
            total += fn(sph(th, ph)) * math.sin(th)
    return total * (math.pi / 2 / n_t) * (2 * math.pi / n_p)


def test_diffuse_closed_forms(orc, rp):
    m = material(rp, rp.MATERIAL_DIFFUSE, refl=0.6)
    wo = sph(0.4, 1.0)
    f, p = bsdf_eval(orc, m, wo, sph(0.9, 2.0))
    assert abs(f - 0.6 / math.pi) < 1e-7 and abs(p - math.cos(0.9) / math.pi) < 1e-7
    assert bsdf_eval(orc, m, wo, sph(2.0, 2.0)) == (0.0, 0.0)  # opposite hemisphere
    # cosine-hemisphere sampling: E[cos] = 2/3 and pdf = cos/pi
    rng = np.random.default_rng(0)
    cos = []
    for _ in range(4000):
        ok, wi, pdf, _ = bsdf_sample(orc, m, wo, 0.5, *rng.random(2).astype(np.float32))
        assert ok and abs(pdf - wi[2] / math.pi) < 1e-6
        cos.append(wi[2])
    assert abs(np.mean(cos) - 2 / 3) < 0.01


@pytest.mark.parametrize("alpha", [0.05, 0.3, 0.7])
def test_ggx_conductor_normalisation_and_consistency(orc, rp, alpha):
    m = material(rp, rp.MATERIAL_CONDUCTOR, rough=alpha, eta=0.2, k=3.0)
    wo = (0.0, 0.0, 1.0)
    # At normal incidence the reflected direction stays above the horizon iff
    # theta_m < 45 deg; the GGX slope CDF P(tan^2 theta_m <= x) = x / (alpha^2 + x)
    # gives the integral of the sampling pdf over the hemisphere: 1 / (1 + alpha^2).
    n = 400 if alpha < 0.1 else 160
    integ = hemisphere_integral(lambda w: bsdf_eval(orc, m, wo, w)[1], n_t=n, n_p=64)
    assert abs(integ - 1 / (1 + alpha * alpha)) < 0.01
    # albedo with F <= 1 never exceeds one
    alb = hemisphere_integral(lambda w: bsdf_eval(orc, m, sph(0.7, 0.3), w)[0] * w[2], n_t=n, n_p=64)
    assert 0 < alb < 1
    # sampled pdf equals evaluated pdf at the sampled direction
    rng = np.random.default_rng(1)
    for _ in range(300):
        w = sph(rng.uniform(0, 1.4), rng.uniform(0, 6.28))
        ok, wi, pdf, delta = bsdf_sample(orc, m, w, 0.5, *rng.random(2).astype(np.float32))
        if not ok:
            continue
        f, p = bsdf_eval(orc, m, w, tuple(wi))
        assert abs(p - pdf) <= 2e-3 * pdf


@pytest.mark.parametrize("alpha", [0.1, 0.4])
def test_rough_dielectric_energy_and_consistency(orc, rp, alpha):
    eta = 1.5
    m = material(rp, rp.MATERIAL_DIELECTRIC, rough=alpha, eta=eta)
    wo = sph(0.5, 0.2)
    # pdf over the full sphere (reflection + transmission) integrates to <= 1, ~1 for small roughness
    up = hemisphere_integral(lambda w: bsdf_eval(orc, m, wo, w)[1], n_t=200, n_p=96)
    dn = hemisphere_integral(lambda w: bsdf_eval(orc, m, wo, w)[1], n_t=200, n_p=96, lower=True)
    assert 0.9 < up + dn < 1.02
    # Generalized reciprocity of radiance BTDFs (Veach 1997, eq. 5.14):
    # f(wi, wo) / eta_wi_side^2 = f(wo, wi) / eta_wo_side^2; here wo is outside
    # (eta 1) and wi inside (eta 1.5), so f(wi, wo) = eta^2 f(wo, wi).
    wi = sph(2.6, 3.0)
    f1, _ = bsdf_eval(orc, m, wo, wi)
    f2, _ = bsdf_eval(orc, m, wi, wo)
    assert f1 > 0 and abs(f2 - f1 * eta * eta) < 2e-3 * f2 + 1e-6
    rng = np.random.default_rng(2)
    for _ in range(300):
        w = sph(rng.uniform(0, 1.4), rng.uniform(0, 6.28))
        ok, wi, pdf, delta = bsdf_sample(orc, m, w, *rng.random(3).astype(np.float32))
        if not ok:
            continue
        _, p = bsdf_eval(orc, m, w, tuple(wi))
        assert abs(p - pdf) <= 3e-3 * pdf


def test_smooth_dielectric_split(orc, rp):
    m = material(rp, rp.MATERIAL_DIELECTRIC, rough=0.0, eta=1.5)
    wo = sph(0.6, 0.0)
    F_ = orc.L.or_fresnel_dielectric(math.cos(0.6), 1.5)
    ok, wi, pdf, delta = bsdf_sample(orc, m, wo, 0.0, 0.3, 0.3)
    assert ok and delta and abs(pdf - F_) < 1e-6 and wi[2] > 0
    ok, wi, pdf, delta = bsdf_sample(orc, m, wo, 0.999, 0.3, 0.3)
    assert ok and delta and abs(pdf - (1 - F_)) < 1e-6 and wi[2] < 0
    # Snell: sin(theta_t) = sin(theta_i) / eta
    assert abs(math.hypot(wi[0], wi[1]) - math.sin(0.6) / 1.5) < 1e-6


# ------------------------------------------------------- Henyey-Greenstein
@pytest.mark.parametrize("g", [-0.7, 0.0, 0.3, 0.7])
def test_hg_normalisation_and_mean(orc, g):
    # 2 pi * integral over cos of p = 1; mean cosine of dot(wo, wi) = -g (PBRT outward convention)
    xs = np.linspace(-1, 1, 200001)
    p = np.array([orc.L.or_hg(float(x), g) for x in xs[::50]])
    xsub = xs[::50]
    assert abs(2 * math.pi * np.trapezoid(p, xsub) - 1) < 2e-3
    assert abs(2 * math.pi * np.trapezoid(p * xsub, xsub) + g) < 2e-3
    rng = np.random.default_rng(3)
    wi, pdf = (F * 3)(), F()
    wo = fa(0.0, 0.0, 1.0)
    cs = []
    for _ in range(20000):
        orc.L.or_hg_sample(wo, g, float(rng.random()), float(rng.random()), wi, C.byref(pdf))
        cs.append(wi[2])
        if len(cs) < 50:
            assert abs(pdf.value - orc.L.or_hg(wi[2], g)) < 1e-4 * pdf.value + 1e-7
    assert abs(np.mean(cs) + g) < 0.015


# ----------------------------------------------------------- triangle test
def brute_force(o, d, tri):
    """Double-precision Moller-Trumbore; returns t or None."""
    o, d = np.asarray(o, np.float64), np.asarray(d, np.float64)
    p0, p1, p2 = (np.asarray(tri[i], np.float64) for i in range(3))
    e1, e2 = p1 - p0, p2 - p0
    pv = np.cross(d, e2)
    det = np.dot(e1, pv)
    if abs(det) < 1e-300:
        return None
    tv = o - p0
    u = np.dot(tv, pv) / det
    qv = np.cross(tv, e1)
    v = np.dot(d, qv) / det
    t = np.dot(e2, qv) / det
    if u < 0 or v < 0 or u + v > 1 or t <= 0:
        return None
    return t


def test_triangle_test_matches_double_brute_force(orc):
    rng = np.random.default_rng(4)
    out = (F * 4)()
    agree = total = 0
    for _ in range(3000):
        tri = rng.normal(0, 1, (3, 3)).astype(np.float32)
        o = rng.normal(0, 3, 3).astype(np.float32)
        target = (tri.mean(0) + rng.normal(0, 0.7, 3)).astype(np.float32)
        d = (target - o).astype(np.float32)
        hit = orc.L.or_intersect_triangle(fa(*o), fa(*d), float("inf"), fa(*tri.ravel()), out)
        ref = brute_force(o, d, tri)
        total += 1
        if (ref is None) == (not hit):
            agree += 1
            if hit:
                assert abs(out[0] - ref) <= 32 * 2 ** -23 * max(1.0, abs(ref))
                b = np.array([out[1], out[2], out[3]])
                p = b @ tri.astype(np.float64)
                assert np.allclose(p, o + ref * d, atol=1e-4 * max(1, np.abs(p).max()))
    # decisions differ only for rays grazing an edge
    assert agree / total > 0.998


def test_triangle_test_watertight_shared_edge(orc):
    # rays aimed exactly at the shared edge of two triangles hit at least one
    a, b, c, d_ = (0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0)
    t1 = np.array([a, b, c], np.float32).ravel()
    t2 = np.array([b, d_, c], np.float32).ravel()
    out = (F * 4)()
    rng = np.random.default_rng(5)
    for _ in range(2000):
        s = np.float32(rng.random())
        pt = np.array([1 - s, s, 0], np.float32)  # on edge b-c
        o = np.array([rng.normal(), rng.normal(), 2 + rng.random()], np.float32)
        d = (pt - o).astype(np.float32)
        h1 = orc.L.or_intersect_triangle(fa(*o), fa(*d), float("inf"), fa(*t1), out)
        h2 = orc.L.or_intersect_triangle(fa(*o), fa(*d), float("inf"), fa(*t2), out)
        assert h1 or h2
