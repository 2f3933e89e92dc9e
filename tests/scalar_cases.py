"""Scenes and closed forms of the reference's scalar-transport tests, built in
Python and shared by the oracle tests and the device parity tests.

* two_plate(): two 2 x 2 diffuse plates one unit apart, the camera between
  them (tests/render_transport_tests.hpp:14-28), with the four golden
  Philox paths of tests/render_transport_tests.hpp:47-68.
* floor(): the 16 x 16 floor seen from 2 units above
  (tests/render_tests.cpp:48-56) and the per-path MIS closed form of
  tests/render_tests.cpp:772-783.
* analytic_xyz(): mean and variance of a one-sample path on the sky or the
  unobstructed floor (tests/render_reference.cpp:48-69, 211-236).
* demo(): the demo scene (src/render.cpp:2060-2106) restated independently
  of the library's synthetic builder.
"""
import numpy as np

import rypath_amd as rp
from tests import transport_cases as tc
from tests.scene_util import SceneBuilder

RHO_TWO_PLATE = 0.75
FLOOR_KNOTS = (0.10, 0.13, 0.22, 0.39, 0.53, 0.68, 0.73, 0.66)
KNOT_NM = np.array([360, 420, 470, 520, 570, 620, 700, 830], np.float64)


def _light(c):
    r = 4 * c
    return r / (1 + r * r)


def _escape(c):
    r2 = 16 * c * c
    return r2 / (1 + r2)


# (sample index, scatterings, expected radiance) for seed 1, one pixel
TWO_PLATE_PROBES = [
    (16, 1, RHO_TWO_PLATE * (_light(0.2605856298185942) + _escape(0.39052519476934366))),
    (42, 1, RHO_TWO_PLATE * (_light(0.2866041691639817) + _escape(0.5992369846717766))),
    (7, 3, RHO_TWO_PLATE * _light(0.4436034472185657) + RHO_TWO_PLATE ** 2 * _light(0.7026132031747483)
     + RHO_TWO_PLATE ** 3 * _light(0.34453408825428117)),
    (65, 4, RHO_TWO_PLATE ** 3 * (_light(0.49632220285984685) + _light(0.2762963058330404))
     + RHO_TWO_PLATE ** 2 * _escape(0.34637618413176924)),
]
PLATE = [(-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1)]
PLATE_TRIS = [(0, 1, 2), (0, 2, 3)]


def two_plate():
    b = SceneBuilder(1, 1, 113)
    o = b.object(b.diffuse(b.constant(RHO_TWO_PLATE)))
    b.mesh(PLATE, PLATE_TRIS, o)
    b.mesh([(x, y + 1, z) for x, y, z in PLATE], PLATE_TRIS, o)
    b.environment(b.constant(1.0))
    b.look_at((0, 0.5, 0), (0, 0, 0), (0, 0, 1), 10)
    d = b.build()
    d.transport = rp.TRANSPORT_SCALAR_DEMO
    return b, d


def floor(knots=FLOOR_KNOTS, width=1, height=1):
    b = SceneBuilder(width, height, 113)
    o = b.object(b.diffuse(b.knots8(knots)))
    b.mesh([(-8, 0, -8), (8, 0, -8), (8, 0, 8), (-8, 0, 8)], PLATE_TRIS, o)
    b.environment(b.constant(1.0))
    b.look_at((0, 2, 0), (0, 0, 0), (0, 0, 1), 30)
    d = b.build()
    d.transport = rp.TRANSPORT_SCALAR_DEMO
    return b, d


def sky():
    """Only the environment is visible: the scene's one small triangle sits
    far above the camera, which looks straight down."""
    b = SceneBuilder(1, 1, 113)
    o = b.object(b.diffuse(b.constant(0.5)))
    b.mesh([(-0.1, 50, -0.1), (0.1, 50, -0.1), (0, 50, 0.1)], [(0, 1, 2)], o)
    b.environment(b.constant(1.0))
    b.look_at((0, 2, 0), (0, 0, 0), (0, 0, 1), 30)
    d = b.build()
    d.transport = rp.TRANSPORT_SCALAR_DEMO
    return b, d


def reflectance(knots, nm):
    return np.interp(nm, KNOT_NM, np.asarray(knots, np.float64))


def u01(word):
    """The reference's Philox uniform: the top 24 bits times 2^-24."""
    return (np.asarray(word, np.uint64) >> np.uint64(8)).astype(np.float64) * 2.0 ** -24


def floor_path_radiance(knots, wavelength, light_words, bsdf_words):
    """One floor scattering under the unit sky: rho (w_light + w_bsdf) with the
    uniform-sphere light direction (z = 1 - 2u, phi = 2 pi v; its cosine to
    the floor normal is sqrt(1 - z^2) sin phi) and the cosine-weighted
    continuation (cos^2 = 1 - u)."""
    z = 1 - 2 * u01(light_words[0])
    phi = 2 * np.pi * u01(light_words[1])
    r = 4 * np.maximum(0.0, np.sqrt(1 - z * z) * np.sin(phi))
    r2 = 16 * (1 - u01(bsdf_words[0]))
    return reflectance(knots, wavelength) * (r / (1 + r * r) + r2 / (1 + r2))


def y_normalizer():
    y = tc.CIE[:, 1]
    return float(((y[1:] + y[:-1]) * 0.5).sum())


def floor_directional_second_moment(intervals=65536):
    """E[(w_light + w_bsdf)^2] over independent light (density 1/2 in the
    cosine, zero below the plane) and cosine-weighted directions."""
    c = (np.arange(intervals) + 0.5) / intervals
    r = 4 * c
    nee = r / (1 + r * r)
    bsdf = r * r / (1 + r * r)
    nee_mean = (0.5 * nee).sum() / intervals
    bsdf_mean = (2 * c * bsdf).sum() / intervals
    square = (0.5 * nee * nee + 2 * c * bsdf * bsdf).sum() / intervals
    assert abs(nee_mean + bsdf_mean - 1) < 1e-9  # the Lambertian identity
    return square + 2 * nee_mean * bsdf_mean


def analytic_xyz(knots=None):
    """Mean and variance of one path's XYZ: the sky (knots None) or the
    unobstructed floor of reflectance `knots`, by three-point Gauss-Legendre
    on every 1 nm interval (exact for the product of two linear spectra)."""
    node = np.sqrt(3.0 / 5.0) / 2
    offsets = np.array([0.5 - node, 0.5, 0.5 + node])
    weights = np.array([5.0 / 18.0, 4.0 / 9.0, 5.0 / 18.0])
    scale = 470.0 / y_normalizer()
    angular = floor_directional_second_moment() if knots is not None else 1.0
    rows = np.arange(470)
    mean, second = np.zeros(3), np.zeros(3)
    for off, w in zip(offsets, weights):
        nm = 360.0 + rows + off
        rho = reflectance(knots, nm) if knots is not None else np.ones_like(nm)
        sensor = scale * ((1 - off) * tc.CIE[rows] + off * tc.CIE[rows + 1])
        signal = rho[:, None] * sensor
        mean += w * signal.sum(0) / 470
        second += w * (signal * signal).sum(0) * angular / 470
    return mean, second - mean * mean


# ------------------------------------------------------------ demo scene
DEMO_KNOTS = [
    (0.54, 0.54, 0.55, 0.58, 0.61, 0.63, 0.65, 0.65),
    (0.69, 0.68, 0.67, 0.65, 0.64, 0.62, 0.61, 0.61),
    (0.34, 0.34, 0.36, 0.43, 0.55, 0.61, 0.63, 0.63),
    (0.10, 0.10, 0.16, 0.52, 0.40, 0.09, 0.06, 0.06),
    (0.07, 0.08, 0.09, 0.13, 0.39, 0.73, 0.72, 0.72),
    (0.32, 0.47, 0.72, 0.23, 0.11, 0.08, 0.09, 0.09),
]
DEMO_QUADS = [
    [(-3, 0, -3), (-3, 0, 3), (3, 0, 3), (3, 0, -3)],
    [(-3, 0, -3), (3, 0, -3), (3, 3.2, -3), (-3, 3.2, -3)],
    [(-3, 0, 3), (-3, 0, -3), (-3, 3.2, -3), (-3, 3.2, 3)],
]
CUBE = [(-0.5, 0, -0.5), (0.5, 0, -0.5), (0.5, 0, 0.5), (-0.5, 0, 0.5),
        (-0.5, 1, -0.5), (0.5, 1, -0.5), (0.5, 1, 0.5), (-0.5, 1, 0.5)]
CUBE_TRIS = [(0, 1, 2), (0, 2, 3), (4, 6, 5), (4, 7, 6), (0, 4, 5), (0, 5, 1),
             (1, 5, 6), (1, 6, 2), (2, 6, 7), (2, 7, 3), (3, 7, 4), (3, 4, 0)]
# (offset, scale (width, height, depth), angle in radians)
DEMO_BOXES = [((-1.25, -0.05, -0.6), (1.25, 2.25, 1.25), 0.18),
              ((0.25, -0.05, 0.8), (1.45, 1.08, 1.3), -0.24),
              ((1.4, -0.05, -1.5), (1.07, 1.55, 1.06), 0.44)]


def box_transform(offset, scale, angle):
    """Rows {c w, 0, s d, x}, {0, h, 0, y}, {-s w, 0, c d, z}."""
    c, s = np.cos(angle), np.sin(angle)
    (x, y, z), (w, h, d) = offset, scale
    return np.array([[c * w, 0, s * d, x], [0, h, 0, y], [-s * w, 0, c * d, z]])


def demo(width=256, height=192):
    b = SceneBuilder(width, height, 113)
    objs = [b.object(b.diffuse(b.knots8(k))) for k in DEMO_KNOTS]
    for q, o in zip(DEMO_QUADS, objs[:3]):
        b.mesh(q, PLATE_TRIS, o)
    for (off, scale, ang), o in zip(DEMO_BOXES, objs[3:]):
        m = box_transform(off, scale, ang)
        b.mesh([m[:, :3] @ np.asarray(p, np.float64) + m[:, 3] for p in CUBE], CUBE_TRIS, o)
    b.environment(b.constant(1.0))
    b.look_at((4.1, 3.0, 6.2), (0.0, 1.0, -0.45), (0, 1, 0), 42)
    d = b.build()
    d.transport = rp.TRANSPORT_SCALAR_DEMO
    return b, d


def scalar_request(pixel_count, sample_count, seed=1, sample_begin=0, independent=True, pixel_begin=0):
    return rp.request(seed=seed, pixel_begin=pixel_begin, pixel_count=pixel_count, sample_begin=sample_begin,
                      sample_count=sample_count,
                      sampler_kind=rp.SAMPLER_INDEPENDENT if independent else rp.SAMPLER_SOBOL)
