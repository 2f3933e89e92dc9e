"""Known-answer checks of the oracle's scalar (one-wavelength) transport.

The demo model (environment-only lighting, uniform-sphere next-event
estimation, power-heuristic MIS, 3/4 roulette from the third scattering) is
pinned by the reference's own fixtures and closed forms:

* the four golden two-plate Philox paths (tests/render_transport_tests.hpp:57-68),
  exact scattering counts and radiance within 3e-6;
* the per-path MIS closed form on an unobstructed floor
  (tests/render_tests.cpp:772-783), within 2e-6 per path;
* the sky and floor means against the analytic moments
  (tests/render_reference.cpp:48-69, 211-236), six analytic standard errors
  plus 2e-5 max(1, |expected|) (tests/render_tests.cpp:528-548);
* the demo scene against the independent float64 BSDF-only estimator with the
  demo sensor, five combined standard errors over 16 replicates
  (tests/render_tests.cpp:575-612).

The scalar model (scene light sampler, power heuristic) is held to the hero
transport's known answers (furnace, environment-lit floor, spot light), whose
expectations use the same nearest-nm sensor.  Planted slips show the gates
catch an unweighted escape and uncompensated roulette.
"""
import numpy as np
import pytest

import rypath_amd as rp
from tests import bruteforce
from tests import scalar_cases as sc
from tests import transport_cases as tc


def _samples(orc, desc, req):
    return orc.scene(desc).sample_scalar(req)


def test_two_plate_golden_paths(orc):
    _, d = sc.two_plate()
    scene = orc.scene(d)
    for sample, scatterings, expected in sc.TWO_PLATE_PROBES:
        r = scene.sample_scalar(sc.scalar_request(1, 1, seed=1, sample_begin=sample))
        assert r["status"][0] == 0
        assert r["scatterings"][0] == scatterings, (sample, r["scatterings"][0])
        assert abs(float(r["radiance"][0]) - expected) <= 3e-6, (sample, float(r["radiance"][0]), expected)


def test_floor_per_path_mis_closed_form(orc):
    _, d = sc.floor()
    seed = 0x6a540294bd881bb9
    r = _samples(orc, d, sc.scalar_request(1, 32, seed=seed))
    assert (r["status"] == 0).all() and (r["scatterings"] == 1).all() and (r["radiance"] > 0).all()
    key = [seed & 0xffffffff, seed >> 32]
    for s in r:
        light = orc.philox([int(s["pixel"]), int(s["sample"]) & 0xffffffff, int(s["sample"]) >> 32, 2], key)
        bsdf = orc.philox([int(s["pixel"]), int(s["sample"]) & 0xffffffff, int(s["sample"]) >> 32, 1], key)
        expected = sc.floor_path_radiance(sc.FLOOR_KNOTS, float(s["wavelength"]), light, bsdf)
        assert abs(float(s["radiance"]) - expected) <= 2e-6, (int(s["sample"]), float(s["radiance"]), expected)


def test_analytic_moments_identities():
    """The restated analytic moments reproduce the reference's own identities
    (tests/render_tests.cpp:345-358): unit Y under the equal-energy sky, the
    Lambertian furnace halving under reflectance 1/2, and independent MIS arms
    adding variance."""
    sky_mean, sky_var = sc.analytic_xyz(None)
    assert abs(sky_mean[1] - 1) < 2e-14
    floor_mean, floor_var = sc.analytic_xyz(sc.FLOOR_KNOTS)
    assert (sky_var > 0).all() and (floor_var > 0).all() and (floor_mean < 0.75 * sky_mean).all()
    grey_mean, grey_var = sc.analytic_xyz((0.5,) * 8)
    assert np.allclose(grey_mean, sky_mean / 2, rtol=0, atol=2e-14)
    assert (grey_var > sky_var / 4).all()
    assert abs(sc.y_normalizer() - 106.8569149167669) < 1e-10


def _analytic_gate(xyz, knots, paths):
    mean, var = sc.analytic_xyz(knots)
    se = np.sqrt(var / paths)
    got = xyz.mean(0)
    tol = 6 * se + 2e-5 * np.maximum(1.0, np.abs(mean))
    return bool(np.all(np.abs(got - mean) <= tol)), f"mean {got} expected {mean} |d|/se {np.abs(got - mean) / se}"


@pytest.mark.parametrize("case", ["sky", "floor"])
def test_sky_and_floor_match_analytic_moments(orc, case):
    paths = 32768
    _, d = sc.sky() if case == "sky" else sc.floor()
    r = _samples(orc, d, sc.scalar_request(1, paths, seed=0x9e314fd9c628b435))
    assert (r["status"] == 0).all()
    assert int(r["scatterings"].max()) == (0 if case == "sky" else 1)
    ok, msg = _analytic_gate(r["xyz"].astype(np.float64), None if case == "sky" else sc.FLOOR_KNOTS, paths)
    assert ok, msg


def test_demo_synthetic_matches_the_restated_scene():
    """The library's demo scene (RP_SYNTH_DEMO) equals the restatement of
    src/render.cpp:2060-2106 built here."""
    syn = rp.Synthetic(rp.SYNTH_DEMO, width=256, height=192)
    d = syn.desc.contents
    b, ref = sc.demo(256, 192)
    assert d.transport == rp.TRANSPORT_SCALAR_DEMO and d.has_environment == 1
    assert d.triangle_count == len(b.indices) == 42
    pos = np.ctypeslib.as_array(d.positions, shape=(3 * d.vertex_count,)).reshape(-1, 3)
    idx = np.ctypeslib.as_array(d.indices, shape=(3 * d.triangle_count,)).reshape(-1, 3)
    tri_obj = np.ctypeslib.as_array(d.tri_object, shape=(d.triangle_count,))
    want = np.asarray(b.positions, np.float64)[np.asarray(b.indices)]
    assert np.abs(pos[idx] - want).max() < 2e-6
    pool = np.ctypeslib.as_array(d.spectra, shape=(d.spectra_count,))
    for t in range(d.triangle_count):
        m = d.materials[d.objects[tri_obj[t]].material]
        mb = ref.materials[b.objects[b.tri_object[t]].material]
        assert m.kind == rp.MATERIAL_DIFFUSE and m.reflectance.kind == rp.SPECTRUM_KNOTS8
        assert np.array_equal(pool[m.reflectance.table:m.reflectance.table + 8],
                              np.asarray(b.pool, np.float32)[mb.reflectance.table:mb.reflectance.table + 8])
    o, corner, dx, dy = b.camera
    cam = d.camera
    for got, want in ((cam.origin, o), (cam.corner, corner), (cam.dx, dx), (cam.dy, dy)):
        assert np.abs(np.array(list(got)) - want).max() < 1e-6


def _demo_vs_bruteforce(orc, replicates=16, width=4, height=4, spp=128):
    syn = rp.Synthetic(rp.SYNTH_DEMO, width=width, height=height)
    scene = orc.scene(syn.desc)
    b, d = sc.demo(width, height)
    ours, ours_max = [], 0
    brute, brute_max = [], 0
    for k in range(replicates):
        r = scene.sample_scalar(sc.scalar_request(width * height, spp, seed=0x71ac6a05eb690500 + k))
        assert (r["status"] == 0).all()
        ours.append(r["xyz"].astype(np.float64).mean(0))
        ours_max = max(ours_max, int(r["scatterings"].max()))
        x, _, deepest = bruteforce.render(b, d, spp, seed=0x56da2309ab9f4000 + k, sensor="demo")
        brute.append(x.mean(0))
        brute_max = max(brute_max, deepest)
    ours, brute = np.array(ours), np.array(brute)
    se = np.sqrt(ours.var(0, ddof=1) / replicates + brute.var(0, ddof=1) / replicates)
    a, e = ours.mean(0), brute.mean(0)
    ok = bool(np.all(np.abs(a - e) <= 5 * se + 2e-5 * np.maximum(1.0, np.abs(e))))
    return ok, ours_max, brute_max, f"oracle {a} brute force {e} |d|/se {np.abs(a - e) / se}"


def test_demo_scene_matches_the_fp64_bsdf_only_estimator(orc):
    ok, ours_max, brute_max, msg = _demo_vs_bruteforce(orc)
    assert ours_max >= 4 and brute_max >= 4
    assert ok, msg


# scalar model against the hero known answers: (case, spp)
SCALAR_CASES = {
    "furnace_d3": (lambda: tc.furnace(3), 512),
    "furnace_d6": (lambda: tc.furnace(6), 512),
    "furnace_d16_bright": (lambda: tc.furnace(16, rho=(0.8, 0.95)), 512),
    "env_floor": (lambda: tc.env_floor(), 512),
    "spot_plane": (lambda: tc.spot_plane(False), 1024),
    "spot_plane_env": (lambda: tc.spot_plane(True), 1024),
}
_cases = {}


def _scalar_case(name):
    if name not in _cases:
        case = SCALAR_CASES[name][0]()
        case.desc.transport = rp.TRANSPORT_SCALAR
        _cases[name] = case
    return _cases[name]


def _scalar_gate(orc, name, independent=False):
    case = _scalar_case(name)
    spp = SCALAR_CASES[name][1]
    w, h = case.builder.width, case.builder.height
    r = orc.scene(case.desc).sample_scalar(sc.scalar_request(w * h, spp, seed=1, independent=independent))
    assert (r["status"] == 0).all(), sorted(set(r["status"]))
    if case.per_pixel:
        return tc.pixel_check(r, case)
    return tc.mean_check(tc.path_xyz(r), case.expected, label=case.label)


@pytest.mark.parametrize("name", list(SCALAR_CASES))
def test_scalar_model_matches_known_answer(orc, name):
    ok, msg = _scalar_gate(orc, name)
    assert ok, msg


def test_scalar_model_independent_sampler(orc):
    ok, msg = _scalar_gate(orc, "furnace_d3", independent=True)
    assert ok, msg


def test_slip_unweighted_escape_fails_the_floor_gates(orc):
    _, d = sc.floor()
    with orc.planted_slip(1):
        r = _samples(orc, d, sc.scalar_request(1, 32768, seed=0x9e314fd9c628b435))
    ok, msg = _analytic_gate(r["xyz"].astype(np.float64), sc.FLOOR_KNOTS, 32768)
    assert not ok, msg


def test_slip_uncompensated_roulette_fails(orc):
    with orc.planted_slip(3):
        ok_furnace, msg2 = _scalar_gate(orc, "furnace_d16_bright")
    assert not ok_furnace, msg2


def test_scalar_scene_refusals():
    """Scenes the scalar estimators cannot render are refused at creation
    (no device needed: validation precedes the device check)."""
    from tests.scene_util import SceneBuilder

    def refused(desc, words):
        with pytest.raises(rp.RypathError) as e:
            rp.Scene(desc)
        assert e.value.code == rp.INVALID and words in str(e.value), str(e.value)

    slab = tc.slab()
    slab.desc.transport = rp.TRANSPORT_SCALAR
    refused(slab.desc, "media")
    spot = tc.spot_plane(True)
    spot.desc.transport = rp.TRANSPORT_SCALAR_DEMO
    refused(spot.desc, "spot lights")
    b = SceneBuilder(2, 2, 4)
    floor = b.object(b.diffuse(b.constant(0.5)))
    lamp = b.object(b.diffuse(b.constant(0.5)), emission=b.emission(b.constant(1.0)))
    b.quad((-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1), floor, facing=(0, 1, 0))
    b.quad((-0.2, 1, -0.2), (0.2, 1, -0.2), (0.2, 1, 0.2), (-0.2, 1, 0.2), lamp, facing=(0, -1, 0))
    b.environment(b.constant(1.0))
    b.look_at((0, 3, 0.1), (0, 0, 0), (0, 1, 0), 30)
    d = b.build()
    d.transport = rp.TRANSPORT_SCALAR_DEMO
    refused(d, "emitters")
    _, d = sc.floor()
    d.has_environment = 0
    refused(d, "environment")
    _, d = sc.floor()
    d.transport = 7
    refused(d, "unknown transport")
    # the conductors of the inverse-rendering scene are hero-only materials
    inv = rp.Synthetic(rp.SYNTH_INVERSE, seed=1, width=8, height=8)
    d = rp.SceneDesc.from_buffer_copy(inv.d)
    d.transport = rp.TRANSPORT_SCALAR
    refused(d, "diffuse surfaces")
