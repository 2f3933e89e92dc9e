"""GPU parity of the scalar (one-wavelength) transport through the C ABI
(rp_sample_scalar, rp_render_frame with RP_FRAME_INDEPENDENT_SAMPLER).

Bars: pixel, sample, random words and wavelength bit-exact with the oracle;
radiance and XYZ within 1e-3 relative on >= 99.9% of paths and pixels; the
reference's golden two-plate paths, the per-path floor MIS closed form and
the sky/floor analytic moments met by the device itself.
"""
import numpy as np
import pytest

from tests import scalar_cases as sc
from tests import transport_cases as tc
from tests.parity_util import AGREE_FRACTION, agreement

pytestmark = pytest.mark.gpu


def _parity(rp, orc, desc, req, label):
    print(f"[scalar] {label}: pixels {req.pixel_begin}+{req.pixel_count} samples {req.sample_begin}+"
          f"{req.sample_count} sampler {req.sampler_kind}", flush=True)
    gpu = rp.Scene(desc).sample_scalar(req)
    cpu = orc.scene(desc).sample_scalar(req)
    for f in ("pixel", "sample", "random_words"):
        assert np.array_equal(gpu[f], cpu[f]), f
    assert np.array_equal(gpu["wavelength"].view(np.uint32), cpu["wavelength"].view(np.uint32))
    assert (gpu["status"] == cpu["status"]).mean() >= AGREE_FRACTION
    assert (gpu["scatterings"] == cpu["scatterings"]).mean() >= AGREE_FRACTION
    frac = agreement(gpu["radiance"][:, None], cpu["radiance"][:, None])
    frac_xyz = agreement(gpu["xyz"], cpu["xyz"])
    exact = float(np.mean(gpu["radiance"].view(np.uint32) == cpu["radiance"].view(np.uint32)))
    print(f"[scalar] {label}: radiance agreement {frac:.6f} xyz {frac_xyz:.6f} bit-exact {exact:.6f}", flush=True)
    assert frac >= AGREE_FRACTION and frac_xyz >= AGREE_FRACTION
    return gpu, cpu


@pytest.mark.parametrize("independent", [True, False])
def test_demo_records(rp, orc, gpu_available, independent):
    syn = rp.Synthetic(rp.SYNTH_DEMO, width=64, height=48)
    gpu, _ = _parity(rp, orc, syn.desc, sc.scalar_request(64 * 48, 4, seed=5, independent=independent), "demo")
    assert (gpu["status"] == 0).all() and int(gpu["scatterings"].max()) >= 4


def test_demo_records_offset_range_and_widths(rp, orc, gpu_available):
    syn = rp.Synthetic(rp.SYNTH_DEMO, width=40, height=30)
    for width in (64, 128, 256):
        req = sc.scalar_request(333, 3, seed=0x1234_5678_9abc, sample_begin=1 << 33, pixel_begin=500)
        req.threadgroup_width = width
        _parity(rp, orc, syn.desc, req, f"demo workgroup {width}")


@pytest.mark.parametrize("name", ["furnace_d3", "spot_plane_env"])
def test_scalar_model_records(rp, orc, gpu_available, name):
    case = {"furnace_d3": lambda: tc.furnace(3), "spot_plane_env": lambda: tc.spot_plane(True)}[name]()
    case.desc.transport = rp.TRANSPORT_SCALAR
    w, h = case.builder.width, case.builder.height
    _parity(rp, orc, case.desc, sc.scalar_request(w * h, 16, seed=3, independent=False), name)


def test_two_plate_golden_paths_on_device(rp, gpu_available):
    _, d = sc.two_plate()
    scene = rp.Scene(d)
    for sample, scatterings, expected in sc.TWO_PLATE_PROBES:
        r = scene.sample_scalar(sc.scalar_request(1, 1, seed=1, sample_begin=sample))
        print(f"[scalar] two-plate sample {sample}: {int(r['scatterings'][0])} scatterings, "
              f"radiance {float(r['radiance'][0]):.8f} expected {expected:.8f}", flush=True)
        assert r["status"][0] == 0 and r["scatterings"][0] == scatterings
        assert abs(float(r["radiance"][0]) - expected) <= 3e-6


def test_floor_closed_form_and_analytic_moments_on_device(rp, orc, gpu_available):
    _, d = sc.floor()
    seed = 0x6a540294bd881bb9
    r = rp.Scene(d).sample_scalar(sc.scalar_request(1, 32, seed=seed))
    assert (r["scatterings"] == 1).all()
    key = [seed & 0xffffffff, seed >> 32]
    for s in r:
        ctr = [int(s["pixel"]), int(s["sample"]) & 0xffffffff, int(s["sample"]) >> 32]
        expected = sc.floor_path_radiance(sc.FLOOR_KNOTS, float(s["wavelength"]), orc.philox(ctr + [2], key),
                                          orc.philox(ctr + [1], key))
        assert abs(float(s["radiance"]) - expected) <= 2e-6
    for case, knots in (("sky", None), ("floor", sc.FLOOR_KNOTS)):
        _, d = sc.sky() if case == "sky" else sc.floor()
        r = rp.Scene(d).sample_scalar(sc.scalar_request(1, 32768, seed=0x9e314fd9c628b435))
        mean, var = sc.analytic_xyz(knots)
        got = r["xyz"].astype(np.float64).mean(0)
        se = np.sqrt(var / 32768)
        print(f"[scalar] {case}: mean {got} analytic {mean} |d|/se {np.abs(got - mean) / se}", flush=True)
        assert np.all(np.abs(got - mean) <= 6 * se + 2e-5 * np.maximum(1.0, np.abs(mean)))


def test_demo_film_vs_oracle(rp, orc, gpu_available):
    syn = rp.Synthetic(rp.SYNTH_DEMO, width=64, height=48)
    cfg = rp.frame_config(8, mode=rp.MODE_MEGAKERNEL, seed=1, randomization=rp.RANDOMIZE_NONE,
                          flags=rp.FRAME_INDEPENDENT_SAMPLER)
    img, st = rp.Scene(syn.desc).render_frame(cfg)
    ref, _ = orc.scene(syn.desc).render_frame(8, seed=1, randomization=rp.RANDOMIZE_NONE,
                                              flags=rp.FRAME_INDEPENDENT_SAMPLER)
    frac = agreement(img.reshape(-1, 3), ref.reshape(-1, 3))
    print(f"[scalar] demo film agreement {frac:.6f}, {st['failed_paths']} failed", flush=True)
    assert frac >= AGREE_FRACTION and st["failed_paths"] == 0


def test_scalar_frame_refusals(rp, gpu_available):
    syn = rp.Synthetic(rp.SYNTH_DEMO, width=16, height=16)
    scene = rp.Scene(syn.desc)
    with pytest.raises(rp.RypathError):
        scene.render_frame(rp.frame_config(2, mode=rp.MODE_WAVEFRONT))
    with pytest.raises(rp.RypathError):
        scene.sample(rp.request(seed=1, pixel_count=4, sample_count=1))
    hero = rp.Scene(rp.Synthetic(rp.SYNTH_CORNELL, seed=1, width=16, height=16).desc)
    with pytest.raises(rp.RypathError):
        hero.render_frame(rp.frame_config(2, mode=rp.MODE_MEGAKERNEL, flags=rp.FRAME_INDEPENDENT_SAMPLER))
    with pytest.raises(rp.RypathError):
        hero.sample_scalar(sc.scalar_request(4, 1))
