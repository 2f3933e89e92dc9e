"""GPU tests of the host-side callers around the path: a PBRT-v4 file driven
through the HIP path (parity with the oracle on the translated scene), the
asynchronous frame API (FrameSession::submit / Frame::wait semantics,
src/frame_session.hpp:145-212) and the `rypath` command line."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tests import exr_util, parity_util

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
CORNELL = ROOT / "rypath_amd" / "scenes" / "cornell.pbrt"
CLI = ROOT / "rypath_amd" / "rypath"

# XYZ -> linear sRGB (D65) rows, src/render.cpp:2282-2296
M_SRGB = np.array([[3.240444291419, -1.537133830545, -0.498529890988],
                   [-0.969266752560, 1.876012242980, 0.041556048487],
                   [0.055643519672, -0.204026238798, 1.057226873771]])


def test_pbrt_cornell_parity_with_oracle(rp, orc, gpu_available):
    p = rp.PbrtScene(CORNELL)
    d = p.d
    d.camera.width = d.camera.height = 48  # smaller film, same scene
    scene = rp.Scene(p.desc)
    req = rp.request(seed=p.info["seed"], pixel_begin=0, pixel_count=48 * 48, sample_count=4)
    gpu = scene.sample(req)
    cpu = orc.scene(p.desc).sample(req)
    print(f"[pbrt] {len(gpu)} paths", flush=True)
    assert np.array_equal(gpu["sobol_index"], cpu["sobol_index"])
    assert np.array_equal(gpu["primary_triangle"], cpu["primary_triangle"])
    frac = parity_util.agreement(gpu["radiance"], cpu["radiance"])
    assert frac >= parity_util.AGREE_FRACTION, frac


def test_frames_match_synchronous_render(rp, gpu_available):
    s = rp.Synthetic(rp.SYNTH_CORNELL, seed=1, width=64, height=48)
    scene = rp.Scene(s.desc)
    cfgs = [rp.frame_config(8, mode=rp.MODE_WAVEFRONT, sample_begin=8 * i, seed=1) for i in range(3)]
    ref = [scene.render_frame(c)[0] for c in cfgs]
    f0 = scene.submit(cfgs[0])
    f1 = scene.submit(cfgs[1])
    with pytest.raises(rp.RypathError) as busy:
        scene.submit(cfgs[2])  # two retained frames: the third submit is refused
    assert busy.value.code == rp.BUSY
    img1, st1 = f1.wait()
    img0, st0 = f0.wait()
    assert f0.poll() and f1.poll()
    assert np.array_equal(img0, ref[0]) and np.array_equal(img1, ref[1])
    assert st0["paths"] == 64 * 48 * 8 and st0["seconds"] > 0
    f0.release()  # frees a slot
    f2 = scene.submit(cfgs[2])
    del scene  # frames outlive the scene
    img2, _ = f2.wait()
    assert np.array_equal(img2, ref[2])
    f1.release()
    f2.release()
    print("[frames] async frames equal synchronous frames; busy on the third retained frame", flush=True)


def test_frame_submit_validates_config(rp, gpu_available):
    s = rp.Synthetic(rp.SYNTH_CORNELL, seed=1, width=16, height=16)
    scene = rp.Scene(s.desc)
    bad = rp.frame_config(4, mode=7)
    with pytest.raises(rp.RypathError) as e:
        scene.submit(bad)
    assert e.value.code == rp.INVALID


def run_cli(*args):
    return subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=120)


def test_cli_render_demo_writes_film(rp, gpu_available, tmp_path):
    out = tmp_path / "demo.exr"
    r = run_cli("--render-demo", "--output", out, "--width", 64, "--height", 48, "--samples", 16, "--timings")
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and out.exists()
    rgb, e = exr_util.read_rgb(out)
    assert (e["width"], e["height"]) == (64, 48)
    assert e["strings"]["rypath.samples"] == "16"
    assert "Sampler: philox4x32-10-v1" in r.stdout
    # same film through the library: the demo scene, scalar megakernel, Philox
    s = rp.Synthetic(rp.SYNTH_DEMO, width=64, height=48)
    xyz, _ = rp.Scene(s.desc).render_frame(rp.frame_config(
        16, mode=rp.MODE_MEGAKERNEL, seed=1, randomization=rp.RANDOMIZE_NONE, flags=rp.FRAME_INDEPENDENT_SAMPLER))
    np.testing.assert_allclose(rgb, (xyz.astype(np.float64) @ M_SRGB.T).astype(np.float32), rtol=1e-5, atol=1e-6)


def test_cli_render_pbrt(gpu_available, tmp_path):
    out = tmp_path / "cornell.exr"
    r = run_cli("--render-pbrt", CORNELL, "--output", out, "--samples", 4, "--half")
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    rgb, e = exr_util.read_rgb(out)
    assert (e["width"], e["height"]) == (256, 256)
    assert {t for _, t in e["channels"]} == {1}
    assert np.isfinite(rgb).all() and rgb.mean() > 0


def test_checkpoint_resume_equals_uninterrupted(rp, gpu_available, tmp_path):
    """Accumulate 8 + 8 samples, checkpoint, read back, accumulate 16 more:
    the film equals one 32-sample frame bit for bit (the Neumaier state is
    the whole accumulator), and the running means equal shorter frames."""
    s = rp.Synthetic(rp.SYNTH_VOLUME, seed=3, width=40, height=30, volume_subdivision=3)
    scene = rp.Scene(s.desc)
    ref32, _ = scene.render_frame(rp.frame_config(32, mode=rp.MODE_WAVEFRONT, seed=4))
    ref16, _ = scene.render_frame(rp.frame_config(16, mode=rp.MODE_WAVEFRONT, seed=4))
    film = rp.Film(40, 30)
    for begin in (0, 8):
        running, _ = film.accumulate(scene, rp.frame_config(8, mode=rp.MODE_WAVEFRONT, sample_begin=begin, seed=4))
    assert np.array_equal(running, ref16) and np.array_equal(film.resolve(), ref16)
    path = tmp_path / "film.ckpt"
    film.write(path)
    resumed = rp.Film.read(path)
    info = resumed.info()
    assert info["samples"] == 16 and info["scene_identity"] == scene.identity() and info["segments"] == 1
    out, _ = resumed.accumulate(scene, rp.frame_config(16, mode=rp.MODE_MEGAKERNEL, sample_begin=16, seed=4))
    assert np.array_equal(out, ref32) and np.array_equal(resumed.resolve(), ref32)
    print("[checkpoint] resumed film equals the uninterrupted 32-sample frame", flush=True)
    # continuation rules
    with pytest.raises(rp.RypathError):
        resumed.accumulate(scene, rp.frame_config(4, sample_begin=16, seed=4))  # not at the film's end
    with pytest.raises(rp.RypathError):
        resumed.accumulate(scene, rp.frame_config(4, sample_begin=32, seed=5))  # another seed
    other = rp.Scene(rp.Synthetic(rp.SYNTH_VOLUME, seed=4, width=40, height=30, volume_subdivision=3).desc)
    with pytest.raises(rp.RypathError) as e:
        resumed.accumulate(other, rp.frame_config(4, sample_begin=32, seed=4))
    assert "another scene" in str(e.value)


def test_checkpoint_records_transport_and_sampler(rp, gpu_available, tmp_path):
    """Format-3 files written after accumulation carry the film model, packet
    flag and sampler of the transport that rendered them, and a film of one
    sampler does not continue with the other."""
    from tests.test_checkpoint import decode3

    demo = rp.Scene(rp.Synthetic(rp.SYNTH_DEMO, width=24, height=16).desc)
    film = rp.Film(24, 16)
    film.set_job(8, 4)
    cfg = rp.frame_config(4, mode=rp.MODE_MEGAKERNEL, seed=3, randomization=rp.RANDOMIZE_NONE,
                          flags=rp.FRAME_INDEPENDENT_SAMPLER)
    film.accumulate(demo, cfg)
    film.write(tmp_path / "demo.ckpt")
    d = decode3((tmp_path / "demo.ckpt").read_bytes())
    assert (d["sampler"], d["model"], d["packet"], d["job"], d["batch"]) == (0, 0, 0, 8, 4)
    assert (d["pixels"]["count"] == 4).all() and (d["pixels"]["w"] == 0).all()
    seg = d["segments"][0]
    assert seg[:2] == (0, 4) and seg[2].startswith("rypath_amd") and "gfx950" in seg[4]
    with pytest.raises(rp.RypathError) as e:  # Sobol samples onto an independent-sampler film
        film.accumulate(demo, rp.frame_config(4, mode=rp.MODE_MEGAKERNEL, sample_begin=4, seed=3,
                                              randomization=rp.RANDOMIZE_NONE))
    assert "sampler" in str(e.value)
    film.accumulate(demo, rp.frame_config(8, mode=rp.MODE_MEGAKERNEL, sample_begin=4, seed=3,
                                          randomization=rp.RANDOMIZE_NONE, flags=rp.FRAME_INDEPENDENT_SAMPLER))
    film.write(tmp_path / "demo12.ckpt")  # grown past its job of 8: the count is written as the job
    d = decode3((tmp_path / "demo12.ckpt").read_bytes())
    assert (d["job"], d["segments"][0][:2]) == (12, (0, 12))
    cornell = rp.Scene(rp.Synthetic(rp.SYNTH_CORNELL, seed=1, width=24, height=16).desc)
    hero = rp.Film(24, 16)
    hero.accumulate(cornell, rp.frame_config(3, mode=rp.MODE_WAVEFRONT, seed=3))
    hero.write(tmp_path / "hero.ckpt")
    d = decode3((tmp_path / "hero.ckpt").read_bytes())
    assert (d["sampler"], d["model"], d["packet"], d["job"]) == (1, 1, 1, 3)
    assert (d["pixels"]["w"][:, 0] == 3).all()
    back = rp.Film.read(tmp_path / "hero.ckpt")
    np.testing.assert_array_equal(back.resolve(), hero.resolve())


def test_cli_checkpoint_and_resume(gpu_available, tmp_path):
    ck = tmp_path / "run.ckpt"
    full = tmp_path / "full.exr"
    assert run_cli("--render-demo", "--output", full, "--width", 32, "--height", 32, "--samples", 24).returncode == 0
    part = tmp_path / "part.exr"
    r = run_cli("--render-demo", "--output", part, "--width", 32, "--height", 32, "--samples", 8,
                "--checkpoint", ck, "--batch-size", 4)
    assert r.returncode == 0 and ck.exists()
    resumed = tmp_path / "resumed.exr"
    r = run_cli("--render-demo", "--output", resumed, "--width", 32, "--height", 32, "--samples", 24,
                "--resume", ck, "--checkpoint", ck)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    a, _ = exr_util.read_rgb(full)
    b, e = exr_util.read_rgb(resumed)
    assert np.array_equal(a, b)
    assert len(e["strings"]["rypath.checkpointIdentity"]) == 64


def test_film_identity_follows_parameter_changes(rp, gpu_available):
    """rp_set_params rewrites values the scene identity hashes: a film
    accumulated before the change cannot be continued after it, whichever
    call computed the identity first."""
    s = rp.Synthetic(rp.SYNTH_INVERSE, seed=2, width=16, height=16)
    scene = rp.Scene(s.desc)
    params, values = rp.inverse_params(s)
    film = rp.Film(16, 16)
    film.accumulate(scene, rp.frame_config(2, seed=4), film=False)
    before = scene.identity()
    film.check(scene, 4, rp.RANDOMIZE_FASTOWEN)
    scene.set_params(params, values * np.float32(0.9))
    assert scene.identity() != before
    with pytest.raises(rp.RypathError) as e:
        film.check(scene, 4, rp.RANDOMIZE_FASTOWEN)
    assert "another scene" in str(e.value)
    with pytest.raises(rp.RypathError):
        film.accumulate(scene, rp.frame_config(2, sample_begin=2, seed=4), film=False)
    scene.set_params(params, values)
    assert scene.identity() == before
    film.accumulate(scene, rp.frame_config(2, sample_begin=2, seed=4), film=False)
    assert film.info()["samples"] == 4
