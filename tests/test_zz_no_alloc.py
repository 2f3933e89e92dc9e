"""rp_render_frame_device_noalloc against rp_render_frame_device (megakernel
mode), called through the C ABI: the films must be bit-identical and the ray
counts equal, on two calls in a row, the second with a smaller film.  The
first call (1920x1080, 2 spp) fills the static scratch array with its film
and runs in two sample chunks; the second (64x48 Cornell box, 5 spp) must
not see anything the first left behind.  Sizes beyond the array and the
queue modes are refused before anything runs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _frames(rp, scene, cfg):
    import torch

    n = scene.width * scene.height * 3
    a = torch.full((n,), float("nan"), dtype=torch.float32, device="cuda")
    b = torch.full((n,), float("nan"), dtype=torch.float32, device="cuda")
    st_a = scene.render_frame_device(cfg, a.data_ptr())
    st_b = scene.render_frame_device_noalloc(cfg, b.data_ptr())
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy(), st_a, st_b


def test_noalloc_frame_equals_frame_device_twice(rp, gpu_available):
    cases = [(rp.Synthetic(rp.SYNTH_TERRAIN, seed=1, width=1920, height=1080), 2, 3),
             (rp.Synthetic(rp.SYNTH_CORNELL, seed=1, width=64, height=48), 5, 7)]
    for synth, spp, seed in cases:
        scene = rp.Scene(synth.desc)
        cfg = rp.frame_config(spp, mode=rp.MODE_MEGAKERNEL, seed=seed)
        print(f"[noalloc] {scene.width}x{scene.height} {spp} spp seed {seed}", flush=True)
        a, b, st_a, st_b = _frames(rp, scene, cfg)
        print(f"[noalloc] frame_device {st_a} noalloc {st_b}", flush=True)
        assert np.isfinite(a).all()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        for k in ("paths", "extension_rays", "shadow_rays", "failed_paths"):
            assert st_a[k] == st_b[k], k


def test_noalloc_refuses_what_it_cannot_hold(rp, gpu_available):
    import torch

    too_big = rp.Scene(rp.Synthetic(rp.SYNTH_CORNELL, seed=1, width=2048, height=1100).desc)
    out = torch.zeros(2048 * 1100 * 3, dtype=torch.float32, device="cuda")
    with pytest.raises(rp.RypathError) as e:
        too_big.render_frame_device_noalloc(rp.frame_config(1, mode=rp.MODE_MEGAKERNEL), out.data_ptr())
    assert e.value.code == rp.INVALID
    small = rp.Scene(rp.Synthetic(rp.SYNTH_CORNELL, seed=1, width=32, height=32).desc)
    with pytest.raises(rp.RypathError) as e:
        small.render_frame_device_noalloc(rp.frame_config(1, mode=rp.MODE_WAVEFRONT), out.data_ptr())
    assert e.value.code == rp.INVALID
