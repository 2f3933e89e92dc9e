"""CPU-side checks of the C ABI library: it loads, exports every symbol the
headers declare, validates inputs and reports errors without a GPU."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[\w\s\*]+?\b(rp_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(rp):
    lib = rp.library()
    declared = declared_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), f"{name} declared in include/ but not exported"
    assert set(rp.EXPORTED) <= declared


def test_api_version(rp):
    assert rp.library().rp_api_version() == 1


def test_struct_layouts(rp):
    # sizes fixed by include/rp_api.h
    assert C.sizeof(rp.Spectrum) == 24
    assert C.sizeof(rp.Material) == 88
    assert C.sizeof(rp.Camera) == 64
    assert C.sizeof(rp.Request) == 72
    assert C.sizeof(rp.FrameConfig) == 48
    assert C.sizeof(rp.FrameStats) == 144
    assert C.sizeof(rp.BatchConfig) == 32
    assert rp.RECORD_DTYPE.itemsize == 112
    assert C.sizeof(rp.Ray) == 32 and C.sizeof(rp.Hit) == 16


def test_scene_create_rejects_invalid_input(rp):
    s = rp.Synthetic(rp.SYNTH_CORNELL, seed=1, width=8, height=8)
    d = rp.SceneDesc.from_buffer_copy(s.d)
    d.api_version = 99
    with pytest.raises(rp.RypathError) as e:
        rp.Scene(d)
    assert e.value.code == rp.INVALID and "api_version" in str(e.value)
    d = rp.SceneDesc.from_buffer_copy(s.d)
    d.max_depth = 500
    with pytest.raises(rp.RypathError) as e:
        rp.Scene(d)
    assert "max_depth" in str(e.value)


def test_scene_create_without_device(rp):
    if rp.device_count() > 0:
        pytest.skip("a HIP device is present")
    s = rp.Synthetic(rp.SYNTH_CORNELL, seed=1, width=8, height=8)
    with pytest.raises(rp.RypathError) as e:
        rp.Scene(s.desc)
    assert e.value.code == rp.NO_DEVICE


def test_null_arguments_return_invalid(rp):
    lib = rp.library()
    assert lib.rp_scene_create(None, None, None) == rp.INVALID
    assert lib.rp_sample(None, None, None, 0) == rp.INVALID
    assert lib.rp_batch_create(None, 0, None) == rp.INVALID
    buf = C.create_string_buffer(256)
    lib.rp_last_error(buf, 256)
    assert b"rp_batch_create" in buf.value


@pytest.mark.parametrize("kind,tris", [(0, 36), (1, None), (3, None), (4, None)])
def test_synthetic_scenes(rp, kind, tris):
    opts = dict(width=16, height=16)
    if kind == 1:
        opts.update(grid=64, spheres=4, sphere_subdivision=2)
    if kind == 3:
        opts.update(volume_subdivision=2)
    s = rp.Synthetic(kind, seed=1, **opts)
    d = s.d
    if tris:
        assert d.triangle_count == tris
    pos, idx = s.triangles()
    assert np.isfinite(pos).all() and idx.max() < d.vertex_count
    objs = np.ctypeslib.as_array(d.tri_object, shape=(d.triangle_count,))
    assert objs.max() < d.object_count


def test_terrain_config_sizes(rp):
    s = rp.Synthetic(rp.SYNTH_TERRAIN, seed=1)
    d = s.d
    assert d.camera.width == 1920 and d.camera.height == 1080 and d.max_depth == 8
    assert d.triangle_count == 1023 * 1023 * 2 + 64 * 5120 + 32
    kinds = [d.materials[i].kind for i in range(d.material_count)]
    assert set(kinds) <= {0, 1, 2}


def test_probe_and_device_backward_reject_invalid_arguments(rp):
    """Argument checks that run before any device call."""
    with pytest.raises(rp.RypathError) as e:
        rp.probe_read_bandwidth(100, 1)  # below 4096 bytes
    assert e.value.code == rp.INVALID
    with pytest.raises(rp.RypathError) as e:
        rp.probe_read_bandwidth(4096 + 8, 1)  # not a multiple of 16
    assert e.value.code == rp.INVALID
    with pytest.raises(rp.RypathError) as e:
        rp.probe_read_bandwidth(1 << 20, 0)  # no pass
    assert e.value.code == rp.INVALID
    lib = rp.library()
    g = (C.c_float * 4)()
    cfg = rp.frame_config(1)
    assert lib.rp_backward_device(None, C.byref(cfg), None, 0, C.c_void_p(1), g, None) == rp.INVALID
    buf = C.create_string_buffer(256)
    lib.rp_last_error(buf, len(buf))
    assert b"rp_backward_device" in buf.value
