"""ctypes bindings of the gfx950 path tracer's C ABI (include/rp_api.h).

The HIP library rypath_amd/librypath_amd.so is required: importing this
package raises if it is missing, and every rendering call runs on the GPU
through the C ABI (there is no CPU fallback).
"""
import atexit
import ctypes as C
import itertools
import weakref
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "librypath_amd.so"

if not _LIB_PATH.exists():
    raise ImportError(f"{_LIB_PATH} is missing: build it with `python __graft_entry__.py` (hipcc, gfx950)")

# PyTorch wheels ship their own HIP runtime.  Loading torch first makes this
# library bind to that same runtime, so one process holds one HIP runtime
# (a second runtime initialised after ours leaves torch without devices).
try:
    import torch  # noqa: F401
except ImportError:
    pass

_lib = C.CDLL(str(_LIB_PATH))

# Live native handles, destroyed newest first by an exit hook that runs
# while the module's globals still exist (a batch before its scenes, a
# frame before the library goes away).  Objects collected earlier release
# their handle in __del__ and leave the registry.
_live = {}
_seq = itertools.count()


def _track(obj, destroy):
    key = next(_seq)
    obj._reg = key
    _live[key] = (weakref.ref(obj), destroy)


def _release(obj):
    h = getattr(obj, "_h", None)
    entry = _live.pop(getattr(obj, "_reg", None), None)
    if h and entry is not None:
        obj._h = None
        getattr(_lib, entry[1])(h)


@atexit.register
def _teardown():
    for key in sorted(_live, reverse=True):
        obj = _live[key][0]()
        if obj is not None:
            _release(obj)
    _live.clear()

NONE = 0xFFFFFFFF
SUCCESS, NOT_BUILT, NO_DEVICE, FAILURE, INVALID, BUSY = range(6)
SPECTRUM_CONSTANT, SPECTRUM_SIGMOID, SPECTRUM_TABLE, SPECTRUM_CAUCHY, SPECTRUM_LINEAR, SPECTRUM_KNOTS8 = range(6)
TRANSPORT_HERO, TRANSPORT_SCALAR, TRANSPORT_SCALAR_DEMO = range(3)
FRAME_INDEPENDENT_SAMPLER = 8
MATERIAL_DIFFUSE, MATERIAL_CONDUCTOR, MATERIAL_DIELECTRIC, MATERIAL_INTERFACE = range(4)
RANDOMIZE_NONE, RANDOMIZE_PERMUTE, RANDOMIZE_FASTOWEN, RANDOMIZE_OWEN = range(4)
MODE_MEGAKERNEL, MODE_WAVEFRONT, MODE_WAVEFRONT_SORTED = range(3)
SYNTH_CORNELL, SYNTH_TERRAIN, SYNTH_BATCH, SYNTH_VOLUME, SYNTH_INVERSE, SYNTH_DEMO = range(6)
PARAM_REFLECTANCE_BIN, PARAM_ROUGHNESS, PARAM_SIGMA_S_BIN = range(3)
SCHED_PERSISTENT, SCHED_WAVEFRONT = range(2)
SAMPLER_INDEPENDENT, SAMPLER_SOBOL, SAMPLER_HALTON, SAMPLER_ZSOBOL = range(4)
LIGHT_SAMPLER_BVH, LIGHT_SAMPLER_POWER, LIGHT_SAMPLER_UNIFORM = range(3)
u32, u64, f32, f64 = C.c_uint32, C.c_uint64, C.c_float, C.c_double


class Spectrum(C.Structure):
    _fields_ = [("kind", u32), ("c", f32 * 3), ("scale", f32), ("table", u32)]


class Material(C.Structure):
    _fields_ = [("kind", u32), ("remap_roughness", u32), ("roughness_u", f32), ("roughness_v", f32),
                ("reflectance", Spectrum), ("eta", Spectrum), ("k", Spectrum)]


class Object(C.Structure):
    _fields_ = [("material", u32), ("emission", u32), ("medium_inside", u32), ("medium_outside", u32)]


class Emission(C.Structure):
    _fields_ = [("radiance", Spectrum), ("two_sided", u32), ("pad", u32 * 3)]


class SpotLight(C.Structure):
    _fields_ = [("position", f32 * 3), ("direction", f32 * 3), ("cos_falloff_start", f32),
                ("cos_falloff_end", f32), ("intensity", Spectrum)]


class Medium(C.Structure):
    _fields_ = [("sigma_a", Spectrum), ("sigma_s", Spectrum), ("g", f32), ("pad", u32 * 3)]


class Camera(C.Structure):
    _fields_ = [("origin", f32 * 3), ("corner", f32 * 3), ("dx", f32 * 3), ("dy", f32 * 3),
                ("width", u32), ("height", u32), ("medium", u32), ("pad", u32)]


class SceneDesc(C.Structure):
    _fields_ = [("api_version", u32), ("vertex_count", u32), ("triangle_count", u32), ("object_count", u32),
                ("positions", C.POINTER(f32)), ("indices", C.POINTER(u32)), ("tri_object", C.POINTER(u32)),
                ("objects", C.POINTER(Object)),
                ("material_count", u32), ("emission_count", u32), ("medium_count", u32), ("spot_count", u32),
                ("materials", C.POINTER(Material)), ("emissions", C.POINTER(Emission)),
                ("media", C.POINTER(Medium)), ("spots", C.POINTER(SpotLight)), ("spectra", C.POINTER(f32)),
                ("spectra_count", u32), ("has_environment", u32), ("environment", Spectrum),
                ("camera", Camera), ("max_depth", u32), ("hero_wavelengths", u32), ("light_sampler", u32),
                ("transport", u32)]


class BuildOptions(C.Structure):
    _fields_ = [("leaf_size", u32), ("sah_bins", u32), ("threads", u32), ("device", u32)]


class BuildInfo(C.Structure):
    _fields_ = [("build_seconds", f64), ("upload_seconds", f64), ("node_count", u64), ("leaf_count", u64),
                ("device_bytes", u64), ("max_depth", u32), ("max_stack", u32), ("emitter_count", u32),
                ("light_count", u32)]


class Request(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("seed", u64), ("material_seed", C.c_int32),
                ("pixel_begin", u32), ("pixel_count", u32), ("sample_count", u32), ("sample_begin", u64),
                ("total_samples", u64), ("threadgroup_width", u32), ("emergency_limit", u32),
                ("sampler_kind", u32), ("randomization", u32), ("boundary_limit", u32), ("pad", u32)]


class FrameConfig(C.Structure):
    _fields_ = [("mode", u32), ("spp", u32), ("sample_begin", u64), ("seed", u64), ("randomization", u32),
                ("capacity", u32), ("group_size", u32), ("sort_bits", u32), ("flags", u32), ("tail_paths", u32)]


class FrameStats(C.Structure):
    _fields_ = [("seconds", f64), ("paths", u64), ("extension_rays", u64), ("shadow_rays", u64),
                ("failed_paths", u64), ("trace_seconds", f64), ("extend_launches", u64), ("node_fetches", u64),
                ("triangle_tests", u64), ("extend_rays", u64), ("shadow_node_fetches", u64),
                ("shadow_triangle_tests", u64), ("stack_spills", u64), ("shadow_seconds", f64),
                ("shadow_launches", u64), ("shadow_kernel_rays", u64), ("fused_launches", u64),
                ("fused_shadow_rays", u64)]


class Ray(C.Structure):
    _fields_ = [("origin", f32 * 3), ("tmax", f32), ("direction", f32 * 3), ("pad", u32)]


class Hit(C.Structure):
    _fields_ = [("triangle", u32), ("t", f32), ("b1", f32), ("b2", f32)]


class SyntheticOptions(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("max_depth", u32), ("grid", u32), ("spheres", u32),
                ("sphere_subdivision", u32), ("target_triangles", u32), ("volume_subdivision", u32),
                ("hg_g", f32), ("bins", u32), ("pad", u32)]


class BatchItem(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("spp", u32), ("pad", u32), ("seed", u64)]


class BatchConfig(C.Structure):
    _fields_ = [("capacity", u32), ("group_size", u32), ("sort_bits", u32), ("use_graph", u32),
                ("randomization", u32), ("scheduler", u32), ("tail_paths", u32), ("flags", u32)]


class BatchStats(C.Structure):
    _fields_ = [("seconds", f64), ("paths", u64), ("extension_rays", u64), ("shadow_rays", u64),
                ("failed_paths", u64), ("graph_launches", u32), ("pad", u32), ("mixed_scene_lanes", u64)]


class Param(C.Structure):
    _fields_ = [("kind", u32), ("target", u32), ("bin", u32), ("bins", u32)]


RECORD_DTYPE = np.dtype([
    ("pixel", "<u4"), ("status", "<u4"), ("sample", "<u8"), ("sobol_index", "<u8"),
    ("wavelengths", "<f4", 4), ("wavelength_pdfs", "<f4", 4), ("radiance", "<f4", 4), ("xyz", "<f4", 3),
    ("primary_triangle", "<u4"), ("vertices", "<u4"), ("camera_words", "<u4", 2), ("filter_weight", "<f4"),
    ("splat_begin", "<u4"), ("splat_count", "<u4"),
])
assert RECORD_DTYPE.itemsize == 112
RENDER_SAMPLE_DTYPE = np.dtype([
    ("pixel", "<u4"), ("status", "<u4"), ("sample", "<u8"), ("wavelength", "<f4"), ("radiance", "<f4"),
    ("scatterings", "<u4"), ("random_words", "<u4", 4), ("xyz", "<f4", 3),
])
assert RENDER_SAMPLE_DTYPE.itemsize == 56
RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("tmax", "<f4"), ("direction", "<f4", 3), ("pad", "<u4")])
HIT_DTYPE = np.dtype([("triangle", "<u4"), ("t", "<f4"), ("b1", "<f4"), ("b2", "<f4")])


def _sig(name, res, args):
    fn = getattr(_lib, name)
    fn.restype = res
    fn.argtypes = args
    return fn


_sig("rp_api_version", u32, [])
_sig("rp_device_count", C.c_int, [])
_sig("rp_last_error", C.c_size_t, [C.c_char_p, C.c_size_t])
_sig("rp_scene_create", C.c_int, [C.POINTER(SceneDesc), C.POINTER(BuildOptions), C.POINTER(C.c_void_p)])
_sig("rp_scene_info", C.c_int, [C.c_void_p, C.POINTER(BuildInfo)])
_sig("rp_scene_destroy", None, [C.c_void_p])
_sig("rp_sample", C.c_int, [C.c_void_p, C.POINTER(Request), C.c_void_p, u64])
_sig("rp_render_frame", C.c_int, [C.c_void_p, C.POINTER(FrameConfig), C.c_void_p, C.POINTER(FrameStats)])
_sig("rp_render_frame_device", C.c_int, [C.c_void_p, C.POINTER(FrameConfig), C.c_void_p, C.POINTER(FrameStats)])
_sig("rp_sample_scalar", C.c_int, [C.c_void_p, C.POINTER(Request), C.c_void_p, u64])
_sig("rp_render_frame_device_noalloc", C.c_int,
     [C.c_void_p, C.POINTER(FrameConfig), C.c_void_p, C.POINTER(FrameStats)])
_sig("rp_trace_closest", C.c_int, [C.c_void_p, C.c_void_p, u64, C.c_void_p, C.POINTER(f64)])
_sig("rp_trace_any", C.c_int, [C.c_void_p, C.c_void_p, u64, C.c_void_p, C.POINTER(f64)])
_sig("rp_probe_read_bandwidth", C.c_int, [C.c_int, u64, u32, C.POINTER(f64)])
_sig("rp_batch_create", C.c_int, [C.POINTER(BatchItem), u32, C.POINTER(C.c_void_p)])
_sig("rp_batch_render", C.c_int, [C.c_void_p, C.POINTER(BatchConfig), C.POINTER(C.c_void_p), C.POINTER(BatchStats)])
_sig("rp_batch_destroy", None, [C.c_void_p])
_sig("rp_autotune", C.c_int, [C.c_void_p, C.POINTER(BatchConfig), C.POINTER(f64)])
_sig("rp_set_params", C.c_int, [C.c_void_p, C.POINTER(Param), u32, C.POINTER(f32)])
_sig("rp_backward", C.c_int, [C.c_void_p, C.POINTER(FrameConfig), C.POINTER(Param), u32, C.POINTER(f32),
                              C.POINTER(f32), C.POINTER(FrameStats)])
_sig("rp_backward_device", C.c_int, [C.c_void_p, C.POINTER(FrameConfig), C.POINTER(Param), u32, C.c_void_p,
                                     C.POINTER(f32), C.POINTER(FrameStats)])
_sig("rp_synthetic_create", C.c_int, [u32, u64, C.POINTER(SyntheticOptions), C.POINTER(C.c_void_p)])
_sig("rp_synthetic_desc", C.POINTER(SceneDesc), [C.c_void_p])
_sig("rp_synthetic_destroy", None, [C.c_void_p])


class PbrtInfo(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("spp", u32), ("randomization", u32), ("seed", u64),
                ("max_depth", u32), ("volumetric", u32), ("output", C.c_char * 256)]


_sig("rp_pbrt_load", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)])
_sig("rp_pbrt_load_string", C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)])
_sig("rp_pbrt_desc", C.POINTER(SceneDesc), [C.c_void_p])
_sig("rp_pbrt_get_info", C.c_int, [C.c_void_p, C.POINTER(PbrtInfo)])
_sig("rp_pbrt_destroy", None, [C.c_void_p])
_sig("rp_write_exr", C.c_int, [C.c_char_p, u32, u32, C.POINTER(f32), f64, u32, C.POINTER(C.c_char_p),
                               C.POINTER(C.c_char_p), u32])
class FilmInfo(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("samples", u64), ("seed", u64), ("randomization", u32),
                ("segments", u32), ("scene_identity", C.c_uint8 * 32), ("sampler_kind", u32), ("model", u32),
                ("job_samples", u64), ("batch_size", u64), ("packet", u32), ("format", u32)]


_sig("rp_film_create", C.c_int, [u32, u32, C.POINTER(C.c_void_p)])
_sig("rp_film_check", C.c_int, [C.c_void_p, C.c_void_p, u64, u32])
_sig("rp_film_accumulate", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(FrameConfig), C.c_void_p,
                                     C.POINTER(FrameStats)])
_sig("rp_film_resolve", C.c_int, [C.c_void_p, C.c_void_p])
_sig("rp_film_get_info", C.c_int, [C.c_void_p, C.POINTER(FilmInfo)])
_sig("rp_film_set_job", C.c_int, [C.c_void_p, u64, u64])
_sig("rp_checkpoint_write", C.c_int, [C.c_void_p, C.c_char_p])
_sig("rp_checkpoint_read", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)])
_sig("rp_film_destroy", None, [C.c_void_p])
_sig("rp_scene_identity", C.c_int, [C.c_void_p, C.POINTER(C.c_uint8)])
_sig("rp_sha256", None, [C.c_void_p, C.c_size_t, C.POINTER(C.c_uint8)])
_sig("rp_frame_submit", C.c_int, [C.c_void_p, C.POINTER(FrameConfig), C.POINTER(C.c_void_p)])
_sig("rp_frame_poll", C.c_int, [C.c_void_p, C.POINTER(u32)])
_sig("rp_frame_wait", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(FrameStats)])
_sig("rp_frame_release", None, [C.c_void_p])

EXPORTED = [
    "rp_api_version", "rp_device_count", "rp_last_error", "rp_scene_create", "rp_scene_info", "rp_scene_destroy",
    "rp_sample", "rp_sample_scalar", "rp_render_frame", "rp_render_frame_device", "rp_render_frame_device_noalloc", "rp_trace_closest", "rp_trace_any",
    "rp_probe_read_bandwidth",
    "rp_batch_create", "rp_batch_render", "rp_batch_destroy", "rp_autotune", "rp_set_params", "rp_backward",
    "rp_backward_device",
    "rp_synthetic_create", "rp_synthetic_desc", "rp_synthetic_destroy",
    "rp_pbrt_load", "rp_pbrt_load_string", "rp_pbrt_desc", "rp_pbrt_get_info", "rp_pbrt_destroy",
    "rp_write_exr", "rp_frame_submit", "rp_frame_poll", "rp_frame_wait", "rp_frame_release",
    "rp_film_create", "rp_film_check", "rp_film_accumulate", "rp_film_resolve", "rp_film_get_info", "rp_film_set_job", "rp_checkpoint_write",
    "rp_checkpoint_read", "rp_film_destroy", "rp_scene_identity", "rp_sha256",
]


class RypathError(RuntimeError):
    def __init__(self, where, code):
        buf = C.create_string_buffer(4096)
        _lib.rp_last_error(buf, len(buf))
        super().__init__(f"{where} failed with status {code}: {buf.value.decode(errors='replace')}")
        self.code = code


def _check(where, code):
    if code != SUCCESS:
        raise RypathError(where, code)


def library():
    return _lib


def device_count():
    return _lib.rp_device_count()


class Synthetic:
    """Owned synthetic scene description (include/rp_synthetic.h)."""

    def __init__(self, kind, seed=1, **options):
        o = SyntheticOptions()
        for k, v in options.items():
            setattr(o, k, v)
        h = C.c_void_p()
        _check("rp_synthetic_create", _lib.rp_synthetic_create(kind, seed, C.byref(o), C.byref(h)))
        self._h = h
        _track(self, "rp_synthetic_destroy")
        self.desc = _lib.rp_synthetic_desc(h)
        self.desc._owner = self  # the pointer keeps the storage it points into alive

    def __del__(self):
        if getattr(self, "_h", None):  # the exit hook has released every handle before module teardown
            _release(self)

    @property
    def d(self):
        c = self.desc.contents
        c._owner = self  # the description points into this object's storage
        return c

    def triangles(self):
        d = self.d
        pos = np.ctypeslib.as_array(d.positions, shape=(d.vertex_count, 3)).copy()
        idx = np.ctypeslib.as_array(d.indices, shape=(d.triangle_count, 3)).copy()
        return pos, idx


def batch_mix(count=64, seed=2026, resolution_divisor=1):
    """The uneven batch workload: `count` seeded scenes built like the large
    terrain frame, triangle counts log-uniform in [5k, 500k], square
    resolutions drawn from {128, 256, 512, 1024} (each divided by
    `resolution_divisor`), samples per pixel drawn from {4, 16, 64}, depth 8.
    Returns a list of (Synthetic, spp, sample seed), in draw order."""
    rng = np.random.default_rng(seed)
    mix = []
    for i in range(count):
        tris = int(np.exp(rng.uniform(np.log(5e3), np.log(5e5))))
        res = max(1, int(rng.choice([128, 256, 512, 1024])) // resolution_divisor)
        spp = int(rng.choice([4, 16, 64]))
        synth = Synthetic(SYNTH_BATCH, seed=1000 + i, width=res, height=res, target_triangles=tris)
        mix.append((synth, spp, 1 + i))
    return mix


def request(seed=1, pixel_begin=0, pixel_count=1, sample_begin=0, sample_count=1,
            randomization=RANDOMIZE_FASTOWEN, emergency_limit=0, boundary_limit=0, width=0, height=0,
            material_seed=0, total_samples=0, threadgroup_width=0, sampler_kind=None):
    """An rp_request (RenderRequest); width/height 0 take the scene camera's."""
    r = Request()
    r.width, r.height, r.seed, r.material_seed = width, height, seed, material_seed
    r.randomization, r.pixel_begin, r.pixel_count = randomization, pixel_begin, pixel_count
    r.sample_begin, r.sample_count, r.total_samples = sample_begin, sample_count, total_samples
    r.threadgroup_width, r.emergency_limit, r.boundary_limit = threadgroup_width, emergency_limit, boundary_limit
    r.sampler_kind = SAMPLER_SOBOL if sampler_kind is None else sampler_kind
    return r


FRAME_COUNT_TRAFFIC = 1
FRAME_NO_ROULETTE = 2
FRAME_NO_TAIL = 4
BATCH_NO_TAIL = 1


def frame_config(spp, mode=MODE_WAVEFRONT, sample_begin=0, seed=1, randomization=RANDOMIZE_FASTOWEN,
                 capacity=0, group_size=0, sort_bits=0, flags=0, tail_paths=0):
    c = FrameConfig()
    c.flags, c.tail_paths = flags, tail_paths
    c.mode, c.spp, c.sample_begin, c.seed, c.randomization = mode, spp, sample_begin, seed, randomization
    c.capacity, c.group_size, c.sort_bits = capacity, group_size, sort_bits
    return c


class Scene:
    """Device scene: BVH built on the host, uploaded to HBM (rp_scene_create)."""

    def __init__(self, desc, leaf_size=0, sah_bins=0, device=0):
        o = BuildOptions()
        o.leaf_size, o.sah_bins, o.device = leaf_size, sah_bins, device
        h = C.c_void_p()
        ptr = desc if isinstance(desc, C._Pointer) else C.pointer(desc)
        _check("rp_scene_create", _lib.rp_scene_create(ptr, C.byref(o), C.byref(h)))
        self._h = h
        _track(self, "rp_scene_destroy")
        cam = ptr.contents.camera
        self.width, self.height = cam.width, cam.height
        # the parameter values last written through set_params (their order
        # and values), which the autograd harness compares against
        self.param_state = None

    @property
    def handle(self):
        return self._h

    def __del__(self):
        if getattr(self, "_h", None):  # the exit hook has released every handle before module teardown
            _release(self)

    def info(self):
        i = BuildInfo()
        _check("rp_scene_info", _lib.rp_scene_info(self._h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in BuildInfo._fields_}

    def sample(self, req):
        n = req.pixel_count * req.sample_count
        out = np.zeros(n, dtype=RECORD_DTYPE)
        _check("rp_sample", _lib.rp_sample(self._h, C.byref(req), out.ctypes.data_as(C.c_void_p), n))
        return out

    def sample_scalar(self, req):
        """RenderSample records of a scalar-transport scene (rp_sample_scalar)."""
        n = req.pixel_count * req.sample_count
        out = np.zeros(n, dtype=RENDER_SAMPLE_DTYPE)
        _check("rp_sample_scalar", _lib.rp_sample_scalar(self._h, C.byref(req), out.ctypes.data_as(C.c_void_p), n))
        return out

    def render_frame(self, cfg):
        out = np.zeros((self.height, self.width, 3), dtype=np.float32)
        st = FrameStats()
        _check("rp_render_frame", _lib.rp_render_frame(self._h, C.byref(cfg), out.ctypes.data_as(C.c_void_p),
                                                        C.byref(st)))
        return out, {f: getattr(st, f) for f, _ in FrameStats._fields_}

    def render_frame_device(self, cfg, device_ptr):
        st = FrameStats()
        _check("rp_render_frame_device", _lib.rp_render_frame_device(self._h, C.byref(cfg), C.c_void_p(device_ptr),
                                                                      C.byref(st)))
        return {f: getattr(st, f) for f, _ in FrameStats._fields_}

    def render_frame_device_noalloc(self, cfg, device_ptr):
        """rp_render_frame_device_noalloc: the megakernel frame with every
        scratch buffer in the library's static device array (no allocation)."""
        st = FrameStats()
        _check("rp_render_frame_device_noalloc",
               _lib.rp_render_frame_device_noalloc(self._h, C.byref(cfg), C.c_void_p(device_ptr), C.byref(st)))
        return {f: getattr(st, f) for f, _ in FrameStats._fields_}

    def set_params(self, params, values):
        """Writes parameter values (rp_set_params); params: list of (kind, target, bin, bins)."""
        arr = params_array(params)
        vals = np.ascontiguousarray(values, dtype=np.float32)
        assert len(vals) == len(arr)
        self.param_state = None
        _check("rp_set_params", _lib.rp_set_params(self._h, arr, len(arr), vals.ctypes.data_as(C.POINTER(f32))))
        self.param_state = (tuple(tuple(p) for p in params), vals.tobytes())

    def backward(self, cfg, params, dl_dxyz):
        """Path-replay adjoint: d(sum dl_dxyz * XYZ)/d params for the frame `cfg` (rp_backward)."""
        arr = params_array(params)
        dl = np.ascontiguousarray(dl_dxyz, dtype=np.float32).reshape(-1)
        assert dl.size == self.width * self.height * 3
        g = np.zeros(len(arr), np.float32)
        st = FrameStats()
        _check("rp_backward", _lib.rp_backward(self._h, C.byref(cfg), arr, len(arr), dl.ctypes.data_as(C.POINTER(f32)),
                                                g.ctypes.data_as(C.POINTER(f32)), C.byref(st)))
        return g, {f: getattr(st, f) for f, _ in FrameStats._fields_}

    def backward_device(self, cfg, params, dl_ptr):
        """rp_backward with dL/dXYZ (H*W*3 float32, contiguous) in device memory
        at `dl_ptr`, e.g. a CUDA tensor's data_ptr()."""
        arr = params_array(params)
        g = np.zeros(len(arr), np.float32)
        st = FrameStats()
        _check("rp_backward_device", _lib.rp_backward_device(self._h, C.byref(cfg), arr, len(arr), C.c_void_p(dl_ptr),
                                                              g.ctypes.data_as(C.POINTER(f32)), C.byref(st)))
        return g, {f: getattr(st, f) for f, _ in FrameStats._fields_}

    def trace_closest(self, rays):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.zeros(len(rays), dtype=HIT_DTYPE)
        sec = f64()
        _check("rp_trace_closest", _lib.rp_trace_closest(self._h, rays.ctypes.data_as(C.c_void_p), len(rays),
                                                          hits.ctypes.data_as(C.c_void_p), C.byref(sec)))
        return hits, sec.value

    def trace_any(self, rays):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        occ = np.zeros(len(rays), dtype=np.uint8)
        sec = f64()
        _check("rp_trace_any", _lib.rp_trace_any(self._h, rays.ctypes.data_as(C.c_void_p), len(rays),
                                                  occ.ctypes.data_as(C.c_void_p), C.byref(sec)))
        return occ, sec.value

    def identity(self):
        d = (C.c_uint8 * 32)()
        _check("rp_scene_identity", _lib.rp_scene_identity(self._h, d))
        return bytes(d)

    def submit(self, cfg):
        """Asynchronous frame (rp_frame_submit); returns a Frame, or raises with code BUSY
        when two frames are already retained."""
        h = C.c_void_p()
        _check("rp_frame_submit", _lib.rp_frame_submit(self._h, C.byref(cfg), C.byref(h)))
        return Frame(h, self.width, self.height)


class Frame:
    """A submitted frame (FrameSession::submit result): poll, wait, release."""

    def __init__(self, handle, width, height):
        self._h, self.width, self.height = handle, width, height
        _track(self, "rp_frame_release")

    def poll(self):
        done = u32()
        _check("rp_frame_poll", _lib.rp_frame_poll(self._h, C.byref(done)))
        return bool(done.value)

    def wait(self, film=True):
        out = np.zeros((self.height, self.width, 3), np.float32) if film else None
        st = FrameStats()
        ptr = out.ctypes.data_as(C.c_void_p) if film else None
        _check("rp_frame_wait", _lib.rp_frame_wait(self._h, ptr, C.byref(st)))
        return out, {f: getattr(st, f) for f, _ in FrameStats._fields_}

    def release(self):
        _release(self)

    def __del__(self):
        if getattr(self, "_h", None):  # the exit hook has released every handle before module teardown
            _release(self)


class Film:
    """Running film with checkpoint/resume (rp_film_*, rp_checkpoint_*)."""

    def __init__(self, width=None, height=None, handle=None):
        if handle is None:
            h = C.c_void_p()
            _check("rp_film_create", _lib.rp_film_create(width, height, C.byref(h)))
            handle = h
        self._h = handle
        _track(self, "rp_film_destroy")

    @classmethod
    def read(cls, path):
        h = C.c_void_p()
        _check("rp_checkpoint_read", _lib.rp_checkpoint_read(str(path).encode(), C.byref(h)))
        return cls(handle=h)

    def write(self, path):
        _check("rp_checkpoint_write", _lib.rp_checkpoint_write(self._h, str(path).encode()))

    def info(self):
        i = FilmInfo()
        _check("rp_film_get_info", _lib.rp_film_get_info(self._h, C.byref(i)))
        out = {f: getattr(i, f) for f, _ in FilmInfo._fields_}
        out["scene_identity"] = bytes(i.scene_identity)
        return out

    def set_job(self, job_samples, batch_size):
        """Records the job's total samples and batch size for checkpoints (rp_film_set_job)."""
        _check("rp_film_set_job", _lib.rp_film_set_job(self._h, job_samples, batch_size))

    def check(self, scene, seed, randomization):
        """Raises unless this film can continue `scene` with (seed, randomization) (rp_film_check)."""
        _check("rp_film_check", _lib.rp_film_check(scene.handle, self._h, seed, randomization))

    def accumulate(self, scene, cfg, film=True):
        i = self.info()
        out = np.zeros((i["height"], i["width"], 3), np.float32) if film else None
        st = FrameStats()
        ptr = out.ctypes.data_as(C.c_void_p) if film else None
        _check("rp_film_accumulate", _lib.rp_film_accumulate(scene.handle, self._h, C.byref(cfg), ptr, C.byref(st)))
        return out, {f: getattr(st, f) for f, _ in FrameStats._fields_}

    def resolve(self):
        i = self.info()
        out = np.zeros((i["height"], i["width"], 3), np.float32)
        _check("rp_film_resolve", _lib.rp_film_resolve(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def __del__(self):
        if getattr(self, "_h", None):  # the exit hook has released every handle before module teardown
            _release(self)


def sha256(data: bytes) -> bytes:
    d = (C.c_uint8 * 32)()
    _lib.rp_sha256(data, len(data), d)
    return bytes(d)


class PbrtScene:
    """PBRT-v4 scene-file subset translated to a scene description (rp_pbrt_load)."""

    def __init__(self, path=None, text=None, base_dir="."):
        h = C.c_void_p()
        if text is not None:
            _check("rp_pbrt_load_string", _lib.rp_pbrt_load_string(text.encode(), str(base_dir).encode(), C.byref(h)))
        else:
            _check("rp_pbrt_load", _lib.rp_pbrt_load(str(path).encode(), C.byref(h)))
        self._h = h
        _track(self, "rp_pbrt_destroy")
        self.desc = _lib.rp_pbrt_desc(h)
        self.desc._owner = self  # the pointer keeps the storage it points into alive
        i = PbrtInfo()
        _check("rp_pbrt_get_info", _lib.rp_pbrt_get_info(h, C.byref(i)))
        self.info = {f: getattr(i, f) for f, _ in PbrtInfo._fields_}
        self.info["output"] = i.output.decode()

    @property
    def d(self):
        c = self.desc.contents
        c._owner = self  # the description points into this object's storage
        return c

    def frame_config(self, mode=MODE_WAVEFRONT, sample_begin=0):
        i = self.info
        return frame_config(i["spp"], mode=mode, sample_begin=sample_begin, seed=i["seed"],
                            randomization=i["randomization"])

    def __del__(self):
        if getattr(self, "_h", None):  # the exit hook has released every handle before module teardown
            _release(self)


def probe_read_bandwidth(bytes_, passes, device=0):
    """Device read bandwidth in GB/s over a `bytes_` buffer read `passes`
    times (rp_probe_read_bandwidth)."""
    out = f64()
    _check("rp_probe_read_bandwidth", _lib.rp_probe_read_bandwidth(device, bytes_, passes, C.byref(out)))
    return out.value


def write_exr(path, xyz, exposure=1.0, half=False, metadata=None):
    """Writes a (H, W, 3) XYZ film as a linear-sRGB OpenEXR file (rp_write_exr)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float32)
    h, w, _ = xyz.shape
    meta = dict(metadata or {})
    keys = (C.c_char_p * max(len(meta), 1))(*[k.encode() for k in meta])
    vals = (C.c_char_p * max(len(meta), 1))(*[str(v).encode() for v in meta.values()])
    _check("rp_write_exr", _lib.rp_write_exr(str(path).encode(), w, h, xyz.ctypes.data_as(C.POINTER(f32)),
                                              float(exposure), 1 if half else 0, keys, vals, len(meta)))


class Batch:
    """Uneven multi-scene work pool (rp_batch_*): one device-resident pool for all items."""

    def __init__(self, items):
        self.items = list(items)  # (Scene, spp, seed)
        arr = (BatchItem * len(self.items))()
        for i, (scene, spp, seed) in enumerate(self.items):
            arr[i].scene, arr[i].spp, arr[i].seed = scene.handle, spp, seed
        h = C.c_void_p()
        _check("rp_batch_create", _lib.rp_batch_create(arr, len(self.items), C.byref(h)))
        self._h = h
        _track(self, "rp_batch_destroy")

    def __del__(self):
        if getattr(self, "_h", None):  # the exit hook has released every handle before module teardown
            _release(self)

    @staticmethod
    def config(scheduler=SCHED_WAVEFRONT, capacity=0, group_size=0, sort_bits=0, use_graph=1,
               randomization=RANDOMIZE_FASTOWEN, tail_paths=0, flags=0):
        c = BatchConfig()
        c.scheduler, c.capacity, c.group_size, c.sort_bits = scheduler, capacity, group_size, sort_bits
        c.use_graph, c.randomization, c.tail_paths, c.flags = use_graph, randomization, tail_paths, flags
        return c

    def render(self, cfg, films=True):
        outs = [np.zeros((s.height, s.width, 3), np.float32) for s, _, _ in self.items] if films else None
        ptrs = (C.c_void_p * len(self.items))()
        if films:
            for i, o in enumerate(outs):
                ptrs[i] = o.ctypes.data
        st = BatchStats()
        _check("rp_batch_render", _lib.rp_batch_render(self._h, C.byref(cfg), ptrs, C.byref(st)))
        return outs, {f: getattr(st, f) for f, _ in BatchStats._fields_ if f != "pad"}

    def autotune(self, randomization=RANDOMIZE_FASTOWEN):
        c = BatchConfig()
        c.randomization = randomization
        sec = f64()
        _check("rp_autotune", _lib.rp_autotune(self._h, C.byref(c), C.byref(sec)))
        return c, sec.value


def config_dict(cfg):
    return {f: getattr(cfg, f) for f, _ in cfg._fields_ if f != "pad"}


def params_array(params):
    arr = (Param * max(len(params), 1))()
    for i, (kind, target, b, bins) in enumerate(params):
        arr[i].kind, arr[i].target, arr[i].bin, arr[i].bins = kind, target, b, bins
    return arr


def inverse_params(synth, bins=16, sigma_bins=8):
    """The 58 parameters of the inverse-rendering Cornell box (RP_SYNTH_INVERSE):
    reflectance bins of the three wall materials, roughness of the two box
    conductors, sigma_s bins of the fog."""
    d = synth.d
    params = []
    for m in range(3):
        params += [(PARAM_REFLECTANCE_BIN, m, b, bins) for b in range(bins)]
    params += [(PARAM_ROUGHNESS, 3, 0, 0), (PARAM_ROUGHNESS, 4, 0, 0)]
    params += [(PARAM_SIGMA_S_BIN, 0, b, sigma_bins) for b in range(sigma_bins)]
    values = []
    spectra = np.ctypeslib.as_array(d.spectra, shape=(d.spectra_count,))
    for m in range(3):
        t = d.materials[m].reflectance.table
        for b in range(bins):
            i = next(i for i in range(471) if i * bins // 471 == b)
            values.append(float(spectra[t + i]))
    values += [d.materials[3].roughness_u, d.materials[4].roughness_u]
    t = d.media[0].sigma_s.table
    for b in range(sigma_bins):
        i = next(i for i in range(471) if i * sigma_bins // 471 == b)
        values.append(float(spectra[t + i]))
    return params, np.array(values, np.float32)
