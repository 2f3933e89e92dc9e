"""ctypes access to the oracle (oracle/_build/liboracle.so).

Test infrastructure: the oracle is the CPU restatement of the reference
transport used as the parity checker and as the timed CPU baseline.
"""
import contextlib
import ctypes as C
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "oracle" / "_build" / "liboracle.so"
DATA = ROOT / "rypath_amd" / "data"


class Oracle:
    def __init__(self):
        if not LIB.exists():
            raise FileNotFoundError(f"{LIB} missing: run `make -C oracle`")
        import rypath_amd as rp  # for shared ctypes layouts

        self.rp = rp
        L = C.CDLL(str(LIB))
        self.L = L
        L.or_init.argtypes = [C.c_char_p, C.c_char_p]
        rc = L.or_init(str(DATA / "sobol_tables.bin").encode(), str(DATA / "cie_xyz_1931_2deg.csv").encode())
        if rc != 0:
            raise RuntimeError("oracle table load failed")
        u32, u64, f32 = C.c_uint32, C.c_uint64, C.c_float
        sig = {
            "or_sobol_index": (u64, [u32, u32, u32, u64]),
            "or_sobol_word": (u32, [u64, u32, u32, u32]),
            "or_dimension_hash": (u32, [u32, u32]),
            "or_sobol_uniform": (f32, [u32]),
            "or_camera_uniform": (f32, [u32, u32, u32]),
            "or_philox": (None, [C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)]),
            "or_philox_u01": (f32, [u32]),
            "or_sample_wavelengths": (None, [f32, u32, C.POINTER(f32), C.POINTER(f32)]),
            "or_cie_row": (None, [u32, C.POINTER(f32)]),
            "or_sincos_turns": (None, [f32, C.POINTER(f32), C.POINTER(f32)]),
            "or_exp": (f32, [f32]),
            "or_log": (f32, [f32]),
            "or_bsdf_eval": (None, [C.POINTER(rp.Material), C.POINTER(f32), f32, C.POINTER(f32), C.POINTER(f32),
                                    C.POINTER(f32), C.POINTER(f32)]),
            "or_bsdf_sample": (C.c_int, [C.POINTER(rp.Material), C.POINTER(f32), f32, C.POINTER(f32), f32, f32, f32,
                                         C.POINTER(f32), C.POINTER(f32), C.POINTER(C.c_int)]),
            "or_fresnel_complex": (f32, [f32, f32, f32]),
            "or_fresnel_dielectric": (f32, [f32, f32]),
            "or_hg": (f32, [f32, f32]),
            "or_hg_sample": (None, [C.POINTER(f32), f32, f32, f32, C.POINTER(f32), C.POINTER(f32)]),
            "or_intersect_triangle": (C.c_int, [C.POINTER(f32), C.POINTER(f32), f32, C.POINTER(f32), C.POINTER(f32)]),
            "or_scene_create": (C.c_void_p, [C.POINTER(rp.SceneDesc)]),
            "or_scene_destroy": (None, [C.c_void_p]),
            "or_scene_set_sampling": (None, [C.c_void_p, C.c_void_p]),
            "or_light_count": (u32, [C.c_void_p]),
            "or_light_table": (None, [C.c_void_p] + [C.c_void_p] * 6),
            "or_light_node_count": (u32, [C.c_void_p]),
            "or_light_inputs": (None, [C.c_void_p] + [C.c_void_p] * 9),
            "or_light_alias": (None, [C.c_void_p, C.c_void_p]),
            "or_light_nodes": (None, [C.c_void_p] + [C.c_void_p] * 10),
            "or_spectrum_power_table": (C.c_double, [C.POINTER(f32)]),
            "or_patch_rectangle": (C.c_int, [C.POINTER(C.c_double)]),
            "or_normal_cone": (None, [C.POINTER(C.c_double), u32, C.POINTER(f32), C.POINTER(f32)]),
            "or_light_select": (u32, [C.c_void_p, C.POINTER(f32), C.POINTER(f32), f32, C.POINTER(f32)]),
            "or_light_pmf": (f32, [C.c_void_p, C.POINTER(f32), C.POINTER(f32), u32]),
            "or_light_sample": (u32, [C.c_void_p, C.POINTER(f32), C.POINTER(f32), f32, f32, f32, C.POINTER(f32),
                                      C.POINTER(f32), C.POINTER(C.c_int)]),
            "or_area_light_pdf": (f32, [C.c_void_p, u32, C.POINTER(f32), C.POINTER(f32), C.POINTER(f32),
                                        C.POINTER(f32)]),
            "or_atan2": (f32, [f32, f32]),
            "or_asin": (f32, [f32]),
            "or_trace_closest": (None, [C.c_void_p, C.c_void_p, u64, C.c_void_p, u32]),
            "or_sample": (C.c_int, [C.c_void_p, C.POINTER(rp.Request), C.c_void_p, u32]),
            "or_render_frame": (C.c_int, [C.c_void_p, u64, u32, u64, u32, C.c_void_p, u32, C.POINTER(u64)]),
            "or_render_frame_flags": (C.c_int, [C.c_void_p, u64, u32, u64, u32, C.c_void_p, u32, C.POINTER(u64),
                                                u32]),
            "or_render_frame_f64": (C.c_int, [C.c_void_p, u64, u32, u64, u32, C.c_void_p, u32, C.POINTER(u64),
                                              u32]),
            "or_sample_scalar": (C.c_int, [C.c_void_p, C.POINTER(rp.Request), C.c_void_p, u32]),
            "or_set_roulette": (None, [C.c_int]),
            "or_plant_slip": (None, [C.c_int]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args

    # --- test switches
    @contextlib.contextmanager
    def roulette(self, enabled):
        """Run the block with the oracle's roulette on or off."""
        self.L.or_set_roulette(1 if enabled else 0)
        try:
            yield
        finally:
            self.L.or_set_roulette(1)

    @contextlib.contextmanager
    def planted_slip(self, slip_id):
        """Run the block with one named estimator mistake planted in the oracle
        (the SLIP_* values of or_transport.cpp)."""
        self.L.or_plant_slip(slip_id)
        try:
            yield
        finally:
            self.L.or_plant_slip(0)

    # --- sampling probes
    def sobol_index(self, m, x, y, s):
        return self.L.or_sobol_index(m, x, y, s)

    def sobol_word(self, index, dim, mode, seed):
        return self.L.or_sobol_word(index, dim, mode, seed)

    def philox(self, ctr, key):
        c = (C.c_uint32 * 4)(*ctr)
        k = (C.c_uint32 * 2)(*key)
        o = (C.c_uint32 * 4)()
        self.L.or_philox(c, k, o)
        return list(o)

    def wavelengths(self, u, mode=0):
        lam, pdf = (C.c_float * 4)(), (C.c_float * 4)()
        self.L.or_sample_wavelengths(u, mode, lam, pdf)
        return list(lam), list(pdf)

    # --- scenes
    def scene(self, desc):
        ptr = desc if isinstance(desc, C._Pointer) else C.pointer(desc)
        h = self.L.or_scene_create(ptr)
        return OracleScene(self, h, ptr.contents.camera.width, ptr.contents.camera.height)


class OracleScene:
    def __init__(self, orc, h, w, hgt):
        self.orc, self.h, self.width, self.height = orc, h, w, hgt

    def __del__(self):
        if getattr(self, "h", None):
            self.orc.L.or_scene_destroy(self.h)
            self.h = None

    def set_sampling(self, other):
        """Test switch: `other` (an oracle scene of the same geometry, or None)
        drives this scene's sampling decisions and densities."""
        self._sampling = other  # keeps it alive while referenced
        self.orc.L.or_scene_set_sampling(self.h, other.h if other is not None else None)

    def light_table(self):
        n = self.orc.L.or_light_count(self.h)
        arrs = [np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.zeros(n, np.float32), np.zeros(n, np.float32),
                np.zeros(n, np.float32), np.zeros(n, np.uint32)]
        self.orc.L.or_light_table(self.h, *[a.ctypes.data_as(C.c_void_p) for a in arrs])
        return dict(zip(["kind", "index", "pmf", "threshold", "inv_area", "leaf"], arrs))

    def light_node_count(self):
        return self.orc.L.or_light_node_count(self.h)

    def light_inputs(self):
        """Per light, what the tables were built from (LightSamplingInput)."""
        n = self.orc.L.or_light_count(self.h)
        f, d = np.float32, np.float64
        out = {"power": np.zeros(n, d), "bounds_power": np.zeros(n, d), "low": np.zeros((n, 3), f),
               "high": np.zeros((n, 3), f), "dir": np.zeros((n, 3), f), "cos_outer": np.zeros(n, f),
               "cos_emission": np.zeros(n, f), "two_sided": np.zeros(n, np.uint8), "infinite": np.zeros(n, np.uint8)}
        self.orc.L.or_light_inputs(self.h, *[a.ctypes.data_as(C.c_void_p) for a in out.values()])
        return out

    def light_alias(self):
        a = np.zeros(self.orc.L.or_light_count(self.h), np.uint32)
        self.orc.L.or_light_alias(self.h, a.ctypes.data_as(C.c_void_p))
        return a

    def light_nodes(self):
        m = self.orc.L.or_light_node_count(self.h)
        f, u = np.float32, np.uint32
        out = {"low": np.zeros((m, 3), f), "high": np.zeros((m, 3), f), "dir": np.zeros((m, 3), f),
               "power": np.zeros(m, f), "cos_outer": np.zeros(m, f), "cos_emission": np.zeros(m, f),
               "child": np.zeros(m, u), "parent": np.zeros(m, u), "light": np.zeros(m, u), "flags": np.zeros(m, u)}
        self.orc.L.or_light_nodes(self.h, *[a.ctypes.data_as(C.c_void_p) for a in out.values()])
        return out

    @staticmethod
    def _v(x):
        return (C.c_float * 3)(*[float(v) for v in x])

    def light_select(self, p, n, u):
        pmf = C.c_float()
        i = self.orc.L.or_light_select(self.h, self._v(p), self._v(n), u, C.byref(pmf))
        return i, pmf.value

    def light_pmf(self, p, n, light):
        return self.orc.L.or_light_pmf(self.h, self._v(p), self._v(n), light)

    def light_sample(self, p, n, ul, u0, u1):
        out, pdf, delta = (C.c_float * 3)(), C.c_float(), C.c_int()
        i = self.orc.L.or_light_sample(self.h, self._v(p), self._v(n), ul, u0, u1, out, C.byref(pdf),
                                       C.byref(delta))
        return i, np.array(out[:], np.float64), pdf.value, bool(delta.value)

    def area_light_pdf(self, light, p, n, q, ng):
        return self.orc.L.or_area_light_pdf(self.h, light, self._v(p), self._v(n), self._v(q), self._v(ng))

    def trace_closest(self, rays, threads=0):
        rp = self.orc.rp
        rays = np.ascontiguousarray(rays, dtype=rp.RAY_DTYPE)
        hits = np.zeros(len(rays), dtype=rp.HIT_DTYPE)
        self.orc.L.or_trace_closest(self.h, rays.ctypes.data_as(C.c_void_p), len(rays),
                                    hits.ctypes.data_as(C.c_void_p), threads)
        return hits

    def sample(self, req, threads=0):
        rp = self.orc.rp
        n = req.pixel_count * req.sample_count
        out = np.zeros(n, dtype=rp.RECORD_DTYPE)
        self.orc.L.or_sample(self.h, C.byref(req), out.ctypes.data_as(C.c_void_p), threads)
        return out

    def render_frame_f64(self, spp, seed=1, randomization=2, sample_begin=0, threads=0, flags=0):
        """Film means in binary64 (or_render_frame_f64)."""
        out = np.zeros((self.height, self.width, 3), dtype=np.float64)
        rays = C.c_uint64()
        self.orc.L.or_render_frame_f64(self.h, seed, randomization, sample_begin, spp,
                                       out.ctypes.data_as(C.c_void_p), threads, C.byref(rays), flags)
        return out, rays.value

    def render_frame(self, spp, seed=1, randomization=2, sample_begin=0, threads=0, flags=0):
        out = np.zeros((self.height, self.width, 3), dtype=np.float32)
        rays = C.c_uint64()
        self.orc.L.or_render_frame_flags(self.h, seed, randomization, sample_begin, spp,
                                         out.ctypes.data_as(C.c_void_p), threads, C.byref(rays), flags)
        return out, rays.value

    def sample_scalar(self, req, threads=0):
        rp = self.orc.rp
        n = req.pixel_count * req.sample_count
        out = np.zeros(n, dtype=rp.RENDER_SAMPLE_DTYPE)
        self.orc.L.or_sample_scalar(self.h, C.byref(req), out.ctypes.data_as(C.c_void_p), threads)
        return out


def sample(orc, scene, req):
    return scene.sample(req)


def oracle_gradient(orc, oscene, params, dl_dxyz, spp, seed=1, randomization=2, sample_begin=0, threads=0):
    import ctypes as C

    rp = orc.rp
    arr = rp.params_array(params)
    L = orc.L
    L.or_gradient.restype = C.c_int
    L.or_gradient.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, C.c_void_p, C.c_uint32,
                              C.c_void_p, C.c_void_p, C.c_uint32]
    dl = np.ascontiguousarray(dl_dxyz, np.float32).reshape(-1)
    out = np.zeros(len(params), np.float64)
    L.or_gradient(oscene.h, seed, randomization, sample_begin, spp, C.cast(arr, C.c_void_p), len(params),
                  dl.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), threads)
    return out
