"""ctypes access to the reference's own host code (oracle/_ref/libref.so).

Test infrastructure: `make -C oracle ref` builds it from the reference's
sources where they lie (see oracle/ref_driver.cpp).  Where those sources are
absent the library is not built and the tests that use it skip.
"""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "oracle" / "_ref" / "libref.so"


class Ref:
    def __init__(self):
        L = C.CDLL(str(LIB))
        self.L = L
        u8p, u32, u64, f32, f64 = C.POINTER(C.c_uint8), C.c_uint32, C.c_uint64, C.c_float, C.c_double
        fp, dp, up = C.POINTER(f32), C.POINTER(f64), C.POINTER(u32)
        sig = {
            "ref_sobol_index": (u64, [u32, u32, u32, u64]),
            "ref_sobol_word": (u32, [u64, u32, u32, u32]),
            "ref_dimension_hash": (u32, [u32, u32]),
            "ref_sobol_uniform": (f32, [u32]),
            "ref_camera_uniform": (f32, [u32, u32, u32]),
            "ref_hero_uniform": (None, [f32, fp, fp]),
            "ref_hero_visible": (None, [f32, fp, fp]),
            "ref_packet_visible": (None, [f32, fp, fp]),
            "ref_light_spectrum_power": (f64, [fp]),
            "ref_light_patch_rectangle": (C.c_int, [dp]),
            "ref_light_patch_area": (f64, [dp]),
            "ref_light_normal_cone": (None, [dp, u32, fp, fp]),
            "ref_light_positional_cone": (None, [fp, f32, f32, fp, fp, fp, C.POINTER(C.c_int)]),
            "ref_build_light_sampling": (C.c_int, [u32, dp, dp, fp, fp, fp, fp, fp, u8p, u8p, fp, fp, up, up,
                                                   fp, fp, fp, fp, fp, fp, up, up, up, up, C.c_char_p, C.c_size_t]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args

    @staticmethod
    def _p(a, t):
        return a.ctypes.data_as(C.POINTER(t))

    def wavelengths(self, fn, u):
        lam, pdf = (C.c_float * 4)(), (C.c_float * 4)()
        getattr(self.L, fn)(u, lam, pdf)
        return np.array(lam[:], np.float32), np.array(pdf[:], np.float32)

    def build_light_sampling(self, inputs):
        """The reference's alias table and light BVH for the planar inputs of
        OracleScene.light_inputs()."""
        n = len(inputs["power"])
        f, d, u = np.float32, np.float64, np.uint32
        alias = {k: np.zeros(n, t) for k, t in (("threshold", f), ("pmf", f), ("alias", u), ("leaf", u))}
        m = max(2 * n - 1, 1)
        nodes = {"low": np.zeros((m, 3), f), "high": np.zeros((m, 3), f), "dir": np.zeros((m, 3), f),
                 "power": np.zeros(m, f), "cos_outer": np.zeros(m, f), "cos_emission": np.zeros(m, f),
                 "child": np.zeros(m, u), "parent": np.zeros(m, u), "light": np.zeros(m, u), "flags": np.zeros(m, u)}
        err = C.create_string_buffer(256)
        P, fl, dl, ul, bl = self._p, C.c_float, C.c_double, C.c_uint32, C.c_uint8
        arr = {k: np.ascontiguousarray(v) for k, v in inputs.items()}
        count = self.L.ref_build_light_sampling(
            n, P(arr["power"], dl), P(arr["bounds_power"], dl), P(arr["low"], fl), P(arr["high"], fl),
            P(arr["dir"], fl), P(arr["cos_outer"], fl), P(arr["cos_emission"], fl), P(arr["two_sided"], bl),
            P(arr["infinite"], bl), P(alias["threshold"], fl), P(alias["pmf"], fl), P(alias["alias"], ul),
            P(alias["leaf"], ul), P(nodes["low"], fl), P(nodes["high"], fl), P(nodes["dir"], fl),
            P(nodes["power"], fl), P(nodes["cos_outer"], fl), P(nodes["cos_emission"], fl), P(nodes["child"], ul),
            P(nodes["parent"], ul), P(nodes["light"], ul), P(nodes["flags"], ul), err, 256)
        if count < 0:
            raise RuntimeError(err.value.decode())
        return alias, {k: v[:count] for k, v in nodes.items()}


_REF = None


def ref():
    """The loaded reference library, or a pytest skip where it is not built."""
    global _REF
    if _REF is None:
        if not LIB.exists():
            pytest.skip(f"{LIB.relative_to(ROOT)} not built (the reference's sources are absent)")
        _REF = Ref()
    return _REF
