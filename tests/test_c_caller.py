"""A compiled C11 caller of the C ABI (tests/c_caller/abi_caller.c).

CPU: the public headers compile as strict C11 (-Wall -Wextra -pedantic
-Werror) and the program links against the library.  GPU: the program runs,
its request refusals return RP_INVALID, and its record and film checksums
equal the same calls made through ctypes.
"""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "c_caller" / "abi_caller.c"
LIBDIR = ROOT / "rypath_amd"


def build_caller(out_dir):
    exe = Path(out_dir) / "abi_caller"
    cmd = ["cc", "-std=c11", "-Wall", "-Wextra", "-pedantic", "-Werror", "-O1", "-I", str(ROOT / "include"), str(SRC),
           "-o", str(exe), "-L", str(LIBDIR), "-lrypath_amd", f"-Wl,-rpath,{LIBDIR}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_c_caller_compiles_and_links(tmp_path):
    exe = build_caller(tmp_path)
    assert exe.exists()


@pytest.mark.gpu
def test_c_caller_matches_ctypes(rp, gpu_available, tmp_path):
    exe = build_caller(tmp_path)
    width, spp = 32, 4
    print(f"[c caller] {width}x{width}, {spp} spp", flush=True)
    r = subprocess.run([str(exe), str(width), str(spp)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    f = r.stdout.split()
    got = dict(zip(f[0::2], f[1::2]))
    synth = rp.Synthetic(rp.SYNTH_CORNELL, seed=1, width=width, height=width)
    scene = rp.Scene(synth.desc)
    recs = scene.sample(rp.request(seed=1, pixel_count=width * width, sample_count=spp))
    xor = 0
    for v in recs["sobol_index"]:
        xor ^= int(v)
    assert int(got["records"]) == width * width * spp
    # the C program sums in order in double precision; so does sum() over Python floats
    assert float(got["xyz_y_sum"]) == sum(float(v) for v in recs["xyz"][:, 1])
    assert int(got["sobol_xor"]) == xor
    assert int(got["failed"]) == int((recs["status"] != 0).sum())
    img, st = scene.render_frame(rp.frame_config(spp, mode=rp.MODE_WAVEFRONT, seed=1))
    assert float(got["film_sum"]) == sum(float(v) for v in img.reshape(-1))
    assert int(got["paths"]) == st["paths"]
