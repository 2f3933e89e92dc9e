"""Known-answer transport checks on the HIP path (through the C ABI).

The same closed-form and quadrature scenes as tests/test_oracle_transport.py
(furnace, environment-lit floor, spot over a floor, absorbing slab, single
scattering for g in {-0.7, 0, 0.3, 0.7}) rendered by the device renderer and
held to the same gates, plus the open box against the independent float64
BSDF-only estimator.  The device always runs with roulette on.  Each case's
records are also compared with the oracle's on the same request: the two
estimators must agree path by path, so a device-only mistake cannot hide
behind the statistical gate.
"""
import numpy as np
import pytest

from tests import test_oracle_transport as ot
from tests import transport_cases as tc
from tests.parity_util import AGREE_FRACTION, agreement

pytestmark = pytest.mark.gpu

GPU_CASES = [name for name, (_, _, roulette) in ot.CASES.items() if roulette]


def _gpu_records(rp, case, spp):
    w, h = case.builder.width, case.builder.height
    req = rp.request(seed=1, pixel_begin=0, pixel_count=w * h, sample_count=spp)
    print(f"[transport] {case.label}: {w}x{h} px, {spp} spp", flush=True)
    return rp.Scene(case.desc).sample(req), req


@pytest.mark.parametrize("name", GPU_CASES)
def test_device_matches_known_answer(rp, orc, gpu_available, name):
    case = ot._case(name)
    spp = ot.CASES[name][1]
    gpu, req = _gpu_records(rp, case, spp)
    assert (gpu["status"] == 0).all()
    ok, msg = ot.gate(gpu, case)
    print(f"[transport] {msg}", flush=True)
    assert ok, msg
    cpu = orc.scene(case.desc).sample(req)
    assert np.array_equal(gpu["primary_triangle"], cpu["primary_triangle"])
    assert agreement(gpu["xyz"], cpu["xyz"]) >= AGREE_FRACTION


def test_device_matches_fp64_bsdf_only_estimator(rp, gpu_available):
    from tests import bruteforce

    case = ot._case_open_box()
    b = case.builder
    gpu, _ = _gpu_records(rp, case, 256)
    assert (gpu["status"] == 0).all()
    assert int(gpu["vertices"].max()) >= 5
    gm, gse = ot._region_stats(tc.path_xyz(gpu), gpu["pixel"].astype(np.int64), b.width, b.height, 8)
    bx, bpix, deepest = bruteforce.render(b, case.desc, 1024, seed=0x5eed)
    assert deepest >= 4
    bm, bse = ot._region_stats(bx, bpix, b.width, b.height, 8)
    comb = np.sqrt(gse ** 2 + bse ** 2)
    tol = 5.0 * comb + 2e-5 * np.maximum(1.0, np.abs(bm))
    print(f"[transport] open box: device {gm.tolist()} brute force {bm.tolist()} "
          f"|d|/se {(np.abs(gm - bm) / comb).max():.2f}", flush=True)
    assert np.all(np.abs(gm - bm) <= tol)
