"""Known-answer checks of the oracle's composed hero transport (CPU).

The expectations come from closed forms and float64 quadrature in
tests/transport_cases.py, which shares no code with the oracle or the device:
a white furnace (sum_k rho^k Le) in a cube and in a long tunnel tiled into
unit quad lights (far tiles seen at grazing angles), an environment-lit
floor (rho E), a spot
light over a floor (inverse square, cosine and cone smoothstep, per pixel),
an absorbing slab (Beer-Lambert, per nm) and single scattering in a
homogeneous medium under a spot light (Henyey-Greenstein, per pixel) for
g in {-0.7, 0, 0.3, 0.7}.  The gates are the reference's analytic gates
(docs/rendering.md:179-192: six standard errors plus 2e-5 max(1, |expected|)).

Each planted slip (or_plant_slip, a test switch of the oracle) reproduces one
plausible estimator mistake -- a dropped MIS weight, a missing light pmf,
uncompensated roulette, a flipped phase-function sign, a missing
transmittance, a wrong environment pdf, a missing falloff or cosine, hero
transmittance on every lane, a light-BVH pmf without the receiver normal, an
unwarped quad proposal density -- and the slip tests require that at least one
known-answer gate fails with it.  Every result is deterministic (fixed Sobol
seed), so these are exact regression checks, not flaky statistics.
"""
import pytest

import rypath_amd as rp
from tests import transport_cases as tc

# (case factory, samples per pixel, roulette on)
CASES = {
    "furnace_d1": (lambda: tc.furnace(1), 512, True),
    "furnace_d3": (lambda: tc.furnace(3), 512, True),
    "furnace_d6": (lambda: tc.furnace(6), 512, True),
    "furnace_d6_no_roulette": (lambda: tc.furnace(6), 512, False),
    "furnace_tunnel_d1": (lambda: tc.furnace_tunnel(1), 2048, True),
    "furnace_tunnel_d3": (lambda: tc.furnace_tunnel(3), 2048, True),
    "env_floor": (lambda: tc.env_floor(), 512, True),
    "spot_plane": (lambda: tc.spot_plane(False), 1024, True),
    "spot_plane_env": (lambda: tc.spot_plane(True), 1024, True),
    "slab": (lambda: tc.slab(), 2048, True),
    "single_scatter_g-0.7": (lambda: tc.single_scatter(-0.7), 4096, True),
    "single_scatter_g0": (lambda: tc.single_scatter(0.0), 4096, True),
    "single_scatter_g0.3": (lambda: tc.single_scatter(0.3), 4096, True),
    "single_scatter_g0.7": (lambda: tc.single_scatter(0.7), 4096, True),
}

# slip id -> (name, cases that must contain at least one failing gate)
SLIPS = {
    1: ("emission reached by BSDF sampling without its MIS weight", ["furnace_d1", "furnace_d3", "furnace_d6"]),
    2: ("area-light pdf without the selection pmf", ["furnace_d3", "furnace_d6"]),
    3: ("roulette survivors not compensated", ["furnace_d6", "env_floor"]),
    4: ("phase function evaluated at -cos", ["single_scatter_g-0.7", "single_scatter_g0.7"]),
    5: ("medium NEE without transmittance", ["single_scatter_g0", "single_scatter_g0.3"]),
    6: ("environment MIS pdf over the hemisphere", ["env_floor", "spot_plane_env"]),
    7: ("spot light without cone falloff", ["spot_plane", "spot_plane_env"]),
    8: ("surface NEE without the receiver cosine", ["spot_plane", "env_floor", "furnace_d3"]),
    9: ("hero sigma_t for every lane's transmittance", ["slab"]),
    10: ("continuation weight without |cos|", ["furnace_d3", "env_floor"]),
    11: ("light-BVH pmf at emission hits without the receiver normal", ["furnace_tunnel_d1", "furnace_tunnel_d3"]),
    12: ("quad proposal density without its cosine warp", ["furnace_tunnel_d1", "furnace_d1", "furnace_d3"]),
}

_built = {}


def _case(name):
    if name not in _built:
        _built[name] = CASES[name][0]()
    return _built[name]


def _oracle_records(orc, name, spp):
    case = _case(name)
    w, h = case.builder.width, case.builder.height
    req = rp.request(seed=1, pixel_begin=0, pixel_count=w * h, sample_count=spp)
    return orc.scene(case.desc).sample(req)


def gate(records, case):
    if case.per_pixel:
        return tc.pixel_check(records, case)
    return tc.mean_check(tc.path_xyz(records), case.expected, label=case.label)


def run_case(orc, name):
    _, spp, roulette = CASES[name]
    with orc.roulette(roulette):
        recs = _oracle_records(orc, name, spp)
    assert (recs["status"] == 0).all(), f"{name}: failed paths {sorted(set(recs['status']))}"
    return gate(recs, _case(name))


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_matches_known_answer(orc, name):
    ok, msg = run_case(orc, name)
    assert ok, msg


@pytest.mark.parametrize("slip", list(SLIPS))
def test_planted_slip_fails_a_known_answer_gate(orc, slip):
    what, names = SLIPS[slip]
    results = {}
    with orc.planted_slip(slip):
        for name in names:
            _, spp, roulette = CASES[name]
            with orc.roulette(roulette):
                recs = _oracle_records(orc, name, spp)
            results[name] = gate(recs, _case(name))
    assert not all(ok for ok, _ in results.values()), (
        f"slip {slip} ({what}) passed every gate: " + "; ".join(m for _, m in results.values()))


def test_furnace_expectation_is_the_neumann_series():
    """The furnace expectation at depth D is the D-term Neumann series; with
    rho constant it is Le (1 - rho^(D+1)) / (1 - rho) per wavelength."""
    import numpy as np

    b_case = tc.furnace(4, rho=(0.5, 0.5), le=(2.0, 2.0))
    cie_int = tc.xyz_integral(lambda lam: np.ones_like(lam))
    expected = cie_int * 2.0 * (1 - 0.5 ** 5) / (1 - 0.5)
    assert np.allclose(b_case.expected, expected, rtol=1e-12)


def _region_stats(xyz, pix, w, h, region):
    """Mean and standard error of per-path XYZ per region x region block of pixels."""
    import numpy as np

    rx, ry = (pix % w) // region, (pix // w) // region
    key = ry * (w // region) + rx
    nreg = (w // region) * (h // region)
    mean, se = np.zeros((nreg, 3)), np.zeros((nreg, 3))
    for k in range(nreg):
        v = xyz[key == k]
        mean[k] = v.mean(0)
        se[k] = v.std(0, ddof=1) / np.sqrt(len(v))
    return mean, se


def _compare_with_bruteforce(orc, case, spp, brute_ppp, region=8, nsig=5.0):
    import numpy as np

    from tests import bruteforce

    b = case.builder
    recs = _oracle_records_for(orc, case, spp)
    assert (recs["status"] == 0).all()
    om, ose = _region_stats(tc.path_xyz(recs), recs["pixel"].astype(np.int64), b.width, b.height, region)
    bx, bpix, deepest = bruteforce.render(b, case.desc, brute_ppp, seed=0x5eed)
    bm, bse = _region_stats(bx, bpix, b.width, b.height, region)
    tol = nsig * np.sqrt(ose ** 2 + bse ** 2) + 2e-5 * np.maximum(1.0, np.abs(bm))
    ok = bool(np.all(np.abs(om - bm) <= tol))
    return ok, deepest, recs, (f"oracle {om.tolist()} brute force {bm.tolist()} "
                               f"|d|/se {(np.abs(om - bm) / np.sqrt(ose ** 2 + bse ** 2)).max():.2f}")


def _oracle_records_for(orc, case, spp):
    w, h = case.builder.width, case.builder.height
    req = rp.request(seed=1, pixel_begin=0, pixel_count=w * h, sample_count=spp)
    return orc.scene(case.desc).sample(req)


def test_oracle_matches_fp64_bsdf_only_estimator(orc):
    """Multi-bounce transport against the independent float64 BSDF-only
    estimator (the reference's multibounce gate, docs/rendering.md:184-186:
    region means within five combined standard errors, a path reaching a
    fourth hit)."""
    case = _case_open_box()
    ok, deepest, recs, msg = _compare_with_bruteforce(orc, case, spp=256, brute_ppp=1024)
    assert deepest >= 4 and int(recs["vertices"].max()) >= 5
    assert ok, msg


@pytest.mark.parametrize("slip", [6, 10])
def test_planted_slip_fails_the_bruteforce_gate(orc, slip):
    case = _case_open_box()
    with orc.planted_slip(slip):
        ok, _, _, msg = _compare_with_bruteforce(orc, case, spp=256, brute_ppp=1024)
    assert not ok, f"slip {slip} ({SLIPS[slip][0]}) passed: {msg}"


def _case_open_box():
    if "open_box" not in _built:
        _built["open_box"] = tc.open_box()
    return _built["open_box"]
