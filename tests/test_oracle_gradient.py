"""Pins of the oracle's forward-mode derivative (or_gradient), which the GPU
path-replay adjoint is checked against bit-for-bit in the GPU tests:

* the analytic d log f / d roughness of the GGX conductor against central
  differences of the BSDF evaluation itself;
* with roulette switched off in the oracle (or_set_roulette, a test switch)
  the renderer's sampling no longer depends on reflectances, so central
  differences of the rendered loss with common random numbers pin the
  reflectance derivative to curvature accuracy;
* roughness and sigma_s move the sampled directions and free-flight
  distances, so their central differences are compared statistically over
  independent sample sets.
"""
import ctypes as C
import math

import numpy as np
import pytest

from tests import grad_util, oracle_lib
from tests.test_oracle_pins import bsdf_eval, fa, material, sph

W = 12


@pytest.fixture()
def no_roulette(orc):
    orc.L.or_set_roulette.argtypes = [C.c_int]
    orc.L.or_set_roulette(0)
    yield
    orc.L.or_set_roulette(1)


@pytest.mark.parametrize("remap,r", [(0, 0.15), (0, 0.5), (1, 0.09), (1, 0.3)])
def test_conductor_dlogf_matches_bsdf_differences(orc, rp, remap, r):
    orc.L.or_conductor_dlogf.restype = C.c_double
    orc.L.or_conductor_dlogf.argtypes = [C.POINTER(rp.Material), C.c_float, C.POINTER(C.c_float),
                                         C.POINTER(C.c_float)]
    rng = np.random.default_rng(int(r * 100) + remap)
    checked = 0
    for _ in range(40):
        wo = sph(rng.uniform(0.05, 1.4), rng.uniform(0, 2 * math.pi))
        wi = sph(rng.uniform(0.05, 1.4), rng.uniform(0, 2 * math.pi))
        m = material(rp, rp.MATERIAL_CONDUCTOR, rough=r, eta=0.3, k=2.5, remap=remap)
        analytic = orc.L.or_conductor_dlogf(C.byref(m), 550.0, fa(*wo), fa(*wi))
        h = 1e-3 * r
        lf = []
        for rr in (r + h, r - h):
            mm = material(rp, rp.MATERIAL_CONDUCTOR, rough=rr, eta=0.3, k=2.5, remap=remap)
            f, _ = bsdf_eval(orc, mm, wo, wi)
            lf.append(math.log(f))
        fd = (lf[0] - lf[1]) / (2 * h)
        assert abs(analytic - fd) <= 2e-3 * max(1.0, abs(fd)), (wo, wi, analytic, fd)
        checked += 1
    assert checked == 40


def loss_and_grad(orc, rp, dp, synth, params, w, v, h, spp, seed):
    losses = []
    for sgn in (+1, -1):
        dp.apply((dp.values + sgn * h * v).astype(np.float32))
        s = orc.scene(synth.desc)
        img, _ = s.render_frame(spp, seed=seed)
        del s
        losses.append(float((img.astype(np.float64) * w).sum()))
    dp.restore()
    s = orc.scene(synth.desc)
    g = oracle_lib.oracle_gradient(orc, s, params, w, spp=spp, seed=seed)
    return (losses[0] - losses[1]) / (2 * h), float(g @ v)


def inverse_scene(rp):
    synth = rp.Synthetic(rp.SYNTH_INVERSE, seed=2, width=W, height=W)
    params, values = rp.inverse_params(synth)
    return synth, params, values, grad_util.DescParams(rp, synth, params, values)


def test_reflectance_derivative_exact_without_roulette(orc, rp, no_roulette):
    synth, params, values, dp = inverse_scene(rp)
    w = np.random.default_rng(1).uniform(0.5, 1.5, (W, W, 3)).astype(np.float32)
    v = grad_util.direction(values, "walls", 3)
    for h in (1e-2, 1e-3):
        fd, ad = loss_and_grad(orc, rp, dp, synth, params, w, v, h, 64, 5)
        assert abs(fd - ad) <= 2e-3 * abs(ad), (h, fd, ad)


@pytest.mark.parametrize("group", ["walls", "roughness", "sigma_s"])
def test_forward_mode_matches_frozen_sampling_differences_at_1e_3(orc, rp, no_roulette, group):
    """At the stated step h = 1e-3 (every parameter of the group moved by
    0.1 % of its value, along the signs of an independent pilot gradient),
    central differences of the forward render with the sampling frozen at
    the base parameters (every path fixed, see
    grad_util.frozen_sampling_differences) equal the forward-mode derivative
    of the detached estimator to within 2e-3 relative, for each group: the
    2 % gradient target met with a margin of ten."""
    synth, params, values, dp = inverse_scene(rp)
    w = np.random.default_rng(1).uniform(0.5, 1.5, (W, W, 3)).astype(np.float32)
    pilot = oracle_lib.oracle_gradient(orc, orc.scene(synth.desc), params, w, spp=64, seed=3)
    v = grad_util.steepest_relative_direction(values, group, pilot)
    fd, base = grad_util.frozen_sampling_differences(orc, dp, synth, w, v, 1e-3, 256, 7)
    ad = float(oracle_lib.oracle_gradient(orc, base, params, w, spp=256, seed=7) @ v)
    print(f"[{group}] forward mode {ad:.6e} frozen-sampling central difference {fd:.6e} "
          f"rel {abs(fd - ad) / abs(ad):.2e}")
    assert ad != 0 and abs(fd - ad) <= 2e-3 * abs(ad)


def test_frozen_sampling_differences_detect_a_wrong_derivative(orc, rp, no_roulette):
    """The frozen-sampling quotient is sensitive to the derivative itself: a
    direction that moves roughness gives a quotient far from the forward-mode
    derivative along another group."""
    synth, params, values, dp = inverse_scene(rp)
    w = np.random.default_rng(1).uniform(0.5, 1.5, (W, W, 3)).astype(np.float32)
    v = grad_util.direction(values, "roughness", 4)
    fd, base = grad_util.frozen_sampling_differences(orc, dp, synth, w, v, 1e-3, 64, 7)
    g = oracle_lib.oracle_gradient(orc, base, params, w, spp=64, seed=7)
    wrong = g.copy()
    wrong[48:50] *= 1.05  # a 5 % error in d log f / d roughness
    assert abs(fd - float(g @ v)) <= 2e-3 * abs(float(g @ v))
    assert abs(fd - float(wrong @ v)) > 2e-2 * abs(float(g @ v))


@pytest.mark.parametrize("group", ["roughness", "sigma_s"])
def test_sampling_dependent_derivatives_statistically(orc, rp, no_roulette, group):
    synth, params, values, dp = inverse_scene(rp)
    w = np.random.default_rng(1).uniform(0.5, 1.5, (W, W, 3)).astype(np.float32)
    v = grad_util.direction(values, group, 4)
    fds, ads = [], []
    # a larger step lowers the variance that sample-moving discontinuities
    # add (about 1 / (spp * h)); the loss is smooth in expectation
    h = 0.02 if group == "roughness" else 0.1
    for seed in range(6):
        fd, ad = loss_and_grad(orc, rp, dp, synth, params, w, v, h, 1024, 11 + seed)
        fds.append(fd)
        ads.append(ad)
    fm, fse = grad_util.mean_se(fds)
    am, ase = grad_util.mean_se(ads)
    print(f"[{group}] adjoint {am:.3f} +- {ase:.3f}  central differences {fm:.3f} +- {fse:.3f}")
    assert abs(fm - am) <= 4 * math.hypot(fse, ase) + 0.03 * abs(am)
