"""Shared helpers for derivative checks of the inverse-rendering scene: write
parameter values into a scene description's tables (the host-side meaning of
rp_set_params) and estimate directional derivatives by central differences,
over several independent sample sets or with the sampling frozen."""
import numpy as np


class DescParams:
    """Applies parameter vectors to an owned scene description (Synthetic)."""

    def __init__(self, rp, synth, params, values):
        self.rp, self.synth, self.params, self.values = rp, synth, params, np.asarray(values)
        d = synth.d
        self.d = d
        self.pool = np.ctypeslib.as_array(d.spectra, shape=(d.spectra_count,))
        self.orig_pool = self.pool.copy()
        self.orig_rough = {t: (d.materials[t].roughness_u, d.materials[t].roughness_v)
                           for k, t, _, _ in params if k == rp.PARAM_ROUGHNESS}

    def apply(self, vals):
        rp, d = self.rp, self.d
        self.pool[:] = self.orig_pool
        for t, (u, v) in self.orig_rough.items():
            d.materials[t].roughness_u, d.materials[t].roughness_v = u, v
        for (kind, target, b, bins), v in zip(self.params, vals):
            if kind == rp.PARAM_ROUGHNESS:
                d.materials[target].roughness_u = d.materials[target].roughness_v = float(v)
                continue
            spec = d.materials[target].reflectance if kind == rp.PARAM_REFLECTANCE_BIN else d.media[target].sigma_s
            for i in range(471):
                if i * bins // 471 == b:
                    self.pool[spec.table + i] = v

    def restore(self):
        self.apply(self.values)


GROUPS = {"walls": slice(0, 48), "roughness": slice(48, 50), "sigma_s": slice(50, 58)}


def direction(values, group, seed):
    rng = np.random.default_rng(seed)
    sl = GROUPS[group]
    v = np.zeros(len(values))
    v[sl] = rng.normal(0, 1, sl.stop - sl.start) * np.abs(values[sl])
    return v


def mean_se(x):
    x = np.asarray(x, np.float64)
    return float(x.mean()), float(x.std(ddof=1) / np.sqrt(len(x)))


def frozen_sampling_differences(orc, dp, synth, w, v, h, spp, seed):
    """Central difference (L(theta + h v) - L(theta - h v)) / 2h of the loss
    L = sum(w * film) of the oracle's forward render (binary64 film means),
    with every sampling
    decision and density (BSDF and free-flight sampling, hero-lane pdfs, MIS
    weights) taken at theta (or_scene_set_sampling) and only the evaluated
    values moved.  Every path stays the same, so the quotient has no
    discontinuity noise, and since the estimator's expectation is the
    rendered integral at every parameter value, its derivative is the
    derivative of the image.  Returns (difference quotient, base scene)."""
    dp.restore()
    base = orc.scene(synth.desc)
    losses = []
    for sgn in (+1, -1):
        dp.apply((dp.values + sgn * h * v).astype(np.float32))
        s = orc.scene(synth.desc)
        s.set_sampling(base)
        img, _ = s.render_frame_f64(spp, seed=seed)
        s.set_sampling(None)
        del s
        losses.append(float((img.astype(np.float64) * w).sum()))
    dp.restore()
    return (losses[0] - losses[1]) / (2 * h), base


def steepest_relative_direction(values, group, pilot):
    """Direction for a relative step in one group: v_i = sign(pilot_i) |theta_i|,
    so theta + h v moves every parameter of the group by the fraction h of its
    value, each the way that `pilot` (a gradient from an independent sample
    set) says increases the loss.  In a random direction a group's terms can
    cancel to a directional derivative far below what the group carries."""
    sl = GROUPS[group]
    v = np.zeros(len(values))
    v[sl] = np.sign(pilot[sl]) * np.abs(values[sl])
    return v
