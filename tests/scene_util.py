"""Scene descriptions built directly in Python for the known-answer transport
tests: quads, boxes, emitters, media, spot lights and a look-at camera,
assembled into an rp_scene_desc whose arrays this object keeps alive."""
import ctypes as C

import numpy as np

import rypath_amd as rp

WAVELENGTHS = np.arange(360, 831, dtype=np.float64)


def spectrum(kind=rp.SPECTRUM_CONSTANT, c=(0.0, 0.0, 0.0), scale=1.0, table=0):
    s = rp.Spectrum()
    s.kind, s.scale, s.table = kind, scale, table
    for i in range(3):
        s.c[i] = c[i]
    return s


class SceneBuilder:
    def __init__(self, width, height, max_depth):
        self.width, self.height, self.max_depth = width, height, max_depth
        self.positions, self.indices, self.tri_object = [], [], []
        self.objects, self.materials, self.emissions, self.media, self.spots = [], [], [], [], []
        self.pool = []
        self.env = None
        self.camera = None
        self.camera_medium = rp.NONE
        self.tables = []  # (offset, values) of TABLE spectra, for the expectations

    # --- spectra
    def constant(self, v):
        return spectrum(rp.SPECTRUM_CONSTANT, (v, 0, 0))

    def linear(self, a, b):
        return spectrum(rp.SPECTRUM_LINEAR, (a, b, 0))

    def table(self, values):
        values = np.asarray(values, np.float32)
        assert values.shape == (471,)
        off = len(self.pool)
        self.pool.extend(values.tolist())
        return spectrum(rp.SPECTRUM_TABLE, table=off)

    def knots8(self, values):
        """Piecewise-linear spectrum through 8 values at 360, 420, 470, 520,
        570, 620, 700 and 830 nm (the reference's spectral material)."""
        values = np.asarray(values, np.float32)
        assert values.shape == (8,)
        off = len(self.pool)
        self.pool.extend(values.tolist())
        return spectrum(rp.SPECTRUM_KNOTS8, table=off)

    def binned(self, bins):
        """471-sample table, piecewise constant on len(bins) equal wavelength bins
        (nm index i belongs to bin i * n // 471)."""
        n = len(bins)
        return self.table([bins[i * n // 471] for i in range(471)])

    # --- materials, emitters, media, objects
    def diffuse(self, refl):
        m = rp.Material()
        m.kind = rp.MATERIAL_DIFFUSE
        m.reflectance = refl
        m.eta = spectrum(rp.SPECTRUM_CONSTANT, (1.5, 0, 0))
        m.k = spectrum(rp.SPECTRUM_CONSTANT, (0, 0, 0))
        self.materials.append(m)
        return len(self.materials) - 1

    def interface(self):
        m = rp.Material()
        m.kind = rp.MATERIAL_INTERFACE
        m.reflectance = spectrum()
        m.eta = spectrum(rp.SPECTRUM_CONSTANT, (1.0, 0, 0))
        m.k = spectrum()
        self.materials.append(m)
        return len(self.materials) - 1

    def emission(self, radiance, two_sided=False):
        e = rp.Emission()
        e.radiance, e.two_sided = radiance, 1 if two_sided else 0
        self.emissions.append(e)
        return len(self.emissions) - 1

    def medium(self, sigma_a, sigma_s, g):
        m = rp.Medium()
        m.sigma_a, m.sigma_s, m.g = sigma_a, sigma_s, g
        self.media.append(m)
        return len(self.media) - 1

    def object(self, material, emission=rp.NONE, inside=rp.NONE, outside=rp.NONE):
        o = rp.Object()
        o.material, o.emission, o.medium_inside, o.medium_outside = material, emission, inside, outside
        self.objects.append(o)
        return len(self.objects) - 1

    def spot(self, position, direction, cos_start, cos_end, intensity):
        s = rp.SpotLight()
        d = np.asarray(direction, np.float64)
        d = d / np.linalg.norm(d)
        for i in range(3):
            s.position[i], s.direction[i] = position[i], d[i]
        s.cos_falloff_start, s.cos_falloff_end, s.intensity = cos_start, cos_end, intensity
        self.spots.append(s)

    def environment(self, radiance):
        self.env = radiance

    # --- geometry
    def quad(self, p0, p1, p2, p3, obj, facing=None):
        """Quad p0 p1 p2 p3 (in order around its boundary) as two triangles.
        With `facing`, the winding is chosen so that the geometric normal
        (v1 - v0) x (v2 - v0) points along `facing`."""
        p = [np.asarray(v, np.float64) for v in (p0, p1, p2, p3)]
        if facing is not None and np.dot(np.cross(p[1] - p[0], p[2] - p[0]), facing) < 0:
            p = [p[0], p[3], p[2], p[1]]
        base = len(self.positions)
        self.positions.extend(v.tolist() for v in p)
        self.indices += [[base, base + 1, base + 2], [base, base + 2, base + 3]]
        self.tri_object += [obj, obj]

    def mesh(self, points, triangles, obj):
        """Triangles over `points` exactly as given (no winding change)."""
        base = len(self.positions)
        self.positions.extend(np.asarray(v, np.float64).tolist() for v in points)
        self.indices += [[base + a, base + b, base + c] for a, b, c in triangles]
        self.tri_object += [obj] * len(triangles)

    def box(self, lo, hi, obj, inward=True):
        lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
        c = (lo + hi) / 2
        for axis in range(3):
            for side in (lo[axis], hi[axis]):
                u, v = (axis + 1) % 3, (axis + 2) % 3
                corners = []
                for a, b in ((lo[u], lo[v]), (hi[u], lo[v]), (hi[u], hi[v]), (lo[u], hi[v])):
                    q = np.zeros(3)
                    q[axis], q[u], q[v] = side, a, b
                    corners.append(q)
                out = np.zeros(3)
                out[axis] = 1.0 if side > c[axis] else -1.0
                self.quad(*corners, obj, facing=-out if inward else out)

    def look_at(self, origin, target, up, vfov_deg):
        """Pinhole camera: direction = corner + x dx + y dy for x, y in [0, 1),
        the image's first row at the top."""
        o, t, u = (np.asarray(v, np.float64) for v in (origin, target, up))
        f = (t - o) / np.linalg.norm(t - o)
        r = np.cross(f, u)
        r /= np.linalg.norm(r)
        up2 = np.cross(r, f)
        th = np.tan(np.radians(vfov_deg) / 2)
        asp = self.width / self.height
        corner = f - th * asp * r + th * up2
        self.camera = (o, corner, 2 * th * asp * r, -2 * th * up2)

    # --- assembly
    def build(self, hero_wavelengths=0):
        d = rp.SceneDesc()
        keep = []

        def arr(values, ctype, dtype):
            a = np.ascontiguousarray(values, dtype=dtype).reshape(-1)
            if a.size == 0:
                a = np.zeros(1, dtype)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(ctype))

        def structs(items, cls):
            a = (cls * max(len(items), 1))(*items)
            keep.append(a)
            return a

        d.api_version = 1
        d.vertex_count, d.triangle_count, d.object_count = len(self.positions), len(self.indices), len(self.objects)
        d.positions = arr(self.positions, C.c_float, np.float32)
        d.indices = arr(self.indices, C.c_uint32, np.uint32)
        d.tri_object = arr(self.tri_object, C.c_uint32, np.uint32)
        d.objects = structs(self.objects, rp.Object)
        d.material_count, d.emission_count = len(self.materials), len(self.emissions)
        d.medium_count, d.spot_count = len(self.media), len(self.spots)
        d.materials = structs(self.materials, rp.Material)
        d.emissions = structs(self.emissions, rp.Emission)
        d.media = structs(self.media, rp.Medium)
        d.spots = structs(self.spots, rp.SpotLight)
        d.spectra = arr(self.pool, C.c_float, np.float32)
        d.spectra_count = len(self.pool)
        d.has_environment = 1 if self.env is not None else 0
        d.environment = self.env if self.env is not None else spectrum()
        o, corner, dx, dy = self.camera
        for i in range(3):
            d.camera.origin[i], d.camera.corner[i], d.camera.dx[i], d.camera.dy[i] = o[i], corner[i], dx[i], dy[i]
        d.camera.width, d.camera.height, d.camera.medium = self.width, self.height, self.camera_medium
        d.max_depth, d.hero_wavelengths = self.max_depth, hero_wavelengths
        d._keep = keep
        return d

    def camera_rays(self, sub=4):
        """float64 unit camera directions on a sub x sub grid inside every
        pixel: array (height, width, sub*sub, 3)."""
        o, corner, dx, dy = self.camera
        off = (np.arange(sub) + 0.5) / sub
        px = (np.arange(self.width)[:, None] + off[None, :]) / self.width    # (W, sub)
        py = (np.arange(self.height)[:, None] + off[None, :]) / self.height  # (H, sub)
        X = px[None, :, None, :]                                             # (1, W, 1, sub)
        Y = py[:, None, :, None]                                             # (H, 1, sub, 1)
        d = corner + X[..., None] * dx + Y[..., None] * dy                   # (H, W, sub, sub, 3)
        d = d / np.linalg.norm(d, axis=-1, keepdims=True)
        return d.reshape(self.height, self.width, sub * sub, 3)
