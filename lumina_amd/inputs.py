"""Input preparation for the CDA solve: dipole geometry, material permittivity
and polarisabilities.

These are the reference's L5/L6 inputs to the solve (not the hot path
itself), restated in numpy so that the solver session receives exactly the
positions and polarisability blocks the reference's runner would build:

* cubic-lattice discretisation of primitives -- lumina-geometry
  ``discretise.rs:27-58`` (lattice centred on the bounding-box midpoint),
  ``primitives.rs:80-311`` (``contains`` / ``bounding_box``), and the object-local
  recentring of ``scene.rs:256-273``;
* Johnson & Christy Au/Ag tables with the Numerical-Recipes natural cubic
  spline -- lumina-materials ``johnson_christy.rs:51-186, 266-279`` and
  ``spline.rs:32-90``;
* Clausius-Mossotti + radiative-reaction polarisability (``types.rs:243-259``),
  the ellipsoid tensor (``types.rs:286-317``) and depolarisation factors
  (``ellipsoid.rs:66-108``), assembled per wavelength as in
  ``lumina-cli/src/runner.rs:150-186``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

# ---------------------------------------------------------------------------
# Geometry
# ---------------------------------------------------------------------------


@dataclass(frozen=True)
class Sphere:
    radius: float

    def bbox(self):
        r = self.radius
        return np.array([-r, -r, -r]), np.array([r, r, r])

    def contains(self, p):
        # primitives.rs:82-87
        dx, dy, dz = p[:, 0], p[:, 1], p[:, 2]
        return dx * dx + dy * dy + dz * dz <= self.radius * self.radius


@dataclass(frozen=True)
class Cuboid:
    half_extents: tuple

    def bbox(self):
        h = np.asarray(self.half_extents, dtype=np.float64)
        return -h, h.copy()

    def contains(self, p):
        # primitives.rs:88-92
        h = self.half_extents
        return ((np.abs(p[:, 0]) <= h[0]) & (np.abs(p[:, 1]) <= h[1])
                & (np.abs(p[:, 2]) <= h[2]))


@dataclass(frozen=True)
class Ellipsoid:
    semi_axes: tuple

    def bbox(self):
        s = np.asarray(self.semi_axes, dtype=np.float64)
        return -s, s.copy()

    def contains(self, p):
        # primitives.rs:93-98
        a = self.semi_axes
        dx = p[:, 0] / a[0]
        dy = p[:, 1] / a[1]
        dz = p[:, 2] / a[2]
        return dx * dx + dy * dy + dz * dz <= 1.0


def _unit(axis):
    ax = np.asarray(axis, dtype=np.float64)
    ln = math.sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2])
    return ax / ln, ln


@dataclass(frozen=True)
class Cylinder:
    radius: float
    length: float
    axis: tuple = (0.0, 0.0, 1.0)

    def bbox(self):
        # primitives.rs:252-283 (base centre at the origin, scene.rs:286-290)
        ax, ln = _unit(self.axis)
        if ln <= 1e-15:
            ax = np.array([0.0, 0.0, 1.0])
        base = np.zeros(3)
        top = base + ax * self.length
        lo = np.zeros(3)
        hi = np.zeros(3)
        for i in range(3):
            r_ext = self.radius * math.sqrt(max(1.0 - ax[i] * ax[i], 0.0))
            lo[i] = min(base[i], top[i]) - r_ext
            hi[i] = max(base[i], top[i]) + r_ext
        return lo, hi

    def contains(self, p):
        # primitives.rs:99-126
        ax, ln = _unit(self.axis)
        if ln < 1e-15:
            return np.zeros(len(p), dtype=bool)
        dp = p  # base centre at the origin
        proj = dp[:, 0] * ax[0] + dp[:, 1] * ax[1] + dp[:, 2] * ax[2]
        perp_sq = dp[:, 0] * dp[:, 0] + dp[:, 1] * dp[:, 1] + dp[:, 2] * dp[:, 2] - proj * proj
        return (proj >= 0.0) & (proj <= self.length) & (perp_sq <= self.radius * self.radius)


@dataclass(frozen=True)
class Helix:
    radius: float
    pitch: float
    turns: float
    wire_radius: float
    axis: tuple = (0.0, 0.0, 1.0)

    def bbox(self):
        # primitives.rs:284-312
        ax, ln = _unit(self.axis)
        if ln <= 1e-15:
            ax = np.array([0.0, 0.0, 1.0])
        total_h = self.pitch * self.turns
        outer = self.radius + self.wire_radius
        lo = np.zeros(3)
        hi = np.zeros(3)
        for i in range(3):
            r_ext = outer * math.sqrt(max(1.0 - ax[i] * ax[i], 0.0))
            top = 0.0 + total_h * ax[i]
            lo[i] = min(0.0, top) - r_ext - self.wire_radius
            hi[i] = max(0.0, top) + r_ext + self.wire_radius
        return lo, hi

    def contains(self, p):
        # primitives.rs:127-209: nearest centre-line sample over a one-turn
        # window (33 samples) around the axial guess.
        ax, ln = _unit(self.axis)
        if ln < 1e-15:
            return np.zeros(len(p), dtype=bool)
        seed = np.array([1.0, 0.0, 0.0]) if abs(ax[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
        dot_sa = seed[0] * ax[0] + seed[1] * ax[1] + seed[2] * ax[2]
        u_raw = seed - dot_sa * ax
        u_len = math.sqrt(u_raw[0] * u_raw[0] + u_raw[1] * u_raw[1] + u_raw[2] * u_raw[2])
        u = u_raw / u_len
        v = np.array([ax[1] * u[2] - ax[2] * u[1], ax[2] * u[0] - ax[0] * u[2],
                      ax[0] * u[1] - ax[1] * u[0]])
        z = p[:, 0] * ax[0] + p[:, 1] * ax[1] + p[:, 2] * ax[2]
        total_h = self.pitch * self.turns
        inside_z = ~((z < -self.wire_radius) | (z > total_h + self.wire_radius))
        pu = p[:, 0] * u[0] + p[:, 1] * u[1] + p[:, 2] * u[2]
        pv = p[:, 0] * v[0] + p[:, 1] * v[1] + p[:, 2] * v[2]
        two_pi = 2.0 * math.pi
        t_max = two_pi * self.turns
        t_guess = z * two_pi / self.pitch if self.pitch > 1e-15 else np.zeros_like(z)
        t_guess = np.clip(t_guess, 0.0, t_max)
        t_lo = np.maximum(t_guess - two_pi, 0.0)
        t_hi = np.minimum(t_guess + two_pi, t_max)
        best = np.full(len(p), np.inf)
        ns = 32
        for si in range(ns + 1):
            t = t_lo + (t_hi - t_lo) * float(si) / float(ns)
            cz = t * self.pitch / two_pi
            cu = self.radius * np.cos(t)
            cv = self.radius * np.sin(t)
            dz = z - cz
            du = pu - cu
            dv = pv - cv
            d2 = dz * dz + du * du + dv * dv
            best = np.where(d2 < best, d2, best)
        return inside_z & (best <= self.wire_radius * self.wire_radius)


def discretise(shape, spacing: float) -> np.ndarray:
    """Dipole positions (N, 3), object-local and centred.

    discretise.rs:27-58 (cubic lattice about the bbox midpoint, x outer, z
    inner) followed by scene.rs:256-273 (subtract the bbox midpoint).
    """
    assert spacing > 0.0
    lo, hi = shape.bbox()
    cx = 0.5 * (lo + hi)
    n = np.floor((hi - cx) / spacing).astype(np.int64)
    ix = np.arange(-n[0], n[0] + 1, dtype=np.float64)
    iy = np.arange(-n[1], n[1] + 1, dtype=np.float64)
    iz = np.arange(-n[2], n[2] + 1, dtype=np.float64)
    xs = cx[0] + ix * spacing
    ys = cx[1] + iy * spacing
    zs = cx[2] + iz * spacing
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    keep = shape.contains(pts)
    pts = pts[keep]
    if len(pts) == 0:
        raise ValueError("shape produced no dipoles")
    return np.ascontiguousarray(pts - cx[None, :])


# ---------------------------------------------------------------------------
# Materials
# ---------------------------------------------------------------------------

# Johnson & Christy 1972 (lambda/um, n, k), johnson_christy.rs:54-98 (Au) and
# :120-165 (Ag). Tabulated optical constants are data, kept identical.
_JC_AU = [
    (0.1879, 1.28, 1.188), (0.1916, 1.32, 1.203), (0.1953, 1.34, 1.226),
    (0.1993, 1.33, 1.251), (0.2033, 1.33, 1.277), (0.2073, 1.30, 1.304),
    (0.2119, 1.30, 1.350), (0.2164, 1.30, 1.387), (0.2214, 1.30, 1.427),
    (0.2262, 1.31, 1.460), (0.2313, 1.30, 1.497), (0.2371, 1.32, 1.536),
    (0.2426, 1.32, 1.577), (0.2490, 1.33, 1.631), (0.2551, 1.33, 1.688),
    (0.2616, 1.35, 1.749), (0.2689, 1.38, 1.803), (0.2761, 1.43, 1.847),
    (0.2844, 1.47, 1.869), (0.2924, 1.49, 1.878), (0.3009, 1.53, 1.889),
    (0.3107, 1.53, 1.893), (0.3204, 1.54, 1.898), (0.3315, 1.48, 1.883),
    (0.3425, 1.48, 1.871), (0.3542, 1.50, 1.866), (0.3679, 1.48, 1.895),
    (0.3815, 1.46, 1.933), (0.3974, 1.47, 1.952), (0.4133, 1.46, 1.958),
    (0.4305, 1.45, 1.948), (0.4509, 1.38, 1.914), (0.4714, 1.31, 1.849),
    (0.4959, 1.04, 1.833), (0.5209, 0.62, 2.081), (0.5486, 0.43, 2.455),
    (0.5821, 0.29, 2.863), (0.6168, 0.21, 3.272), (0.6595, 0.14, 3.697),
    (0.7045, 0.13, 4.103), (0.7560, 0.14, 4.542), (0.8211, 0.16, 5.083),
    (0.8920, 0.17, 5.663),
]
_JC_AG = [
    (0.1879, 1.07, 1.212), (0.1916, 1.10, 1.232), (0.1953, 1.12, 1.255),
    (0.1993, 1.14, 1.274), (0.2033, 1.15, 1.296), (0.2073, 1.18, 1.312),
    (0.2119, 1.20, 1.326), (0.2164, 1.22, 1.342), (0.2214, 1.27, 1.367),
    (0.2262, 1.31, 1.389), (0.2313, 1.33, 1.414), (0.2371, 1.38, 1.431),
    (0.2426, 1.45, 1.439), (0.2490, 1.57, 1.432), (0.2551, 1.73, 1.416),
    (0.2616, 1.93, 1.419), (0.2689, 2.21, 1.468), (0.2761, 2.54, 1.569),
    (0.2844, 2.80, 1.650), (0.2924, 2.38, 1.340), (0.3009, 1.88, 0.980),
    (0.3107, 1.44, 0.680), (0.3204, 0.88, 0.580), (0.3315, 0.35, 0.960),
    (0.3425, 0.24, 1.620), (0.3542, 0.20, 2.070), (0.3679, 0.18, 2.440),
    (0.3815, 0.16, 2.820), (0.3974, 0.145, 2.950), (0.4133, 0.145, 2.844),
    (0.4305, 0.138, 3.042), (0.4509, 0.134, 3.206), (0.4714, 0.130, 3.380),
    (0.4959, 0.126, 3.567), (0.5209, 0.120, 3.765), (0.5486, 0.114, 3.977),
    (0.5821, 0.107, 4.210), (0.6168, 0.100, 4.458), (0.6595, 0.093, 4.758),
    (0.7045, 0.086, 5.067), (0.7560, 0.080, 5.435), (0.8211, 0.074, 5.882),
    (0.8920, 0.066, 6.372),
]


class NaturalSpline:
    """Numerical-Recipes natural cubic spline (spline.rs:32-90)."""

    def __init__(self, xs, ys):
        xs = [float(v) for v in xs]
        ys = [float(v) for v in ys]
        n = len(xs)
        assert n >= 2 and all(xs[i] > xs[i - 1] for i in range(1, n))
        y2 = [0.0] * n
        u = [0.0] * (n - 1)
        for i in range(1, n - 1):
            sig = (xs[i] - xs[i - 1]) / (xs[i + 1] - xs[i - 1])
            p = sig * y2[i - 1] + 2.0
            y2[i] = (sig - 1.0) / p
            ui = (ys[i + 1] - ys[i]) / (xs[i + 1] - xs[i]) - (ys[i] - ys[i - 1]) / (xs[i] - xs[i - 1])
            u[i] = (6.0 * ui / (xs[i + 1] - xs[i - 1]) - sig * u[i - 1]) / p
        for k in range(n - 3, -1, -1):
            y2[k + 1] = y2[k + 1] * y2[k + 2] + u[k + 1]
        self.xs, self.ys, self.y2 = xs, ys, y2

    def __call__(self, x: float) -> float:
        xs, ys, y2 = self.xs, self.ys, self.y2
        lo, hi = 0, len(xs) - 1
        while hi - lo > 1:
            mid = (lo + hi) // 2
            if xs[mid] > x:
                hi = mid
            else:
                lo = mid
        h = xs[hi] - xs[lo]
        a = (xs[hi] - x) / h
        b = (x - xs[lo]) / h
        return (a * ys[lo] + b * ys[hi]
                + ((a * a * a - a) * y2[lo] + (b * b * b - b) * y2[hi]) * h * h / 6.0)


class JohnsonChristy:
    """epsilon(lambda) from the J&C table (johnson_christy.rs:266-279)."""

    def __init__(self, name: str):
        table = {"Au": _JC_AU, "Ag": _JC_AG}[name]
        self.name = name
        wl = [t[0] * 1000.0 for t in table]
        er = [t[1] * t[1] - t[2] * t[2] for t in table]
        ei = [2.0 * t[1] * t[2] for t in table]
        self.wl = wl
        self.re = NaturalSpline(wl, er)
        self.im = NaturalSpline(wl, ei)

    def epsilon(self, wavelength_nm: float) -> complex:
        if wavelength_nm < self.wl[0] or wavelength_nm > self.wl[-1]:
            raise ValueError(f"{self.name}: wavelength {wavelength_nm} nm outside "
                             f"[{self.wl[0]}, {self.wl[-1]}]")
        return complex(self.re(wavelength_nm), self.im(wavelength_nm))


_MATERIALS = {}


def material(name: str) -> JohnsonChristy:
    if name not in _MATERIALS:
        _MATERIALS[name] = JohnsonChristy(name)
    return _MATERIALS[name]


# ---------------------------------------------------------------------------
# Polarisability
# ---------------------------------------------------------------------------

def _cdiv(a: complex, b: complex) -> complex:
    # num-complex quotient (a * conj(b)) / |b|^2, no scaling
    nsq = b.real * b.real + b.imag * b.imag
    return complex((a.real * b.real + a.imag * b.imag) / nsq,
                   (a.imag * b.real - a.real * b.imag) / nsq)


def _cmul(a: complex, b: complex) -> complex:
    return complex(a.real * b.real - a.imag * b.imag, a.real * b.imag + a.imag * b.real)


def wavenumber(wavelength_nm: float, n_medium: float) -> float:
    """cda/mod.rs:169-171."""
    return 2.0 * math.pi * n_medium / wavelength_nm


def clausius_mossotti(volume: float, eps: complex, eps_m: float) -> complex:
    """types.rs:243-246."""
    num = complex(3.0 * volume * (eps.real - eps_m), 3.0 * volume * eps.imag)
    den = complex(eps.real + 2.0 * eps_m, eps.imag)
    return _cdiv(num, den)


def radiative_correction(alpha_cm: complex, k: float) -> complex:
    """types.rs:256-259 (Draine radiative reaction)."""
    corr = complex(0.0, (k * k * k) / (6.0 * math.pi))
    return _cdiv(alpha_cm, complex(1.0, 0.0) - _cmul(corr, alpha_cm))


_GL16_X = [0.0950125098360623, 0.2816035507792589, 0.4580167776572274,
           0.6178762444026438, 0.7554044083550030, 0.8656312023835650,
           0.9445750230732326, 0.9894009349916499]
_GL16_W = [0.1894506104550685, 0.1826034150449236, 0.1691565193950025,
           0.1495959888165767, 0.1246289712555339, 0.0951585116824928,
           0.0622535239386479, 0.0271524594117541]


def depol_factors(semi_axes) -> tuple:
    """ellipsoid.rs:66-108: L_i via 16-point Gauss-Legendre after the
    substitution s = s_i^2 (1-u^2)/u^2."""
    a, b, c = (float(v) for v in semi_axes)
    a2, b2, c2 = a * a, b * b, c * c
    pre = a * b * c * 0.5

    def integrand(si2, u):
        u2 = u * u
        qa = si2 + (a2 - si2) * u2
        qb = si2 + (b2 - si2) * u2
        qc = si2 + (c2 - si2) * u2
        return 2.0 * u2 / math.sqrt(qa * qb * qc)

    def integ(si2):
        s = 0.0
        for x, w in zip(_GL16_X, _GL16_W):
            w = w * 0.5
            s += w * integrand(si2, (1.0 + x) * 0.5)
            s += w * integrand(si2, (1.0 - x) * 0.5)
        return s

    return (pre * integ(a2), pre * integ(b2), pre * integ(c2))


def ellipsoid_tensor(volume, depol, eps: complex, eps_m: float, k: float,
                     rot=(1, 0, 0, 0, 1, 0, 0, 0, 1)) -> np.ndarray:
    """types.rs:286-317: per-axis CM + RRCM, rotated R diag R^T. Returns a
    row-major complex (9,) block."""
    d_eps = complex(eps.real - eps_m, eps.imag)
    alphas = []
    for L in depol:
        den = complex(eps_m + L * d_eps.real, L * d_eps.imag)
        a_cm = _cdiv(complex(volume * d_eps.real, volume * d_eps.imag), den)
        alphas.append(radiative_correction(a_cm, k))
    t = np.zeros(9, dtype=np.complex128)
    for i in range(3):
        for j in range(3):
            s = 0j
            for q in range(3):
                s += alphas[q] * rot[i * 3 + q] * rot[j * 3 + q]
            t[i * 3 + j] = s
    return t


def isotropic_block(alpha: complex) -> np.ndarray:
    """types.rs:22-32 Dipole::isotropic."""
    t = np.zeros(9, dtype=np.complex128)
    t[0] = t[4] = t[8] = alpha
    return t


def polarisability_block(material_name: str, wavelength_nm: float, spacing: float,
                         n_env: float, depol=None) -> np.ndarray:
    """Per-lambda alpha block of runner.rs:171-183 for a single-material
    particle (uniform over its dipoles)."""
    eps = material(material_name).epsilon(wavelength_nm)
    k = wavenumber(wavelength_nm, n_env)
    eps_m = n_env * n_env
    vol = spacing ** 3
    if depol is None:
        return isotropic_block(radiative_correction(clausius_mossotti(vol, eps, eps_m), k))
    return ellipsoid_tensor(vol, depol, eps, eps_m, k)


def polarisability_from_epsilon(eps: complex, wavelength_nm: float, spacing: float,
                                n_env: float, depol=None) -> np.ndarray:
    """The alpha block of a dipole of permittivity eps (runner.rs:171-183)."""
    k = wavenumber(wavelength_nm, n_env)
    eps_m = n_env * n_env
    vol = spacing ** 3
    if depol is None:
        return isotropic_block(radiative_correction(clausius_mossotti(vol, eps, eps_m), k))
    return ellipsoid_tensor(vol, depol, eps, eps_m, k)


def inverse_polarisability_derivative(eps: complex, spacing: float, n_env: float) -> np.ndarray:
    """d(alpha^-1)/d eps for the CM + radiative-reaction model (types.rs:243-259,
    286-317), as a row-major 3x3 block (no reference counterpart).

    alpha_RR^-1 = alpha^-1 - i k^3 / (6 pi) per principal axis, and
    alpha^-1 = (eps_m + L (eps - eps_m)) / (V (eps - eps_m)) (L = 1/3 for the
    isotropic CM form), so d(alpha^-1)/d eps = -eps_m / (V (eps - eps_m)^2) on
    every axis: a multiple of the identity for any depolarisation factors
    and orientation."""
    eps_m = n_env * n_env
    vol = spacing ** 3
    d = -eps_m / (vol * (complex(eps) - eps_m) ** 2)
    return isotropic_block(d)


# ---------------------------------------------------------------------------
# Workloads named by what they are
# ---------------------------------------------------------------------------

def au_sphere_r10_d2():
    return discretise(Sphere(10.0), 2.0)


def au_cuboid_80x40x40_d1p5():
    return discretise(Cuboid((40.0, 20.0, 20.0)), 1.5)


def au_cylinder_d100_h40_d2():
    return discretise(Cylinder(50.0, 40.0), 2.0)


def au_helix_r25_p40_t4_w8_d1p5():
    return discretise(Helix(25.0, 40.0, 4.0, 8.0), 1.5)


@dataclass
class Particle:
    shape: str
    material: str
    radius: float
    semi_axes: tuple
    positions: np.ndarray
    depol: tuple | None


def batch_particles(count: int = 64, seed: int = 0, spacing: float = 2.0):
    """64 particles from a fixed numpy PCG64 seed: spheres and ellipsoids of
    Au/Ag, radius U(5, 30) nm, ellipsoid semi-axes (r, r/s1, r/s2) with
    s ~ U(1, 2.5), spacing 2 nm."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        shape = "sphere" if rng.random() < 0.5 else "ellipsoid"
        mat = "Au" if rng.random() < 0.5 else "Ag"
        r = float(rng.uniform(5.0, 30.0))
        s1 = float(rng.uniform(1.0, 2.5))
        s2 = float(rng.uniform(1.0, 2.5))
        if shape == "sphere":
            axes = (r, r, r)
            pos = discretise(Sphere(r), spacing)
            depol = None
        else:
            axes = (r, r / s1, r / s2)
            pos = discretise(Ellipsoid(axes), spacing)
            depol = depol_factors(axes)
        out.append(Particle(shape, mat, r, axes, pos, depol))
    return out


def batch_workload(n_particles: int = 64, n_wl: int = 16, n_env: float = 1.33,
                   spacing: float = 2.0, seed: int = 0):
    """The batched-solve workload: ``batch_particles(n_particles, seed)`` in
    water (n_env), each at n_wl wavelengths 400 + t 400 / (n_wl - 1) nm, FCD
    with cell_size = spacing, x-polarised light along +z. One problem dict
    per (particle, wavelength), particle-major, with the keys Batch and
    PairSumSession take (positions, polarisability, k, use_fcd, cell_size)
    plus n, wl, material, shape and particle (its index)."""
    parts = batch_particles(n_particles, seed=seed, spacing=spacing)
    wls = [400.0 + t * 400.0 / (n_wl - 1) for t in range(n_wl)] if n_wl > 1 else [600.0]
    probs = []
    for ip, p in enumerate(parts):
        for lam in wls:
            probs.append(dict(positions=p.positions,
                              polarisability=polarisability_block(p.material, lam, spacing, n_env,
                                                                  depol=p.depol),
                              k=wavenumber(lam, n_env), use_fcd=True, cell_size=spacing,
                              n=len(p.positions), wl=lam, material=p.material, shape=p.shape,
                              particle=ip))
    return probs
