"""Near-field and far-field maps from solved dipole moments (fields.rs),
evaluated on the device through ``lh_pairsum_field_at_points`` and
``lh_pairsum_far_field``.

* ``compute_near_field_map`` -- fields.rs:51-108: the observation grid of a
  ``NearFieldPlane`` (solver/mod.rs:83-96) laid out with ``plane_basis``
  (fields.rs:111-141), |E|^2 at each point.
* ``compute_far_field`` -- fields.rs:156-221: |(I - rhat rhat) F|^2 on an
  n_theta x n_phi grid.
* ``compute_circular_dichroism`` -- fields.rs:236-253: k sum_i
  [Im p_y(x-pol) - Im p_x(y-pol)] (host arithmetic on 3N moments).

The dipoles and the wavenumber are those of the ``PairSumSession``; moments
default to the solution of the session's last solve, already on the device.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, cptr, dptr
from .solver import IncidentField, PairSumSession

__all__ = ["NearFieldPlane", "NearFieldMap", "FarFieldMap", "plane_basis", "field_at_points",
           "compute_near_field_map", "compute_far_field", "compute_circular_dichroism"]


@dataclass
class NearFieldPlane:
    """solver/mod.rs:83-96."""
    centre: tuple = (0.0, 0.0, 0.0)
    normal: tuple = (0.0, 0.0, 1.0)
    half_width: float = 20.0
    half_height: float = 20.0
    nx: int = 50
    ny: int = 50


@dataclass
class NearFieldMap:
    """types.rs:208-219: positions (nx*ny, 3), |E|^2 (nx*ny,), x fastest."""
    positions: np.ndarray
    field_intensity: np.ndarray
    nx: int
    ny: int
    extent: tuple


@dataclass
class FarFieldMap:
    """types.rs:174-187: theta, phi, intensity, each n_theta*n_phi long
    (theta-major)."""
    wavelength_nm: float
    theta: np.ndarray
    phi: np.ndarray
    intensity: np.ndarray
    n_theta: int
    n_phi: int


def _normalise(v):
    n = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    return [v[0] / n, v[1] / n, v[2] / n]


def plane_basis(normal):
    """fields.rs:111-141: u = normalise(seed - (seed.n) n) with seed x unless
    |n_x| >= 0.9 (then y), v = n x u."""
    n = _normalise([float(c) for c in normal])
    seed = [1.0, 0.0, 0.0] if abs(n[0]) < 0.9 else [0.0, 1.0, 0.0]
    dot = seed[0] * n[0] + seed[1] * n[1] + seed[2] * n[2]
    u = _normalise([seed[0] - dot * n[0], seed[1] - dot * n[1], seed[2] - dot * n[2]])
    v = [n[1] * u[2] - n[2] * u[1], n[2] * u[0] - n[0] * u[2], n[0] * u[1] - n[1] * u[0]]
    return u, v


def _moments(session: PairSumSession, moments):
    if moments is None:
        return None
    m = np.ascontiguousarray(np.asarray(moments, dtype=np.complex128)).reshape(-1)
    if m.size != 3 * session.n:
        raise ValueError("moments must hold 3N complex values")
    return m


def field_at_points(session: PairSumSession, points, incident: IncidentField | None = None,
                    moments=None):
    """fields.rs:18-48 at each row of ``points`` (M, 3). Returns (E (M, 3)
    complex, |E|^2 (M,))."""
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64)).reshape(-1, 3)
    m = _moments(session, moments)
    inc = (incident or IncidentField())._c()
    field = np.empty((len(pts), 3), dtype=np.complex128)
    inten = np.empty(len(pts))
    check(session._lib.lh_pairsum_field_at_points(session._h, cptr(m) if m is not None else None,
                                                  C.byref(inc), len(pts), dptr(pts), cptr(field),
                                                  dptr(inten)))
    return field, inten


def plane_points(plane: NearFieldPlane):
    """The observation grid of fields.rs:60-84, x fastest, and the extent of
    fields.rs:97-104."""
    u_hat, v_hat = plane_basis(plane.normal)
    nx, ny = int(plane.nx), int(plane.ny)
    x_min, x_max = -float(plane.half_width), float(plane.half_width)
    y_min, y_max = -float(plane.half_height), float(plane.half_height)
    dx = (x_max - x_min) / (nx - 1) if nx > 1 else 0.0
    dy = (y_max - y_min) / (ny - 1) if ny > 1 else 0.0
    u = x_min + np.arange(nx, dtype=np.float64) * dx
    v = y_min + np.arange(ny, dtype=np.float64) * dy
    uu = np.tile(u, ny)
    vv = np.repeat(v, nx)
    c = [float(x) for x in plane.centre]
    pos = np.empty((nx * ny, 3))
    for a in range(3):
        pos[:, a] = c[a] + uu * u_hat[a] + vv * v_hat[a]
    extent = (c[0] + x_min * u_hat[0] + y_min * v_hat[0],
              c[0] + x_max * u_hat[0] + y_max * v_hat[0],
              c[1] + x_min * u_hat[1] + y_min * v_hat[1],
              c[1] + x_max * u_hat[1] + y_max * v_hat[1])
    return pos, extent


def compute_near_field_map(session: PairSumSession, plane: NearFieldPlane,
                           incident: IncidentField | None = None, moments=None) -> NearFieldMap:
    """fields.rs:51-108 on the session's dipoles and wavenumber."""
    pos, extent = plane_points(plane)
    _, inten = field_at_points(session, pos, incident, moments)
    return NearFieldMap(pos, inten, int(plane.nx), int(plane.ny), extent)


def compute_far_field(session: PairSumSession, n_theta: int, n_phi: int, moments=None,
                      wavelength_nm: float = float("nan")) -> FarFieldMap:
    """fields.rs:156-221 on the session's dipoles and wavenumber."""
    nt, nph = int(n_theta), int(n_phi)
    m = _moments(session, moments)
    theta, phi, inten = np.empty(nt * nph), np.empty(nt * nph), np.empty(nt * nph)
    check(session._lib.lh_pairsum_far_field(session._h, cptr(m) if m is not None else None, nt,
                                            nph, dptr(theta), dptr(phi), dptr(inten)))
    return FarFieldMap(wavelength_nm, theta, phi, inten, nt, nph)


def compute_circular_dichroism(moments_x, moments_y, k: float) -> float:
    """fields.rs:236-253: k sum_i [Im p_y,i(x-pol) - Im p_x,i(y-pol)], summed
    in dipole order."""
    px = np.asarray(moments_x, dtype=np.complex128).reshape(-1, 3)
    py = np.asarray(moments_y, dtype=np.complex128).reshape(-1, 3)
    delta = 0.0
    for a, b in zip(px[:, 1].imag.tolist(), py[:, 0].imag.tolist()):
        delta += a - b
    return k * delta
