"""The CLI wavelength sweep (lumina-cli/src/runner.rs:145-218) as one call:
``run_sweep`` hands the geometry, the per-dipole optical model and the
permittivity tables to ``lh_sweep_solve``, which sets up every wavelength on
the device (spline -> eps -> alpha -> alpha^-1, substrate factor, incident
field) and solves all of them in one batch.

* ``TabulatedMaterial`` -- a MaterialProvider backed by a table
  (johnson_christy.rs:29-45, 266-279); ``johnson_christy("Au" | "Ag")``.
* ``SubstrateSpec`` -- the substrate of a job (substrate.rs:125-137):
  interface height, material, retarded or quasi-static factor.
* ``EllipsoidHint`` -- the anisotropy hint of a dipole (depolarisation
  factors and rotation, types.rs:286-317).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from . import inputs as I
from ._lib import check
from .solver import FP64, IncidentField, LatticeSpec, _spec

__all__ = ["TabulatedMaterial", "johnson_christy", "SubstrateSpec", "EllipsoidHint", "run_sweep"]


@dataclass
class TabulatedMaterial:
    name: str
    wavelength_nm: np.ndarray
    epsilon: np.ndarray


def johnson_christy(name: str) -> TabulatedMaterial:
    """The Johnson-Christy table of ``name`` as eps = (n + ik)^2 at its
    wavelengths (johnson_christy.rs:46-120)."""
    m = I.material(name)
    return TabulatedMaterial(name, np.array(m.wl, dtype=float),
                             np.array(m.re.ys, dtype=float) + 1j * np.array(m.im.ys, dtype=float))


@dataclass
class SubstrateSpec:
    z_interface: float
    material: int              # index into the materials of the sweep
    use_retarded: bool = True


@dataclass
class EllipsoidHint:
    depol: tuple               # (L1, L2, L3)
    rotation: tuple = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)


def run_sweep(backend, positions, wavelengths_nm, materials, material_index=None, spacing=2.0,
              hints=None, hint_index=None, n_env: float = 1.0, use_fcd: bool = True,
              cell_size: float | None = None, substrate: SubstrateSpec | None = None,
              lattice: LatticeSpec | None = None, k_bloch=(0.0, 0.0, 0.0),
              incident: IncidentField | None = None, tolerance: float = 1e-6,
              max_iterations: int = 1000, precision: int = FP64):
    """Cross sections (n_wl, 3) of every wavelength, NaN where the solve did
    not converge (runner.rs:189-203), and the solve statistics.

    ``spacing`` is a scalar or one value per dipole (volume = spacing^3);
    ``material_index`` one material per dipole (default 0); ``hints`` a list
    of EllipsoidHint with ``hint_index`` one index per dipole (-1:
    isotropic). ``cell_size`` defaults to the uniform spacing."""
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1, 3))
    n = len(pos)
    lam = np.ascontiguousarray(np.asarray(wavelengths_nm, dtype=np.float64).reshape(-1))
    spacing_arr = np.asarray(spacing, dtype=np.float64)
    if cell_size is None:
        if spacing_arr.ndim != 0:
            raise ValueError("cell_size is needed with per-dipole spacings")
        cell_size = float(spacing_arr)
    sp, keep = _spec(pos, np.eye(3).reshape(9), 1.0, use_fcd, cell_size, None, precision,
                     lattice, k_bloch)
    if substrate is not None:
        sp.has_substrate = 1
        sp.z_interface = float(substrate.z_interface)
    spec = _lib.SweepSpec()
    spec.geometry = sp
    arrays = [keep]
    if material_index is not None:
        mi = np.ascontiguousarray(np.asarray(material_index, dtype=np.int32).reshape(n))
        spec.material = mi.ctypes.data_as(C.POINTER(C.c_int32))
        arrays.append(mi)
    if spacing_arr.ndim == 0:
        spec.uniform_spacing = float(spacing_arr)
    else:
        sa = np.ascontiguousarray(spacing_arr.reshape(n))
        spec.spacing = sa.ctypes.data_as(C.POINTER(C.c_double))
        arrays.append(sa)
    if hints:
        hi = np.ascontiguousarray(np.asarray(hint_index, dtype=np.int32).reshape(n))
        dep = np.ascontiguousarray(np.array([h.depol for h in hints], dtype=np.float64))
        rot = np.ascontiguousarray(np.array([h.rotation for h in hints], dtype=np.float64))
        spec.hint = hi.ctypes.data_as(C.POINTER(C.c_int32))
        spec.n_hints = len(hints)
        spec.hint_depol = dep.ctypes.data_as(C.POINTER(C.c_double))
        spec.hint_rotation = rot.ctypes.data_as(C.POINTER(C.c_double))
        arrays += [hi, dep, rot]
    spec.n_env = float(n_env)
    if substrate is not None:
        spec.substrate_material = int(substrate.material)
        spec.substrate_retarded = 1 if substrate.use_retarded else 0
    mats = (_lib.Material * len(materials))()
    for q, m in enumerate(materials):
        wl = np.ascontiguousarray(np.asarray(m.wavelength_nm, dtype=np.float64))
        ep = np.ascontiguousarray(np.asarray(m.epsilon, dtype=np.complex128))
        if wl.shape != ep.shape:
            raise ValueError(f"material {m.name}: one permittivity per wavelength")
        mats[q].n_points = len(wl)
        mats[q].wavelength_nm = wl.ctypes.data_as(C.POINTER(C.c_double))
        mats[q].epsilon = C.c_void_p(ep.ctypes.data)
        arrays += [wl, ep]
    inc = (incident or IncidentField())._c()
    cs = np.zeros((len(lam), 3))
    infos = (_lib.SolveInfo * len(lam))()
    lib = backend._lib
    st = lib.lh_sweep_solve(backend._h, C.byref(spec), mats, len(materials),
                            lam.ctypes.data_as(C.POINTER(C.c_double)), len(lam), C.byref(inc),
                            float(tolerance), int(max_iterations),
                            cs.ctypes.data_as(C.POINTER(C.c_double)), infos)
    if st != _lib.LH_CONVERGENCE_FAILURE:
        check(st)
    del arrays
    return cs, [i.as_dict() for i in infos]
