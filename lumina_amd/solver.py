"""Host-side mirror of the reference's device and solver interface for the
CDA solve, on top of the C ABI (liblumina_hip.so).

Names follow the reference:

* ``HipBackend`` -- ``ComputeBackend`` (lumina-compute/src/backend.rs:62-134):
  ``device_info``, ``create_session`` (stored-matrix ``MatvecSession``,
  backend.rs:47-55) and ``create_pairsum_session`` (the on-the-fly matvec of
  assembly.rs:209-318, which the reference calls directly from
  cda/mod.rs:329-345, 375-391).
* ``CdaSolver`` -- ``CdaSolver`` (lumina-core/src/solver/cda/mod.rs:44-96)
  with ``compute_cross_sections`` (:175-236) and ``solve_dipoles``
  (:238-419), always on the on-the-fly GMRES path.
* ``IncidentField``, ``SimulationParams``, ``CrossSections``,
  ``DipoleResponse`` -- types.rs:37-204; ``SubstrateRuntime`` -- substrate.rs:98-146.
* Errors: ``ConvergenceFailure`` (SolverError::ConvergenceFailure,
  solver/mod.rs:18-31), ``LuminaError`` for ComputeError variants.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (LH_PREC_FP64, LH_PREC_MIXED, ConvergenceFailure, LuminaError, check, cptr,
                   dptr)

__all__ = [
    "HipBackend", "MatvecSession", "PairSumSession", "CdaSolver", "IncidentField",
    "SimulationParams", "CrossSections", "DipoleResponse", "SubstrateRuntime", "LatticeSpec", "Batch",
    "ConvergenceFailure", "LuminaError", "substrate_reflection_factor", "batch_compute",
    "FP64", "MIXED",
]

FP64 = LH_PREC_FP64
MIXED = LH_PREC_MIXED


def _c128(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128))


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _sized(a, want: int, what: str) -> np.ndarray:
    """A flat complex copy of ``a`` that must hold exactly ``want`` values:
    the C ABI reads and writes fixed lengths (3N per vector, 9 or 9N per
    polarisability), so a wrong length is refused before any C call."""
    v = _c128(a).reshape(-1)
    if v.size != want:
        raise ValueError(f"{what} has {v.size} values, expected {want}")
    return v


@dataclass
class IncidentField:
    """types.rs:37-74. ``polarisation`` may be complex (circular light)."""
    direction: tuple = (0.0, 0.0, 1.0)
    polarisation: tuple = (1.0, 0.0, 0.0)
    amplitude: float = 1.0

    def _c(self) -> _lib.Incident:
        inc = _lib.Incident()
        for i in range(3):
            inc.direction[i] = float(self.direction[i])
            p = complex(self.polarisation[i])
            inc.polarisation[i] = _lib.c64(p.real, p.imag)
        inc.amplitude = float(self.amplitude)
        return inc


@dataclass
class SubstrateRuntime:
    """SubstrateRuntime::Fresnel (substrate.rs:103-107)."""
    z_interface: float
    delta_eps: complex


def substrate_reflection_factor(eps_sub: complex, eps_env: float, use_retarded: bool = True):
    """substrate.rs:160-197: retarded (n_s - n_e)/(n_s + n_e) or quasi-static
    (eps_s - eps_e)/(eps_s + eps_e)."""
    eps_sub = complex(eps_sub)
    if use_retarded:
        ns = np.sqrt(np.complex128(eps_sub))
        ne = math.sqrt(eps_env)
        return complex((ns - ne) / (ns + ne))
    return complex((eps_sub - eps_env) / (eps_sub + eps_env))


@dataclass
class LatticeSpec:
    """lattice.rs:29-44 (2D lattice in the xy plane). ewald_eta None or <= 0
    selects sqrt(pi / A) (lattice.rs:96-102)."""
    a1: tuple
    a2: tuple
    ewald_eta: float | None = None
    truncation_real: int = 5
    truncation_recip: int = 5


@dataclass
class SimulationParams:
    """types.rs:78-115 (only the fields the solve reads)."""
    environment_n: float = 1.0
    solver_tolerance: float = 1e-6
    max_iterations: int = 1000
    k_bloch: tuple = (0.0, 0.0, 0.0)
    substrate_runtime: SubstrateRuntime | None = None


@dataclass
class CrossSections:
    wavelength_nm: float
    extinction: float
    absorption: float
    scattering: float
    circular_dichroism: float | None = None


@dataclass
class DipoleResponse:
    wavelength_nm: float
    environment_n: float
    moments: np.ndarray        # (N, 3) complex
    local_fields: np.ndarray   # (N, 3) complex
    info: dict = field(default_factory=dict)


def wavenumber(wavelength_nm: float, n_medium: float) -> float:
    """cda/mod.rs:169-171."""
    return 2.0 * math.pi * n_medium / wavelength_nm


def _spec(positions, polarisability, k, use_fcd, cell_size, substrate, precision, lattice=None,
          k_bloch=(0.0, 0.0, 0.0), classify_positions=None):
    pos = _f64(positions).reshape(-1, 3)
    pol = _c128(polarisability)
    n = pos.shape[0]
    if pol.size == 9:
        uniform = 1
        pol = pol.reshape(9)
    elif pol.size == 9 * n:
        uniform = 0
        pol = pol.reshape(n, 9)
    else:
        raise ValueError("polarisability must be (9,) or (N, 9)")
    sp = _lib.PairsumSpec()
    sp.n_dipoles = n
    sp.positions = dptr(pos)
    sp.polarisability = cptr(pol)
    sp.polarisability_uniform = uniform
    sp.k = float(k)
    sp.use_fcd = 1 if use_fcd else 0
    sp.cell_size = float(cell_size)
    sp.has_lattice = 0
    if lattice is not None:
        sp.has_lattice = 1
        for i in range(3):
            sp.a1[i] = float(lattice.a1[i])
            sp.a2[i] = float(lattice.a2[i])
        sp.ewald_eta = float(lattice.ewald_eta or 0.0)
        sp.truncation_real = int(lattice.truncation_real)
        sp.truncation_recip = int(lattice.truncation_recip)
        for i in range(3):
            sp.k_bloch[i] = float(k_bloch[i])
    if substrate is not None:
        sp.has_substrate = 1
        sp.z_interface = float(substrate.z_interface)
        d = complex(substrate.delta_eps)
        sp.delta_eps = _lib.c64(d.real, d.imag)
    sp.precision = int(precision)
    cls = None
    if classify_positions is not None:
        cls = _f64(classify_positions).reshape(-1, 3)
        if cls.shape != pos.shape:
            raise ValueError("classify_positions must have the shape of positions")
        sp.classify_positions = dptr(cls)
    # keep the arrays alive as long as the spec
    return sp, (pos, pol, cls)


class MatvecSession:
    """Stored-matrix session (MatvecSession, backend.rs:47-55)."""

    def __init__(self, backend: "HipBackend", n: int, _handle=None):
        self._lib = backend._lib
        self.n = int(n)
        if _handle is None:
            h = C.c_void_p()
            check(self._lib.lh_session_create(backend._h, self.n, C.byref(h)))
            _handle = h
        self._h = _handle

    def upload_matrix(self, matrix) -> None:
        m = _c128(matrix)
        if m.ndim != 2 or m.shape[0] != m.shape[1]:
            raise ValueError("matrix must be square")
        check(self._lib.lh_session_upload_matrix(self._h, cptr(m), m.shape[0]))

    def matvec(self, x) -> np.ndarray:
        xv = _c128(x).reshape(-1)
        y = np.empty_like(xv)
        check(self._lib.lh_session_matvec(self._h, cptr(xv), cptr(y), xv.size))
        return y

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lh_session_destroy(self._h)
            self._h = None

    __del__ = close


class PairSumSession:
    """On-the-fly pair-sum operator A = alpha^-1 - G (assembly.rs:209-318)
    with device GMRES (iterative.rs:33-160)."""

    def __init__(self, backend: "HipBackend", positions, polarisability, k: float,
                 use_fcd: bool = True, cell_size: float = 3.0,
                 substrate: SubstrateRuntime | None = None, precision: int = FP64,
                 lattice: LatticeSpec | None = None, k_bloch=(0.0, 0.0, 0.0),
                 classify_positions=None):
        self._lib = backend._lib
        sp, keep = _spec(positions, polarisability, k, use_fcd, cell_size, substrate, precision,
                         lattice, k_bloch, classify_positions)
        self.n = int(sp.n_dipoles)
        self.k = float(k)
        self.precision = int(precision)
        h = C.c_void_p()
        check(self._lib.lh_pairsum_create(backend._h, C.byref(sp), C.byref(h)))
        self._h = h
        self._pos = keep[0]

    def set_wavelength(self, k: float, polarisability, delta_eps: complex = 0j):
        pol = _c128(polarisability).reshape(-1)
        if pol.size not in (9, 9 * self.n):
            raise ValueError(f"polarisability has {pol.size} values, expected 9 or {9 * self.n}")
        uniform = 1 if pol.size == 9 else 0
        d = complex(delta_eps)
        check(self._lib.lh_pairsum_set_wavelength(self._h, float(k), cptr(pol), uniform,
                                                  _lib.c64(d.real, d.imag)))
        self.k = float(k)

    def matvec(self, x, precision: int | None = None) -> np.ndarray:
        xv = _c128(x).reshape(-1)
        if xv.size != 3 * self.n:
            raise LuminaError(_lib.LH_DEVICE_ERROR, "vector dimension mismatch")
        y = np.empty_like(xv)
        p = self.precision if precision is None else int(precision)
        check(self._lib.lh_pairsum_matvec(self._h, cptr(xv), cptr(y), p))
        return y

    def solve_gmres(self, rhs, tolerance: float, max_iterations: int):
        """Returns (x, info). Raises ConvergenceFailure (with .x and .info)."""
        b = _sized(rhs, 3 * self.n, "rhs")
        x = np.empty_like(b)
        info = _lib.SolveInfo()
        st = self._lib.lh_pairsum_solve(self._h, cptr(b), float(tolerance), int(max_iterations),
                                        cptr(x), C.byref(info))
        if st == _lib.LH_CONVERGENCE_FAILURE:
            e = ConvergenceFailure(st, self._lib.lh_last_error().decode())
            e.x, e.info = x, info.as_dict()
            raise e
        check(st)
        return x, info.as_dict()

    def incident(self, inc: IncidentField) -> np.ndarray:
        e = np.empty(3 * self.n, dtype=np.complex128)
        ci = inc._c()
        check(self._lib.lh_pairsum_incident(self._h, C.byref(ci), cptr(e)))
        return e

    def cross_sections(self, p, e_inc, amplitude: float = 1.0):
        out = np.zeros(3)
        pv, ev = _sized(p, 3 * self.n, "p"), _sized(e_inc, 3 * self.n, "e_inc")
        check(self._lib.lh_pairsum_cross_sections(self._h, cptr(pv), cptr(ev), float(amplitude),
                                                  dptr(out)))
        return out

    def compute_cross_sections(self, inc: IncidentField, tolerance: float, max_iterations: int):
        """Returns (out[3], p, info)."""
        out = np.zeros(3)
        p = np.empty(3 * self.n, dtype=np.complex128)
        info = _lib.SolveInfo()
        ci = inc._c()
        st = self._lib.lh_pairsum_compute_cross_sections(self._h, C.byref(ci), float(tolerance),
                                                         int(max_iterations), dptr(out),
                                                         cptr(p), C.byref(info))
        if st == _lib.LH_CONVERGENCE_FAILURE:
            e = ConvergenceFailure(st, self._lib.lh_last_error().decode())
            e.info = info.as_dict()
            raise e
        check(st)
        return out, p, info.as_dict()

    def adjoint_grad(self, inc: IncidentField, tolerance: float, max_iterations: int, dainv,
                     positions: bool = True):
        """C_ext and its gradients through the forward and adjoint solves
        (lh_pairsum_adjoint_grad). ``dainv``: d(alpha^-1)/d eps, (9,) or (N, 9).
        Returns (cross_sections[3], grad_eps (N, 2) = dC/dRe eps, dC/dIm eps,
        grad_pos (N, 3) or None, forward info, adjoint info)."""
        dv = _c128(dainv).reshape(-1)
        uniform = 1 if dv.size == 9 else 0
        if not uniform and dv.size != 9 * self.n:
            raise ValueError("dainv must be (9,) or (N, 9)")
        out = np.zeros(3)
        ge = np.zeros(2 * self.n)
        gp = np.zeros(3 * self.n) if positions else None
        fi, ai = _lib.SolveInfo(), _lib.SolveInfo()
        ci = inc._c()
        st = self._lib.lh_pairsum_adjoint_grad(self._h, C.byref(ci), float(tolerance),
                                               int(max_iterations), cptr(dv), uniform, dptr(out),
                                               dptr(ge), dptr(gp) if positions else None,
                                               C.byref(fi), C.byref(ai))
        if st == _lib.LH_CONVERGENCE_FAILURE:
            e = ConvergenceFailure(st, self._lib.lh_last_error().decode())
            e.info = fi.as_dict()
            e.adjoint_info = ai.as_dict()
            raise e
        check(st)
        return (out, ge.reshape(-1, 2), gp.reshape(-1, 3) if positions else None,
                fi.as_dict(), ai.as_dict())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lh_pairsum_destroy(self._h)
            self._h = None

    __del__ = close


class HipBackend:
    """ComputeBackend for one MI355X device (backend.rs:62-134)."""

    def __init__(self, device: int = 0):
        self._lib = _lib.load()
        h = C.c_void_p()
        check(self._lib.lh_backend_create(int(device), C.byref(h)))
        self._h = h

    def device_info(self) -> dict:
        d = _lib.DeviceInfo()
        check(self._lib.lh_backend_device_info(self._h, C.byref(d)))
        return {"name": d.name.decode(), "arch": d.arch.decode(),
                "backend_type": "Gpu" if d.backend_type == 1 else str(d.backend_type),
                "memory_bytes": int(d.memory_bytes), "compute_units": int(d.compute_units)}

    def create_session(self, n: int) -> MatvecSession:
        return MatvecSession(self, n)

    def parallel_matrix_fill(self, rows: int, cols: int, fill_fn) -> np.ndarray:
        """ComputeBackend::parallel_matrix_fill (backend.rs:78-83): entry
        (i, j) = fill_fn(i, j), row-major. A host-side fill of a caller's
        closure, as the reference's GPU backend does it (gpu.rs:611-632)."""
        out = np.empty((int(rows), int(cols)), dtype=np.complex128)
        for i in range(int(rows)):
            for j in range(int(cols)):
                out[i, j] = fill_fn(i, j)
        return out

    def matvec(self, matrix, vector) -> np.ndarray:
        """ComputeBackend::matvec as the GPU backend gives it (gpu.rs:634-655):
        a one-off product through a temporary stored-matrix session. A
        non-square matrix or a vector of another length is LH_DEVICE_ERROR
        (ComputeError::DeviceError), as there."""
        m = _c128(matrix)
        v = _c128(vector).reshape(-1)
        if m.ndim != 2 or m.shape[0] != m.shape[1] or m.shape[0] != v.size:
            shape = "x".join(str(d) for d in m.shape)
            raise LuminaError(_lib.LH_DEVICE_ERROR,
                              f"Dimension mismatch: matrix {shape}, vector {v.size}")
        s = MatvecSession(self, m.shape[0])
        try:
            s.upload_matrix(m)
            return s.matvec(v)
        finally:
            s.close()

    def dense_solve(self, matrix, rhs):
        """ComputeBackend::dense_solve (backend.rs:119-134): the GPU backends
        keep the trait's default, ComputeError::Unavailable."""
        raise LuminaError(_lib.LH_UNAVAILABLE, "Dense solve not implemented for this backend")

    def create_pairsum_session(self, *a, **kw) -> PairSumSession:
        return PairSumSession(self, *a, **kw)

    def assemble_and_create_session(self, positions, k: float, diagonal_blocks) -> MatvecSession:
        """A stored-matrix session holding A = blockdiag(diagonal_blocks) - G,
        assembled on the device (assemble_and_create_session,
        backend.rs:109-117; gpu.rs:508-607). ``diagonal_blocks``: (N, 3, 3)
        or (N, 9) alpha^-1 blocks, row-major."""
        pos = _f64(positions).reshape(-1, 3)
        blocks = _c128(diagonal_blocks).reshape(-1, 9)
        if blocks.shape[0] != pos.shape[0]:
            raise ValueError("one diagonal block per dipole")
        h = C.c_void_p()
        check(self._lib.lh_session_assemble(self._h, dptr(pos), pos.shape[0], float(k),
                                            cptr(blocks), C.byref(h)))
        return MatvecSession(self, 3 * pos.shape[0], _handle=h)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lh_backend_destroy(self._h)
            self._h = None

    __del__ = close


class CdaSolver:
    """CdaSolver (cda/mod.rs:44-96) on the MI355X on-the-fly GMRES path.

    ``iterative_threshold`` and ``matrix_memory_budget`` are accepted for
    interface parity; the device path is always GMRES with the on-the-fly
    matvec (the dispatch of cda/mod.rs:271-391 forced to :329-345/:375-391).
    """

    def __init__(self, backend: HipBackend, iterative_threshold: int = 0,
                 incident_field: IncidentField | None = None, use_fcd: bool = True,
                 cell_size: float = 3.0, lattice: LatticeSpec | None = None,
                 precision: int = FP64, matrix_memory_budget: int = 0):
        self.backend = backend
        self.iterative_threshold = iterative_threshold
        self.incident_field = incident_field or IncidentField()
        self.use_fcd = use_fcd
        self.cell_size = cell_size
        self.lattice = lattice
        self.precision = precision
        self.matrix_memory_budget = matrix_memory_budget

    def _session(self, positions, polarisability, wavelength_nm, params):
        k = wavenumber(wavelength_nm, params.environment_n)
        return PairSumSession(self.backend, positions, polarisability, k, self.use_fcd,
                              self.cell_size, params.substrate_runtime, self.precision,
                              self.lattice, params.k_bloch)

    def compute_cross_sections(self, positions, polarisability, wavelength_nm: float,
                               params: SimulationParams) -> CrossSections:
        if len(positions) == 0:
            raise LuminaError(_lib.LH_INVALID_ARGUMENT, "No dipoles provided")
        s = self._session(positions, polarisability, wavelength_nm, params)
        try:
            out, _, info = s.compute_cross_sections(self.incident_field, params.solver_tolerance,
                                                    params.max_iterations)
        finally:
            s.close()
        cs = CrossSections(wavelength_nm, float(out[0]), float(out[1]), float(out[2]))
        cs.info = info
        return cs

    def solve_dipoles(self, positions, polarisability, wavelength_nm: float,
                      params: SimulationParams) -> DipoleResponse:
        if len(positions) == 0:
            raise LuminaError(_lib.LH_INVALID_ARGUMENT, "No dipoles provided")
        s = self._session(positions, polarisability, wavelength_nm, params)
        try:
            out, p, info = s.compute_cross_sections(self.incident_field, params.solver_tolerance,
                                                    params.max_iterations)
        finally:
            s.close()
        n = len(positions)
        moments = p.reshape(n, 3)
        pol = _c128(polarisability)
        blocks = np.broadcast_to(pol.reshape(-1, 3, 3), (n, 3, 3))
        ainv = np.linalg.inv(blocks)
        local = np.einsum("nab,nb->na", ainv, moments)
        return DipoleResponse(wavelength_nm, params.environment_n, moments, local, info)

    def _response_session(self, positions, polarisability, k: float):
        if len(positions) == 0:
            raise LuminaError(_lib.LH_INVALID_ARGUMENT, "No dipoles provided")
        return PairSumSession(self.backend, positions, polarisability, k, self.use_fcd,
                              self.cell_size, None, FP64)

    def cross_sections_from_response(self, positions, polarisability, response: DipoleResponse,
                                     params: SimulationParams) -> CrossSections:
        """cda/mod.rs:123-166: C_ext, C_abs, C_sca of already solved moments
        (no solve), k from the response's wavelength and params.environment_n,
        the device reductions of lh_pairsum_cross_sections."""
        k = wavenumber(response.wavelength_nm, params.environment_n)
        s = self._response_session(positions, polarisability, k)
        try:
            e_inc = s.incident(self.incident_field)
            out = s.cross_sections(np.asarray(response.moments).reshape(-1), e_inc,
                                   self.incident_field.amplitude)
        finally:
            s.close()
        return CrossSections(response.wavelength_nm, float(out[0]), float(out[1]), float(out[2]))

    def compute_near_field(self, positions, polarisability, response: DipoleResponse, plane):
        """cda/mod.rs:421-430 -> fields.rs:51-108 on the device
        (lh_pairsum_field_at_points): |E|^2 on the plane's grid, k from the
        response's wavelength and medium, this solver's incident field."""
        from .fields import compute_near_field_map
        k = wavenumber(response.wavelength_nm, response.environment_n)
        s = self._response_session(positions, polarisability, k)
        try:
            return compute_near_field_map(s, plane, self.incident_field,
                                          np.asarray(response.moments).reshape(-1))
        finally:
            s.close()

    def compute_far_field(self, positions, polarisability, response: DipoleResponse,
                          n_theta: int, n_phi: int):
        """cda/mod.rs:432-441 -> fields.rs:156-221 on the device
        (lh_pairsum_far_field)."""
        from .fields import compute_far_field
        k = wavenumber(response.wavelength_nm, response.environment_n)
        s = self._response_session(positions, polarisability, k)
        try:
            return compute_far_field(s, n_theta, n_phi, np.asarray(response.moments).reshape(-1),
                                     response.wavelength_nm)
        finally:
            s.close()

    def method_name(self) -> str:
        """cda/mod.rs:443-445."""
        return "Coupled Dipole Approximation (CDA)"


class Batch:
    """A device-resident batch of compute_cross_sections problems
    (lh_batch_create / lh_batch_run / lh_batch_results).

    ``problems``: list of dicts with keys positions, polarisability, k,
    use_fcd, cell_size, incident (IncidentField), substrate (optional),
    lattice (LatticeSpec, optional), k_bloch (optional) and
    classify_positions (optional: positions of the FCD decision).
    Problems passing the same ``positions`` array object share its geometry.
    """

    def __init__(self, backend: HipBackend, problems):
        self._lib = backend._lib
        S = len(problems)
        arr = (_lib.BatchProblem * S)()
        keep = []
        pos_cache = {}
        for q, pr in enumerate(problems):
            pid = id(pr["positions"])
            if pid not in pos_cache:
                pos_cache[pid] = _f64(pr["positions"]).reshape(-1, 3)
            sp, kp = _spec(pos_cache[pid], pr["polarisability"], pr["k"], pr.get("use_fcd", True),
                           pr.get("cell_size", 3.0), pr.get("substrate"), FP64,
                           pr.get("lattice"), pr.get("k_bloch", (0.0, 0.0, 0.0)),
                           pr.get("classify_positions"))
            keep.append(kp)
            arr[q].spec = sp
            arr[q].incident = pr.get("incident", IncidentField())._c()
        self.sizes = [len(pos_cache[id(pr["positions"])]) for pr in problems]
        h = C.c_void_p()
        check(self._lib.lh_batch_create(backend._h, arr, S, C.byref(h)))
        self._h = h
        self.count = S

    def run(self, tolerance: float, max_iterations: int, precision: int = MIXED):
        check(self._lib.lh_batch_run(self._h, float(tolerance), int(max_iterations),
                                     int(precision)))

    def results(self, want_moments: bool = False):
        """(cross_sections (S,3), infos, moments or None). Raises
        ConvergenceFailure (with .cross_sections, .infos) if any failed."""
        S = self.count
        cs = np.zeros((S, 3))
        infos = (_lib.SolveInfo * S)()
        moments, mptr = None, None
        if want_moments:
            moments = [np.empty(3 * n, dtype=np.complex128) for n in self.sizes]
            mptr = (C.c_void_p * S)(*[m.ctypes.data for m in moments])
        st = self._lib.lh_batch_results(self._h, dptr(cs), mptr, infos)
        info_list = [infos[q].as_dict() for q in range(S)]
        if st == _lib.LH_CONVERGENCE_FAILURE:
            e = ConvergenceFailure(st, self._lib.lh_last_error().decode())
            e.cross_sections, e.infos, e.moments = cs, info_list, moments
            raise e
        check(st)
        return cs, info_list, moments

    def set_timing(self, on: bool = True):
        check(self._lib.lh_batch_set_timing(self._h, 1 if on else 0))

    def set_graph_cycles(self, cycles: int):
        """Restart cycles per graph launch of run(): 0 adaptive (1, 2, 4,
        then 8), N > 0 fixed, -1 the whole solve in one launch."""
        check(self._lib.lh_batch_set_graph_cycles(self._h, int(cycles)))

    def kernel_times(self) -> dict:
        t = _lib.KernelTimes()
        check(self._lib.lh_batch_kernel_times(self._h, C.byref(t)))
        return t.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lh_batch_destroy(self._h)
            self._h = None

    __del__ = close


def batch_compute(backend: HipBackend, problems, tolerance: float, max_iterations: int,
                  precision: int = MIXED, want_moments: bool = False):
    """Many compute_cross_sections calls in one batched device solve.
    Returns (cross_sections (S,3), infos list, moments list or None)."""
    b = Batch(backend, problems)
    try:
        b.run(tolerance, max_iterations, precision)
        return b.results(want_moments)
    finally:
        b.close()
