"""ctypes binding of the oracle library (oracle/liblumina_oracle.so).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py). Each wrapper names the
C function it calls; the C functions cite the reference file:line they
restate.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liblumina_oracle.so")

_lib = None


def build():
    """Compile the restatement with oracle/Makefile (g++)."""
    subprocess.run(["make", "-s", "-C", HERE, "liblumina_oracle.so"], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        D = C.POINTER(C.c_double)
        _lib.or_greens_point.argtypes = [D, D, C.c_double, D]
        _lib.or_greens_filtered.argtypes = [D, D, C.c_double, C.c_double, D]
        _lib.or_erfc_real.argtypes = [C.c_double]
        _lib.or_erfc_real.restype = C.c_double
        _lib.or_csqrt_pos_imag.argtypes = [D, D]
        _lib.or_ewald.argtypes = [D, D, C.c_double, C.c_int, C.c_int, D, D, C.c_double, C.c_int, D]
        _lib.or_lattice_eta.argtypes = [D, D, C.c_double]
        _lib.or_lattice_eta.restype = C.c_double
        _lib.or_recip_vectors.argtypes = [D, D, D, D]
        _lib.or_invert3.argtypes = [D, D]
        _lib.or_substrate_factor.argtypes = [D, C.c_double, C.c_int, D]
        _lib.or_matvec.argtypes = [C.c_void_p, D, D, C.c_int]
        _lib.or_matvec_rows.argtypes = [C.c_void_p, D, D, C.c_int64, C.c_int64]
        _lib.or_assemble.argtypes = [C.c_void_p, D]
        _lib.or_solve_gmres.argtypes = [C.c_void_p, D, C.c_double, C.c_int64, D, D, C.c_int]
        _lib.or_solve_gmres.restype = C.c_int
        _lib.or_solve_gmres_dense.argtypes = [D, C.c_int64, D, C.c_double, C.c_int64, D, D]
        _lib.or_solve_gmres_dense.restype = C.c_int
        _lib.or_incident.argtypes = [C.c_int64, D, D, D, C.c_double, C.c_double, D]
        _lib.or_cross_sections.argtypes = [C.c_int64, D, C.c_int, D, D, C.c_double, C.c_double, D]
        _lib.or_adjoint_grad.argtypes = [C.c_void_p, D, D, C.c_double, D, D, D, C.c_int, D, D,
                                         C.POINTER(C.c_int64), C.c_int64]
        _lib.or_adjoint_grad.restype = C.c_int
        _lib.or_greens_point_grad.argtypes = [D, D, C.c_double, D]
        _lib.or_field_at_points.argtypes = [C.c_int64, D, D, C.c_double, D, D, C.c_double,
                                            C.c_int64, D, D, D]
        _lib.or_far_field.argtypes = [C.c_int64, D, D, C.c_double, C.c_int64, C.c_int64, D, D, D]
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _v3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(3))


def _t33(buf):
    return buf.view(np.complex128).reshape(3, 3)


def greens_point(r1, r2, k):
    """greens.rs:34-71."""
    out = np.zeros(18)
    load().or_greens_point(_d(_v3(r1)), _d(_v3(r2)), float(k), _d(out))
    return _t33(out)


def greens_filtered(r_obs, r_src, k, cell_size):
    """greens.rs:94-167."""
    out = np.zeros(18)
    load().or_greens_filtered(_d(_v3(r_obs)), _d(_v3(r_src)), float(k), float(cell_size), _d(out))
    return _t33(out)


def erfc_real(x):
    """ewald.rs:316-327."""
    return load().or_erfc_real(float(x))


def csqrt_positive_imag(z):
    """ewald.rs:334-338."""
    a = np.array([complex(z).real, complex(z).imag])
    o = np.zeros(2)
    load().or_csqrt_pos_imag(_d(a), _d(o))
    return complex(o[0], o[1])


def ewald(a1, a2, r, k_bloch, k, which=0, eta=0.0, n_real=5, n_recip=5):
    """ewald.rs:91-273 (which: 0 evaluate, 1 real_space, 2 recip_space)."""
    out = np.zeros(18)
    load().or_ewald(_d(_v3(a1)), _d(_v3(a2)), float(eta), int(n_real), int(n_recip), _d(_v3(r)),
                    _d(_v3(k_bloch)), float(k), int(which), _d(out))
    return _t33(out)


def lattice_eta(a1, a2, eta=0.0):
    """lattice.rs:96-102."""
    return load().or_lattice_eta(_d(_v3(a1)), _d(_v3(a2)), float(eta))


def recip_vectors(a1, a2):
    """lattice.rs:113-147."""
    b1, b2 = np.zeros(3), np.zeros(3)
    load().or_recip_vectors(_d(_v3(a1)), _d(_v3(a2)), _d(b1), _d(b2))
    return b1, b2


def invert_3x3(m):
    """assembly.rs:335-363."""
    a = np.ascontiguousarray(np.asarray(m, dtype=np.complex128).reshape(9))
    o = np.zeros(9, dtype=np.complex128)
    load().or_invert3(_d(a.view(np.float64)), _d(o.view(np.float64)))
    return o


def substrate_factor(eps_sub, eps_env, retarded=True):
    """substrate.rs:160-197."""
    a = np.array([complex(eps_sub).real, complex(eps_sub).imag])
    o = np.zeros(2)
    load().or_substrate_factor(_d(a), float(eps_env), 1 if retarded else 0, _d(o))
    return complex(o[0], o[1])


class _Spec(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("pos", C.POINTER(C.c_double)), ("alpha", C.POINTER(C.c_double)),
        ("alpha_uniform", C.c_int), ("k", C.c_double), ("use_fcd", C.c_int),
        ("cell_size", C.c_double), ("has_lattice", C.c_int), ("a1", C.c_double * 3),
        ("a2", C.c_double * 3), ("eta", C.c_double), ("n_real", C.c_int), ("n_recip", C.c_int),
        ("kb", C.c_double * 3), ("has_sub", C.c_int), ("z_interface", C.c_double),
        ("delta", C.c_double * 2), ("cls_pos", C.POINTER(C.c_double)),
    ]


class Problem:
    """The arguments of matvec_on_the_fly (assembly.rs:209-218)."""

    def __init__(self, positions, polarisability, k, use_fcd=True, cell_size=3.0,
                 substrate=None, lattice=None, k_bloch=(0.0, 0.0, 0.0), classify_positions=None):
        self.pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1, 3))
        pol = np.ascontiguousarray(np.asarray(polarisability, dtype=np.complex128))
        self.n = self.pos.shape[0]
        self.uniform = 1 if pol.size == 9 else 0
        self.pol = pol.reshape(-1)
        s = _Spec()
        s.n = self.n
        s.pos = _d(self.pos)
        s.alpha = _d(self.pol.view(np.float64))
        s.alpha_uniform = self.uniform
        s.k = float(k)
        s.use_fcd = 1 if use_fcd else 0
        s.cell_size = float(cell_size)
        if lattice is not None:
            s.has_lattice = 1
            for i in range(3):
                s.a1[i] = lattice["a1"][i]
                s.a2[i] = lattice["a2"][i]
            s.eta = float(lattice.get("eta", 0.0) or 0.0)
            s.n_real = int(lattice.get("n_real", 5))
            s.n_recip = int(lattice.get("n_recip", 5))
            for i in range(3):
                s.kb[i] = float(k_bloch[i])
        if substrate is not None:
            s.has_sub = 1
            s.z_interface = float(substrate[0])
            d = complex(substrate[1])
            s.delta[0], s.delta[1] = d.real, d.imag
        if classify_positions is not None:
            # FCD near/far decisions taken on these positions (finite differences)
            self._cls = np.ascontiguousarray(np.asarray(classify_positions,
                                                        dtype=np.float64).reshape(-1, 3))
            s.cls_pos = _d(self._cls)
        self.spec = s
        self.k = float(k)

    def matvec(self, x, nthreads=0):
        xv = np.ascontiguousarray(np.asarray(x, dtype=np.complex128).reshape(-1))
        y = np.zeros_like(xv)
        nt = nthreads or os.cpu_count() or 1
        load().or_matvec(C.byref(self.spec), _d(xv.view(np.float64)), _d(y.view(np.float64)), nt)
        return y

    def matvec_rows(self, x, lo, hi):
        xv = np.ascontiguousarray(np.asarray(x, dtype=np.complex128).reshape(-1))
        y = np.zeros(3 * (hi - lo), dtype=np.complex128)
        load().or_matvec_rows(C.byref(self.spec), _d(xv.view(np.float64)),
                              _d(y.view(np.float64)), int(lo), int(hi))
        return y

    def assemble(self):
        dim = 3 * self.n
        A = np.zeros((dim, dim), dtype=np.complex128)
        load().or_assemble(C.byref(self.spec), _d(A.view(np.float64)))
        return A

    def solve_gmres(self, rhs, tol, max_iter, nthreads=0):
        b = np.ascontiguousarray(np.asarray(rhs, dtype=np.complex128).reshape(-1))
        x = np.zeros_like(b)
        info = np.zeros(5)
        nt = nthreads or os.cpu_count() or 1
        st = load().or_solve_gmres(C.byref(self.spec), _d(b.view(np.float64)), float(tol),
                                   int(max_iter), _d(x.view(np.float64)), _d(info), nt)
        return x, {"iterations": int(info[0]), "cycles": int(info[1]), "matvecs": int(info[2]),
                   "residual": float(info[3]), "status": int(st)}

    def incident(self, direction=(0, 0, 1), polarisation=(1, 0, 0), amplitude=1.0):
        """types.rs:60-74 / assembly.rs:178-194."""
        pol = np.array([[complex(p).real, complex(p).imag] for p in polarisation]).reshape(6)
        out = np.zeros(3 * self.n, dtype=np.complex128)
        load().or_incident(self.n, _d(self.pos), _d(_v3(direction)), _d(np.ascontiguousarray(pol)),
                           float(amplitude), self.k, _d(out.view(np.float64)))
        return out

    def cross_sections(self, p, e_inc, amplitude=1.0):
        """cda/mod.rs:187-235."""
        pv = np.ascontiguousarray(np.asarray(p, dtype=np.complex128).reshape(-1))
        ev = np.ascontiguousarray(np.asarray(e_inc, dtype=np.complex128).reshape(-1))
        out = np.zeros(3)
        load().or_cross_sections(self.n, _d(self.pol.view(np.float64)), self.uniform,
                                 _d(pv.view(np.float64)), _d(ev.view(np.float64)), self.k,
                                 float(amplitude), _d(out))
        return out

    def compute_cross_sections(self, tol, max_iter, direction=(0, 0, 1),
                               polarisation=(1, 0, 0), amplitude=1.0, nthreads=0):
        """cda/mod.rs:175-236 on the on-the-fly GMRES path."""
        e = self.incident(direction, polarisation, amplitude)
        p, info = self.solve_gmres(e, tol, max_iter, nthreads)
        return self.cross_sections(p, e, amplitude), p, info


    def adjoint_grad(self, p, lam, dainv, direction=(0, 0, 1), polarisation=(1, 0, 0),
                     amplitude=1.0, rows=None):
        """C_ext gradients from the forward moments p and the adjoint
        solution lam (A lam = conj(E)): (grad_eps (N, 2) as d/dRe, d/dIm;
        grad_r (N, 3) or None without free space). ``rows`` limits the
        position gradients to those dipoles (the others are left at 0)."""
        pv = np.ascontiguousarray(np.asarray(p, dtype=np.complex128).reshape(-1))
        lv = np.ascontiguousarray(np.asarray(lam, dtype=np.complex128).reshape(-1))
        dv = np.ascontiguousarray(np.asarray(dainv, dtype=np.complex128).reshape(-1))
        uniform = 1 if dv.size == 9 else 0
        pol = np.array([[complex(q).real, complex(q).imag] for q in polarisation]).reshape(6)
        ge = np.zeros(2 * self.n)
        gr = np.zeros(3 * self.n)
        if rows is None:
            rp, nrows = None, 0
        else:
            self._rows = np.ascontiguousarray(np.asarray(rows, dtype=np.int64))
            rp, nrows = self._rows.ctypes.data_as(C.POINTER(C.c_int64)), len(self._rows)
        st = load().or_adjoint_grad(C.byref(self.spec), _d(_v3(direction)),
                                    _d(np.ascontiguousarray(pol)), float(amplitude),
                                    _d(pv.view(np.float64)), _d(lv.view(np.float64)),
                                    _d(dv.view(np.float64)), uniform, _d(ge), _d(gr),
                                    rp, nrows)
        return ge.reshape(-1, 2), (gr.reshape(-1, 3) if st == 0 else None)


def field_at_points(positions, moments, k, points, direction=(0, 0, 1), polarisation=(1, 0, 0),
                    amplitude=1.0):
    """fields.rs:18-48 at each point: (E (m, 3) complex, |E|^2 (m,))."""
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1, 3))
    mv = np.ascontiguousarray(np.asarray(moments, dtype=np.complex128).reshape(-1))
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
    pol = np.array([[complex(q).real, complex(q).imag] for q in polarisation]).reshape(6)
    f = np.zeros(3 * len(pts), dtype=np.complex128)
    inten = np.zeros(len(pts))
    load().or_field_at_points(len(pos), _d(pos), _d(mv.view(np.float64)), float(k),
                              _d(_v3(direction)), _d(np.ascontiguousarray(pol)), float(amplitude),
                              len(pts), _d(pts), _d(f.view(np.float64)), _d(inten))
    return f.reshape(-1, 3), inten


def far_field(positions, moments, k, n_theta, n_phi):
    """fields.rs:156-221: (theta, phi, intensity), each n_theta * n_phi."""
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1, 3))
    mv = np.ascontiguousarray(np.asarray(moments, dtype=np.complex128).reshape(-1))
    m = n_theta * n_phi
    th, ph, inten = np.zeros(m), np.zeros(m), np.zeros(m)
    load().or_far_field(len(pos), _d(pos), _d(mv.view(np.float64)), float(k), int(n_theta),
                        int(n_phi), _d(th), _d(ph), _d(inten))
    return th, ph, inten


def greens_point_grad(ro, rs, k):
    """d G_ab / d (r_o)_c as a (3, 3, 3) complex array [c, a, b]."""
    out = np.zeros(54)
    load().or_greens_point_grad(_d(_v3(ro)), _d(_v3(rs)), float(k), _d(out))
    return out.view(np.complex128).reshape(3, 3, 3)


def solve_gmres_dense(A, rhs, tol, max_iter):
    """iterative.rs:33-160 with the matrix.dot closure of its tests."""
    A = np.ascontiguousarray(np.asarray(A, dtype=np.complex128))
    b = np.ascontiguousarray(np.asarray(rhs, dtype=np.complex128).reshape(-1))
    x = np.zeros_like(b)
    info = np.zeros(5)
    st = load().or_solve_gmres_dense(_d(A.view(np.float64)), b.size, _d(b.view(np.float64)),
                                     float(tol), int(max_iter), _d(x.view(np.float64)), _d(info))
    return x, {"iterations": int(info[0]), "cycles": int(info[1]), "matvecs": int(info[2]),
               "residual": float(info[3]), "status": int(st)}
