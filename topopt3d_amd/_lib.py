"""ctypes binding of libtofe.so, the C ABI declared in include/tofe.h.

The library must have been built in-tree (``python topopt3d_amd/build.py`` or
``__graft_entry__.build()``).  There is no CPU fallback: loading fails loudly when the
HIP library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libtofe.so")

TOFE_OK = 0
TOFE_ERR_ARG, TOFE_ERR_HIP, TOFE_ERR_STATE, TOFE_ERR_NUMERIC, TOFE_ERR_NOMEM = -1, -2, -3, -4, -5

CONVERGED_RTOL = 2
CONVERGED_ATOL = 3
DIVERGED_ITS = -3
DIVERGED_DTOL = -4
DIVERGED_INDEFINITE_PC = -8
DIVERGED_NANORINF = -9
DIVERGED_INDEFINITE_MAT = -10

SMOOTHER_CHEBYSHEV, SMOOTHER_JACOBI = 0, 1
MASS_LUMPED, MASS_CONSISTENT = 0, 1
PREC_FP64, PREC_MIXED = 0, 1

(FIELD_X, FIELD_RHO, FIELD_E, FIELD_DEDZ, FIELD_ES, FIELD_DESDZ, FIELD_V, FIELD_DIAG, FIELD_Y,
 FIELD_RHOQ, FIELD_REDGE, FIELD_MDIAG) = range(12)
INTERP_SIMP, INTERP_SIMP_CUT, INTERP_SIMP_LOGISTIC = 0, 1, 2

# Every entry point include/tofe.h declares.
EXPORTS = (
    "tofe_create", "tofe_destroy", "tofe_last_error", "tofe_sizes", "tofe_get_ke",
    "tofe_set_design", "tofe_set_design_dev", "tofe_set_E", "tofe_get_field", "tofe_assemble",
    "tofe_solve", "tofe_solve_dev", "tofe_apply_K", "tofe_apply_K_dev", "tofe_pc_apply",
    "tofe_pc_apply_dev", "tofe_compliance", "tofe_compliance_dev", "tofe_sensitivity",
    "tofe_filter", "tofe_filter_T", "tofe_filter_pattern", "tofe_set_precision",
    "tofe_get_stats", "tofe_get_level_stencil", "tofe_get_level_lambda", "tofe_time_operator",
    "tofe_time_kernels",
    "tofe_set_penalty", "tofe_chain_filter", "tofe_assemble_mass", "tofe_apply_M",
    "tofe_apply_M_dev", "tofe_element_quadratic", "tofe_petsc_vec_write", "tofe_petsc_vec_read",
    "tofe_petsc_aij_write", "tofe_petsc_aij_read", "tofe_mesh_out", "tofe_print_vals",
    "tofe_set_stress", "tofe_set_stress_dev", "tofe_apply_Ks", "tofe_apply_Ks_dev",
    "tofe_stability_terms", "tofe_element_bilinear", "tofe_compliance_noalloc",
    "tofe_set_mass_form", "tofe_autotune", "tofe_batch_create", "tofe_batch_destroy",
    "tofe_batch_last_error", "tofe_batch_solve_dev", "tofe_batch_get_stats", "tofe_set_springs",
)


class TofeDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("hx", C.c_double), ("hy", C.c_double), ("hz", C.c_double),
        ("E0", C.c_double), ("nu", C.c_double), ("Emin", C.c_double), ("penal", C.c_double),
        ("rmin", C.c_double),
        ("mg_levels", C.c_int), ("smoother", C.c_int), ("nu_pre", C.c_int), ("nu_post", C.c_int),
        ("cheb_lo", C.c_double), ("cheb_hi", C.c_double),
        ("eig_its", C.c_int), ("precision", C.c_int),
        ("fixed_dof", C.c_void_p),
        ("rmax", C.c_double), ("vd_min", C.c_double), ("vd_penal", C.c_double),
        ("interpolation", C.c_int), ("interp_param", C.c_double * 2),
    ]


class TofeStats(C.Structure):
    _fields_ = [
        ("levels", C.c_int), ("coarse_size", C.c_int64), ("its", C.c_int), ("reason", C.c_int),
        ("rnorm", C.c_double), ("bnorm", C.c_double),
        ("setup_seconds", C.c_double), ("solve_seconds", C.c_double), ("fine_form", C.c_int),
        ("coarse_form", C.c_int),
    ]


class TofeKernelTime(C.Structure):
    _fields_ = [("level", C.c_int), ("precision", C.c_int), ("epilogue", C.c_int),
                ("count", C.c_int), ("total_ms", C.c_double)]


class TofeBatchStats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("capture_seconds", C.c_double),
                ("replay_seconds", C.c_double), ("replays", C.c_int), ("iterations", C.c_int),
                ("launches", C.c_int), ("unbatched", C.c_int), ("instance_launches", C.c_int)]


class TofeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tofe error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libtofe.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python topopt3d_amd/build.py or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, dp, i64p, i32p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_int32)
    ip = C.POINTER(C.c_int)
    sig = {
        "tofe_create": ([C.POINTER(TofeDesc), C.POINTER(vp)], C.c_int),
        "tofe_destroy": ([vp], C.c_int),
        "tofe_last_error": ([vp], C.c_char_p),
        "tofe_sizes": ([vp, i64p, i64p, i64p, ip], C.c_int),
        "tofe_get_ke": ([vp, dp], C.c_int),
        "tofe_set_design": ([vp, dp], C.c_int),
        "tofe_set_design_dev": ([vp, vp], C.c_int),
        "tofe_set_E": ([vp, dp], C.c_int),
        "tofe_get_field": ([vp, C.c_int, dp], C.c_int),
        "tofe_assemble": ([vp], C.c_int),
        "tofe_solve": ([vp, dp, dp, C.c_int, C.c_double, C.c_int, ip, ip], C.c_int),
        "tofe_solve_dev": ([vp, vp, vp, C.c_int, C.c_double, C.c_int, ip, ip], C.c_int),
        "tofe_apply_K": ([vp, dp, dp, C.c_int], C.c_int),
        "tofe_apply_K_dev": ([vp, vp, vp, C.c_int], C.c_int),
        "tofe_pc_apply": ([vp, dp, dp, C.c_int], C.c_int),
        "tofe_pc_apply_dev": ([vp, vp, vp, C.c_int], C.c_int),
        "tofe_compliance": ([vp, dp, dp, dp, dp], C.c_int),
        "tofe_compliance_dev": ([vp, vp, vp, dp, vp], C.c_int),
        "tofe_sensitivity": ([vp, dp, dp], C.c_int),
        "tofe_filter": ([vp, dp, dp], C.c_int),
        "tofe_filter_T": ([vp, dp, dp], C.c_int),
        "tofe_filter_pattern": ([vp, i64p, i32p, dp], C.c_int),
        "tofe_set_precision": ([vp, C.c_int], C.c_int),
        "tofe_get_stats": ([vp, C.POINTER(TofeStats)], C.c_int),
        "tofe_get_level_stencil": ([vp, C.c_int, dp, i64p], C.c_int),
        "tofe_get_level_lambda": ([vp, C.c_int, dp], C.c_int),
        "tofe_time_operator": ([vp, C.c_int, C.c_int, C.c_int, C.c_int, dp], C.c_int),
        "tofe_time_kernels": ([vp, C.c_int, C.POINTER(TofeKernelTime), C.c_int, ip], C.c_int),
        "tofe_set_penalty": ([vp, C.c_double, C.c_double], C.c_int),
        "tofe_chain_filter": ([vp, dp, dp], C.c_int),
        "tofe_assemble_mass": ([vp, C.c_double, dp], C.c_int),
        "tofe_apply_M": ([vp, dp, dp, C.c_int], C.c_int),
        "tofe_apply_M_dev": ([vp, vp, vp, C.c_int], C.c_int),
        "tofe_element_quadratic": ([vp, dp, C.c_int, dp, dp], C.c_int),
        "tofe_petsc_vec_write": ([C.c_char_p, dp, C.c_int64, C.c_int], C.c_int),
        "tofe_petsc_vec_read": ([C.c_char_p, dp, i64p, C.c_int], C.c_int),
        "tofe_petsc_aij_write": ([C.c_char_p, C.c_int64, C.c_int64, i64p, i32p, dp, C.c_int],
                                 C.c_int),
        "tofe_petsc_aij_read": ([C.c_char_p, i64p, i64p, i64p, i64p, i32p, dp, C.c_int], C.c_int),
        "tofe_mesh_out": ([vp, C.c_char_p, dp, C.c_int], C.c_int),
        "tofe_print_vals": ([vp, C.c_char_p, C.c_char_p, dp, C.c_int], C.c_int),
        "tofe_set_stress": ([vp, dp], C.c_int),
        "tofe_set_stress_dev": ([vp, vp], C.c_int),
        "tofe_apply_Ks": ([vp, dp, dp, C.c_int], C.c_int),
        "tofe_apply_Ks_dev": ([vp, vp, vp, C.c_int], C.c_int),
        "tofe_stability_terms": ([vp, dp, C.c_int, dp, dp], C.c_int),
        "tofe_element_bilinear": ([vp, dp, C.c_int, dp, dp], C.c_int),
        "tofe_set_mass_form": ([vp, C.c_int], C.c_int),
        "tofe_set_springs": ([vp, dp], C.c_int),
        "tofe_autotune": ([vp, C.c_int, ip, i64p], C.c_int),
        "tofe_batch_create": ([C.POINTER(vp), C.c_int, C.POINTER(vp)], C.c_int),
        "tofe_batch_destroy": ([vp], C.c_int),
        "tofe_batch_last_error": ([vp], C.c_char_p),
        "tofe_batch_solve_dev": ([vp, C.POINTER(vp), C.POINTER(vp), C.c_int, C.c_double, C.c_int,
                                  C.c_int, ip, ip], C.c_int),
        "tofe_batch_get_stats": ([vp, C.POINTER(TofeBatchStats)], C.c_int),
        "tofe_compliance_noalloc": ([C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_double, vp, vp, vp, dp, vp],
                                    C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _chk(rc, what):
    if rc != TOFE_OK:
        raise TofeError(rc, what)


def petsc_vec_write(path, v, index_bytes=4):
    v = _f64(v)
    _chk(lib().tofe_petsc_vec_write(os.fsencode(path), _dp(v), v.size, index_bytes), path)


def petsc_vec_read(path, index_bytes=4):
    n = C.c_int64(0)
    _chk(lib().tofe_petsc_vec_read(os.fsencode(path), None, C.byref(n), index_bytes), path)
    v = np.empty(n.value)
    _chk(lib().tofe_petsc_vec_read(os.fsencode(path), _dp(v), C.byref(n), index_bytes), path)
    return v


def petsc_aij_write(path, rows, cols, rowptr, col, val, index_bytes=4):
    rp = np.ascontiguousarray(rowptr, np.int64)
    cl = np.ascontiguousarray(col, np.int32)
    vl = _f64(val)
    _chk(lib().tofe_petsc_aij_write(os.fsencode(path), rows, cols,
                                    rp.ctypes.data_as(C.POINTER(C.c_int64)),
                                    cl.ctypes.data_as(C.POINTER(C.c_int32)), _dp(vl), index_bytes),
         path)


def petsc_aij_read(path, index_bytes=4):
    """(rows, cols, rowptr, col, val) of a PETSc binary AIJ file."""
    m, n, z = C.c_int64(), C.c_int64(), C.c_int64()
    L = lib()
    _chk(L.tofe_petsc_aij_read(os.fsencode(path), C.byref(m), C.byref(n), C.byref(z), None, None,
                               None, index_bytes), path)
    rp = np.empty(m.value + 1, np.int64)
    cl = np.empty(z.value, np.int32)
    vl = np.empty(z.value)
    _chk(L.tofe_petsc_aij_read(os.fsencode(path), C.byref(m), C.byref(n), C.byref(z),
                               rp.ctypes.data_as(C.POINTER(C.c_int64)),
                               cl.ctypes.data_as(C.POINTER(C.c_int32)), _dp(vl), index_bytes), path)
    return m.value, n.value, rp, cl, vl


class Context:
    """One tofe_ctx: a structured Q1 mesh, its filter and multigrid hierarchy on one GPU."""

    def __init__(self, nel, h=(1.0, 1.0, 1.0), E0=1.0, nu=0.3, Emin=1e-10, penal=3.0,
                 rmin=1.5, mg_levels=0, smoother=SMOOTHER_CHEBYSHEV, nu_pre=2, nu_post=2,
                 cheb_lo=0.1, cheb_hi=1.1, eig_its=10, precision=PREC_FP64, fixed_dof=None,
                 rmax=0.0, vd_min=-1.0, vd_penal=1.0, interpolation=INTERP_SIMP,
                 interp_param=(0.0, 0.0)):
        L = lib()
        self.nel = tuple(int(v) for v in nel)
        d = TofeDesc()
        d.nx, d.ny, d.nz = self.nel
        d.hx, d.hy, d.hz = (float(v) for v in h)
        d.E0, d.nu, d.Emin, d.penal, d.rmin = E0, nu, Emin, penal, rmin
        d.mg_levels, d.smoother, d.nu_pre, d.nu_post = mg_levels, smoother, nu_pre, nu_post
        d.cheb_lo, d.cheb_hi, d.eig_its, d.precision = cheb_lo, cheb_hi, eig_its, precision
        d.rmax, d.vd_min, d.vd_penal, d.interpolation = rmax, vd_min, vd_penal, interpolation
        d.interp_param[0], d.interp_param[1] = interp_param
        self._fixed = None
        if fixed_dof is not None:
            self._fixed = np.ascontiguousarray(fixed_dof, dtype=np.uint8)
            d.fixed_dof = self._fixed.ctypes.data
        self.h = C.c_void_p()
        rc = L.tofe_create(C.byref(d), C.byref(self.h))
        if rc != TOFE_OK:
            raise TofeError(rc, "tofe_create failed")
        ne, nn, nd, lv = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int()
        L.tofe_sizes(self.h, C.byref(ne), C.byref(nn), C.byref(nd), C.byref(lv))
        self.nelem, self.nnode, self.ndof, self.levels = ne.value, nn.value, nd.value, lv.value

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().tofe_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != TOFE_OK:
            msg = lib().tofe_last_error(self.h)
            raise TofeError(rc, msg.decode() if msg else "")

    # -- state --------------------------------------------------------------
    def ke(self):
        KE = np.empty(576)
        self._check(lib().tofe_get_ke(self.h, _dp(KE)))
        return KE.reshape(24, 24)

    def set_design(self, x):
        x = _f64(x)
        assert x.size == self.nelem
        self._check(lib().tofe_set_design(self.h, _dp(x)))

    def set_design_dev(self, ptr):
        self._check(lib().tofe_set_design_dev(self.h, C.c_void_p(ptr)))

    def set_E(self, E):
        E = _f64(E)
        assert E.size == self.nelem
        self._check(lib().tofe_set_E(self.h, _dp(E)))

    def field(self, which):
        n = self.ndof if which in (FIELD_DIAG, FIELD_MDIAG) else self.nelem
        out = np.empty(n)
        self._check(lib().tofe_get_field(self.h, which, _dp(out)))
        return out

    def assemble(self):
        self._check(lib().tofe_assemble(self.h))

    def set_precision(self, p):
        self._check(lib().tofe_set_precision(self.h, p))

    # -- solves and operators -------------------------------------------------
    def solve(self, F, U=None, rtol=1e-8, maxit=1000):
        F = _f64(F)
        nrhs = F.size // self.ndof
        U = np.zeros_like(F) if U is None else np.array(U, dtype=np.float64, copy=True)
        its, reason = C.c_int(), C.c_int()
        self._check(lib().tofe_solve(self.h, _dp(F), _dp(U), nrhs, rtol, maxit,
                                     C.byref(its), C.byref(reason)))
        return U, its.value, reason.value

    def solve_dev(self, F_ptr, U_ptr, nrhs=1, rtol=1e-8, maxit=1000):
        its, reason = C.c_int(), C.c_int()
        self._check(lib().tofe_solve_dev(self.h, C.c_void_p(F_ptr), C.c_void_p(U_ptr), nrhs,
                                         rtol, maxit, C.byref(its), C.byref(reason)))
        return its.value, reason.value

    def apply_K(self, X):
        X = _f64(X)
        Y = np.empty_like(X)
        self._check(lib().tofe_apply_K(self.h, _dp(X), _dp(Y), X.size // self.ndof))
        return Y

    def apply_K_dev(self, X_ptr, Y_ptr, nrhs=1):
        self._check(lib().tofe_apply_K_dev(self.h, C.c_void_p(X_ptr), C.c_void_p(Y_ptr), nrhs))

    def pc_apply(self, R):
        R = _f64(R)
        Z = np.empty_like(R)
        self._check(lib().tofe_pc_apply(self.h, _dp(R), _dp(Z), R.size // self.ndof))
        return Z

    def pc_apply_dev(self, R_ptr, Z_ptr, nrhs=1):
        self._check(lib().tofe_pc_apply_dev(self.h, C.c_void_p(R_ptr), C.c_void_p(Z_ptr), nrhs))

    def apply_Ks_dev(self, X_ptr, Y_ptr, nrhs=1):
        self._check(lib().tofe_apply_Ks_dev(self.h, C.c_void_p(X_ptr), C.c_void_p(Y_ptr), nrhs))

    def apply_M_dev(self, X_ptr, Y_ptr, nrhs=1):
        self._check(lib().tofe_apply_M_dev(self.h, C.c_void_p(X_ptr), C.c_void_p(Y_ptr), nrhs))

    def set_stress_dev(self, U_ptr):
        self._check(lib().tofe_set_stress_dev(self.h, C.c_void_p(U_ptr)))

    def compliance(self, U, F, gradient=True):
        U, F = _f64(U), _f64(F)
        c = C.c_double()
        g = np.empty(self.nelem) if gradient else None
        self._check(lib().tofe_compliance(self.h, _dp(U), _dp(F), C.byref(c), _dp(g)))
        return (c.value, g) if gradient else c.value

    def compliance_dev(self, U_ptr, F_ptr, g_ptr=None):
        c = C.c_double()
        self._check(lib().tofe_compliance_dev(self.h, C.c_void_p(U_ptr), C.c_void_p(F_ptr),
                                              C.byref(c), C.c_void_p(g_ptr) if g_ptr else None))
        return c.value

    def sensitivity(self, U):
        U = _f64(U)
        g = np.empty(self.nelem)
        self._check(lib().tofe_sensitivity(self.h, _dp(U), _dp(g)))
        return g

    def filter(self, x):
        x = _f64(x)
        rho = np.empty_like(x)
        self._check(lib().tofe_filter(self.h, _dp(x), _dp(rho)))
        return rho

    def filter_T(self, g):
        g = _f64(g)
        gx = np.empty_like(g)
        self._check(lib().tofe_filter_T(self.h, _dp(g), _dp(gx)))
        return gx

    def assemble_mass(self, density, mlump=None):
        self._check(lib().tofe_assemble_mass(self.h, density,
                                             None if mlump is None else _dp(_f64(mlump))))

    def set_springs(self, spk):
        """Spring stiffness per DOF (None removes them); takes effect at the next assemble."""
        self._check(lib().tofe_set_springs(self.h, None if spk is None else _dp(_f64(spk))))

    def set_mass_form(self, form):
        self._check(lib().tofe_set_mass_form(self.h, form))

    def apply_M(self, X):
        X = _f64(X)
        nrhs = X.size // self.ndof
        Y = np.empty_like(X)
        self._check(lib().tofe_apply_M(self.h, _dp(X), _dp(Y), nrhs))
        return Y

    def element_quadratic(self, phi, mass=True):
        """(kq, mq) of shape (nvec, nelem): phi_e^T KE phi_e and sum of phi_e^2."""
        phi = _f64(phi)
        nvec = phi.size // self.ndof
        kq = np.empty((nvec, self.nelem))
        mq = np.empty((nvec, self.nelem)) if mass else None
        self._check(lib().tofe_element_quadratic(self.h, _dp(phi), nvec, _dp(kq), _dp(mq)))
        return kq, mq

    def mesh_out(self, directory, origin=(0.0, 0.0, 0.0), index_bytes=4):
        o = _f64(origin)
        self._check(lib().tofe_mesh_out(self.h, os.fsencode(directory), _dp(o), index_bytes))

    def print_vals(self, directory, suffix="", U=None, index_bytes=4):
        U = None if U is None else _f64(U)
        self._check(lib().tofe_print_vals(self.h, os.fsencode(directory), suffix.encode(), _dp(U),
                                          index_bytes))

    def set_stress(self, U):
        self._check(lib().tofe_set_stress(self.h, _dp(_f64(U))))

    def apply_Ks(self, X):
        X = _f64(X)
        nrhs = X.size // self.ndof
        Y = np.empty_like(X)
        self._check(lib().tofe_apply_Ks(self.h, _dp(X), _dp(Y), nrhs))
        return Y

    def stability_terms(self, phi, quadratic=True, rhs=True):
        """(pksp (nvec, nelem), dksdu (nvec, ndof)) of buckling modes phi."""
        phi = _f64(phi)
        nvec = phi.size // self.ndof
        q = np.empty((nvec, self.nelem)) if quadratic else None
        r = np.empty((nvec, self.ndof)) if rhs else None
        self._check(lib().tofe_stability_terms(self.h, _dp(phi), nvec, _dp(q), _dp(r)))
        return q, r

    def element_bilinear(self, A, B):
        A, B = _f64(A), _f64(B)
        nvec = A.size // self.ndof
        out = np.empty((nvec, self.nelem))
        self._check(lib().tofe_element_bilinear(self.h, _dp(A), nvec, _dp(B), _dp(out)))
        return out

    def set_penalty(self, penal, vd_penal=1.0):
        self._check(lib().tofe_set_penalty(self.h, penal, vd_penal))

    def chain_filter(self, dfdV=None, dfdE=None):
        """Chain_Filter on copies of dfdV / dfdE (either may be None); returns both."""
        V = None if dfdV is None else _f64(dfdV).copy()
        E = None if dfdE is None else _f64(dfdE).copy()
        self._check(lib().tofe_chain_filter(self.h, _dp(V), _dp(E)))
        return V, E

    def filter_pattern(self):
        rp = np.empty(self.nelem + 1, np.int64)
        self._check(lib().tofe_filter_pattern(self.h, rp.ctypes.data_as(C.POINTER(C.c_int64)),
                                              None, None))
        nnz = int(rp[-1])
        col = np.empty(nnz, np.int32)
        w = np.empty(nnz)
        self._check(lib().tofe_filter_pattern(self.h, rp.ctypes.data_as(C.POINTER(C.c_int64)),
                                              col.ctypes.data_as(C.POINTER(C.c_int32)), _dp(w)))
        return rp, col, w

    def stats(self):
        st = TofeStats()
        self._check(lib().tofe_get_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in TofeStats._fields_}

    def level_stencil(self, level):
        nn = C.c_int64()
        self._check(lib().tofe_get_level_stencil(self.h, level, None, C.byref(nn)))
        out = np.empty(243 * nn.value)
        self._check(lib().tofe_get_level_stencil(self.h, level, _dp(out), C.byref(nn)))
        return out.reshape(27, 3, 3, nn.value)

    def time_operator(self, level=0, precision=PREC_FP64, epilogue=0, nrep=20):
        """Average ms per launch of one level operator kernel (HIP events)."""
        ms = C.c_double()
        self._check(lib().tofe_time_operator(self.h, level, precision, epilogue, nrep, C.byref(ms)))
        return ms.value

    def time_kernels(self, start):
        """start=True: begin timing every operator launch of the following calls;
        start=False: stop and return a list of dicts (level, precision, epilogue, count,
        total_ms), one per kernel kind."""
        n = C.c_int()
        if start:
            self._check(lib().tofe_time_kernels(self.h, 1, None, 0, C.byref(n)))
            return None
        out = (TofeKernelTime * 256)()
        self._check(lib().tofe_time_kernels(self.h, 0, out, 256, C.byref(n)))
        return [{k: getattr(out[j], k) for k, _ in TofeKernelTime._fields_}
                for j in range(min(n.value, 256))]

    def autotune(self, nrep=10):
        """Measured launch shapes for the current precision: (fine z-chunk, split threshold)."""
        z, t = C.c_int(), C.c_int64()
        self._check(lib().tofe_autotune(self.h, nrep, C.byref(z), C.byref(t)))
        return z.value, t.value

    def level_lambda(self, level):
        lam = C.c_double()
        self._check(lib().tofe_get_level_lambda(self.h, level, C.byref(lam)))
        return lam.value


class Batch:
    """tofe_batch over contexts of independent instances: their solves as one HIP graph of
    chunk iterations, replayed until every right-hand side has finished."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        arr = (C.c_void_p * len(self.contexts))(*[c.h.value for c in self.contexts])
        self.h = C.c_void_p()
        rc = lib().tofe_batch_create(arr, len(self.contexts), C.byref(self.h))
        if rc != TOFE_OK:
            raise TofeError(rc, "tofe_batch_create failed")

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().tofe_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve_dev(self, F_ptrs, U_ptrs, nrhs=1, rtol=1e-8, maxit=1000, chunk=4):
        """(its, reasons) as (n, nrhs) arrays."""
        n = len(self.contexts)
        Fa = (C.c_void_p * n)(*F_ptrs)
        Ua = (C.c_void_p * n)(*U_ptrs)
        its = np.zeros(n * nrhs, np.int32)
        rs = np.zeros(n * nrhs, np.int32)
        rc = lib().tofe_batch_solve_dev(self.h, Fa, Ua, nrhs, rtol, maxit, chunk,
                                        its.ctypes.data_as(C.POINTER(C.c_int)),
                                        rs.ctypes.data_as(C.POINTER(C.c_int)))
        if rc != TOFE_OK:
            raise TofeError(rc, lib().tofe_batch_last_error(self.h).decode())
        return its.reshape(n, nrhs), rs.reshape(n, nrhs)

    def stats(self):
        st = TofeBatchStats()
        lib().tofe_batch_get_stats(self.h, C.byref(st))
        return {k: getattr(st, k) for k, _ in TofeBatchStats._fields_}
