"""ctypes binding of the CPU oracle (oracle/libtofe_oracle.so).

TEST INFRASTRUCTURE ONLY.  The oracle restates TopOpt3D's FE-analysis path in
plain C (see tofe_oracle.c for the reference file:line each function follows).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker -- never as the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtofe_oracle.so")

_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    args = ["make", "-C", _HERE]
    if force:
        args.append("-B")
    subprocess.run(args, check=True, capture_output=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.orc_dmat.argtypes = [C.c_double, C.c_double, _f64p]
    L.orc_ke.argtypes = [C.c_double] * 5 + [_f64p]
    L.orc_filter_build.restype = C.c_int64
    L.orc_filter_build.argtypes = [C.c_int] * 3 + [C.c_double] * 4 + [C.c_int, vp, vp, vp]
    L.orc_filter_csr.restype = vp
    L.orc_filter_csr.argtypes = [C.c_int] * 3 + [C.c_double] * 4 + [C.c_int]
    L.orc_rfilter_csr.restype = vp
    L.orc_rfilter_csr.argtypes = [C.c_int] * 3 + [C.c_double] * 4 + [vp]
    L.orc_matint.argtypes = [vp, vp, vp, _f64p, C.c_double, C.c_double, C.c_double, C.c_int,
                             _f64p, C.c_double] + [_f64p] * 7
    L.orc_chain_filter.argtypes = [vp, vp, _f64p, _f64p, _f64p, vp, vp]
    L.orc_vecpow1.restype = C.c_double
    L.orc_vecpow1.argtypes = [C.c_double, C.c_double]
    L.orc_simp.argtypes = [C.c_int64, _f64p, C.c_double, C.c_double, vp, vp, vp, vp]
    L.orc_csr_free.argtypes = [vp]
    L.orc_csr_info.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.orc_csr_export.argtypes = [vp, _i64p, _i32p, _f64p]
    L.orc_csr_import.restype = vp
    L.orc_csr_import.argtypes = [C.c_int64, C.c_int64, _i64p, _i32p, _f64p]
    L.orc_spmv.argtypes = [vp, _f64p, _f64p]
    L.orc_spmv_t.argtypes = [vp, _f64p, _f64p]
    L.orc_transpose.restype = vp
    L.orc_transpose.argtypes = [vp]
    L.orc_spgemm.restype = vp
    L.orc_spgemm.argtypes = [vp, vp]
    L.orc_ptap.restype = vp
    L.orc_ptap.argtypes = [vp, vp]
    L.orc_assemble_K.restype = vp
    L.orc_assemble_K.argtypes = [C.c_int] * 3 + [_f64p, _f64p, vp]
    L.orc_build_P.restype = vp
    L.orc_build_P.argtypes = [C.c_int] * 3
    L.orc_mg_create.restype = vp
    L.orc_mg_create.argtypes = [vp] + [C.c_int] * 4 + [C.c_int] * 3 + [C.c_double] * 2 + [C.c_int]
    L.orc_mg_levels.restype = C.c_int
    L.orc_mg_levels.argtypes = [vp]
    L.orc_mg_A.restype = vp
    L.orc_mg_A.argtypes = [vp, C.c_int]
    L.orc_mg_P.restype = vp
    L.orc_mg_P.argtypes = [vp, C.c_int]
    L.orc_mg_lambda.restype = C.c_double
    L.orc_mg_lambda.argtypes = [vp, C.c_int]
    L.orc_mg_dims.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    L.orc_mg_vcycle.argtypes = [vp, _f64p, _f64p]
    L.orc_pcg.restype = C.c_int
    L.orc_pcg.argtypes = [vp, _f64p, _f64p, C.c_double, C.c_int, C.POINTER(C.c_int),
                          C.POINTER(C.c_double)]
    L.orc_mg_free.argtypes = [vp]
    L.orc_mg_setup_seconds.restype = C.c_double
    L.orc_mg_setup_seconds.argtypes = [vp]
    L.orc_dot.restype = C.c_double
    L.orc_dot.argtypes = [C.c_int64, _f64p, _f64p]
    L.orc_assemble_Ks.restype = vp
    L.orc_assemble_Ks.argtypes = [C.c_int] * 3 + [C.c_double] * 5 + [_f64p, _f64p, vp]
    L.orc_ks_quadratic.argtypes = [C.c_int] * 3 + [C.c_double] * 5 + [_f64p, _f64p, C.c_int, _f64p]
    L.orc_stability_rhs.argtypes = [C.c_int] * 3 + [C.c_double] * 5 + [_f64p, _f64p, C.c_int, vp,
                                                                       _f64p]
    L.orc_mass_diag.argtypes = [C.c_int] * 3 + [_f64p, C.c_double, vp, vp, _f64p]
    L.orc_sensitivity.argtypes = [C.c_int] * 3 + [_f64p, _f64p, _f64p, _f64p]
    L.orc_power_start.argtypes = [C.c_int64, _f64p]
    L.orc_num_threads.restype = C.c_int
    L.orc_lanczos_lambda.restype = C.c_double
    L.orc_lanczos_lambda.argtypes = [vp, _f64p, C.c_int]
    L.orc_tridiag_max_eig.restype = C.c_double
    L.orc_tridiag_max_eig.argtypes = [C.c_int, _f64p, _f64p]
    L.orc_smooth.argtypes = [vp, _f64p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                             C.c_int, _f64p, _f64p, _f64p, _f64p]
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class CSR:
    """Owning handle of an orc_csr."""

    def __init__(self, handle, owned=True):
        if not handle:
            raise RuntimeError("oracle returned a NULL matrix")
        self.h = handle
        self.owned = owned

    def __del__(self):
        if getattr(self, "owned", False) and self.h and _lib is not None:
            _lib.orc_csr_free(self.h)
            self.h = None

    @property
    def shape(self):
        n, m, nnz = C.c_int64(), C.c_int64(), C.c_int64()
        lib().orc_csr_info(self.h, C.byref(n), C.byref(m), C.byref(nnz))
        return n.value, m.value, nnz.value

    def arrays(self):
        n, m, nnz = self.shape
        rp = np.empty(n + 1, np.int64)
        col = np.empty(max(nnz, 1), np.int32)
        val = np.empty(max(nnz, 1), np.float64)
        lib().orc_csr_export(self.h, rp, col, val)
        return rp, col[:nnz], val[:nnz]

    def to_scipy(self):
        import scipy.sparse as sp

        n, m, _ = self.shape
        rp, col, val = self.arrays()
        return sp.csr_matrix((val, col, rp), shape=(n, m))

    @classmethod
    def from_scipy(cls, A):
        A = A.tocsr()
        A.sort_indices()
        rp = np.ascontiguousarray(A.indptr, np.int64)
        col = np.ascontiguousarray(A.indices, np.int32)
        val = np.ascontiguousarray(A.data, np.float64)
        return cls(lib().orc_csr_import(A.shape[0], A.shape[1], rp, col, val))

    def matvec(self, x):
        n, m, _ = self.shape
        y = np.empty(n)
        lib().orc_spmv(self.h, np.ascontiguousarray(x, np.float64), y)
        return y

    def rmatvec(self, x):
        n, m, _ = self.shape
        y = np.empty(m)
        lib().orc_spmv_t(self.h, np.ascontiguousarray(x, np.float64), y)
        return y


def dmat(E0=1.0, nu=0.3):
    D = np.empty(36)
    lib().orc_dmat(E0, nu, D)
    return D.reshape(6, 6)


def ke(E0=1.0, nu=0.3, h=(1.0, 1.0, 1.0)):
    KE = np.empty(576)
    lib().orc_ke(E0, nu, h[0], h[1], h[2], KE)
    return KE.reshape(24, 24)


def filter_csr(nx, ny, nz, R, h=(1.0, 1.0, 1.0), scale=True):
    return CSR(lib().orc_filter_csr(nx, ny, nz, h[0], h[1], h[2], R, 1 if scale else 0))


def rfilter_csr(nx, ny, nz, Rmax, h=(1.0, 1.0, 1.0)):
    """Unit-weight max-length-scale filter R and REdge = max rowsum - rowsum."""
    REdge = np.empty(nx * ny * nz)
    A = CSR(lib().orc_rfilter_csr(nx, ny, nz, h[0], h[1], h[2], Rmax, REdge.ctypes.data_as(C.c_void_p)))
    return A, REdge


def matint(P, R, REdge, x, penal=3.0, vdPenal=1.0, vdMin=-1.0, interp=0, param=(0.0, 0.0),
           eps=1e-10):
    """Full MatIntFnc: returns dict rho, y, rhoq, E, dEdz, Es, dEsdz."""
    ne = np.asarray(x).size
    out = {k: np.empty(ne) for k in ("rho", "y", "rhoq", "E", "dEdz", "Es", "dEsdz")}
    Rh = None if R is None else R.h
    RE = None if REdge is None else np.ascontiguousarray(REdge, np.float64).ctypes.data_as(C.c_void_p)
    lib().orc_matint(P.h, Rh, RE, np.ascontiguousarray(x, np.float64), penal, vdPenal, vdMin, interp,
                     np.ascontiguousarray(param, np.float64), eps, out["rho"], out["y"], out["rhoq"],
                     out["E"], out["dEdz"], out["Es"], out["dEsdz"])
    return out


def chain_filter(P, R, rho, y, rhoq, dfdV=None, dfdE=None):
    """Full Chain_Filter on copies of dfdV / dfdE."""
    V = None if dfdV is None else np.array(dfdV, np.float64, copy=True)
    Ee = None if dfdE is None else np.array(dfdE, np.float64, copy=True)
    lib().orc_chain_filter(P.h, None if R is None else R.h, np.ascontiguousarray(rho, np.float64),
                           np.ascontiguousarray(y, np.float64), np.ascontiguousarray(rhoq, np.float64),
                           _ptr(V), _ptr(Ee))
    return V, Ee


def vecpow(x, p):
    return lib().orc_vecpow1(x, p)


def simp(rho, penal=3.0, eps=1e-10):
    rho = np.ascontiguousarray(rho, np.float64)
    ne = rho.size
    E, dEdz, Es, dEsdz = (np.empty(ne) for _ in range(4))
    lib().orc_simp(ne, rho, penal, eps, _ptr(E), _ptr(dEdz), _ptr(Es), _ptr(dEsdz))
    return E, dEdz, Es, dEsdz


def assemble_K(nx, ny, nz, KE, E, fixed_dof=None):
    KE = np.ascontiguousarray(KE, np.float64).reshape(-1)
    E = np.ascontiguousarray(E, np.float64)
    fx = None if fixed_dof is None else np.ascontiguousarray(fixed_dof, np.uint8)
    return CSR(lib().orc_assemble_K(nx, ny, nz, KE, E, _ptr(fx)))


def build_P(Nfx, Nfy, Nfz):
    return CSR(lib().orc_build_P(Nfx, Nfy, Nfz))


def ptap(A: CSR, P: CSR):
    return CSR(lib().orc_ptap(A.h, P.h))


def sensitivity(nx, ny, nz, KE, U, dEdz):
    g = np.empty(nx * ny * nz)
    lib().orc_sensitivity(nx, ny, nz, np.ascontiguousarray(KE, np.float64).reshape(-1),
                          np.ascontiguousarray(U, np.float64),
                          np.ascontiguousarray(dEdz, np.float64), g)
    return g


def mass_diag(nx, ny, nz, V, density, h=(1.0, 1.0, 1.0), mlump=None, fixed_dof=None):
    """Lumped mass diagonal (Dynamic.cpp:181-236, 44-45)."""
    elem = 1.0
    for v in h:
        elem *= v
    m0 = 1.0 / 2.0 ** 3 / 3 * elem * density
    M = np.empty(3 * (nx + 1) * (ny + 1) * (nz + 1))
    ml = None if mlump is None else np.ascontiguousarray(mlump, np.float64)
    fx = None if fixed_dof is None else np.ascontiguousarray(fixed_dof, np.uint8)
    lib().orc_mass_diag(nx, ny, nz, np.ascontiguousarray(V, np.float64), m0,
                        None if ml is None else ml.ctypes.data_as(C.c_void_p),
                        None if fx is None else fx.ctypes.data_as(C.c_void_p), M)
    return M


def assemble_Ks(nx, ny, nz, U, Es, fixed_dof=None, E0=1.0, nu=0.3, h=(1.0, 1.0, 1.0)):
    """Stress stiffness Ks = sum_e (-Es_e) ks_e(u_e) with zeroed fixed rows/columns."""
    fx = None if fixed_dof is None else np.ascontiguousarray(fixed_dof, np.uint8)
    return CSR(lib().orc_assemble_Ks(nx, ny, nz, E0, nu, h[0], h[1], h[2],
                                     np.ascontiguousarray(U, np.float64),
                                     np.ascontiguousarray(Es, np.float64),
                                     None if fx is None else fx.ctypes.data_as(C.c_void_p)))


def ks_quadratic(nx, ny, nz, U, phi, E0=1.0, nu=0.3, h=(1.0, 1.0, 1.0)):
    phi = np.ascontiguousarray(phi, np.float64)
    ne = nx * ny * nz
    nvec = phi.size // (3 * (nx + 1) * (ny + 1) * (nz + 1))
    out = np.empty(nvec * ne)
    lib().orc_ks_quadratic(nx, ny, nz, E0, nu, h[0], h[1], h[2], np.ascontiguousarray(U, np.float64),
                           phi, nvec, out)
    return out.reshape(nvec, ne)


def stability_rhs(nx, ny, nz, Es, phi, fixed_dof=None, E0=1.0, nu=0.3, h=(1.0, 1.0, 1.0)):
    phi = np.ascontiguousarray(phi, np.float64)
    ndof = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    nvec = phi.size // ndof
    out = np.empty(nvec * ndof)
    fx = None if fixed_dof is None else np.ascontiguousarray(fixed_dof, np.uint8)
    lib().orc_stability_rhs(nx, ny, nz, E0, nu, h[0], h[1], h[2], np.ascontiguousarray(Es, np.float64),
                            phi, nvec, None if fx is None else fx.ctypes.data_as(C.c_void_p), out)
    return out.reshape(nvec, ndof)


def lanczos_lambda(A: CSR, dinv, its):
    """lambda_max(D^-1 A) estimate from `its` Jacobi-CG (Lanczos) steps."""
    return lib().orc_lanczos_lambda(A.h, np.ascontiguousarray(dinv, np.float64), its)


def tridiag_max_eig(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return lib().orc_tridiag_max_eig(a.size, a, b)


def smooth(A: CSR, dinv, lam, b, x0=None, steps=2, lo=0.1, hi=1.1, smoother=0):
    """`steps` level-smoother steps (Chebyshev 0 / Jacobi 1) from x0 (None: zero guess)."""
    n = A.shape[0]
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64, copy=True)
    d, r = np.zeros(n), np.zeros(n)
    lib().orc_smooth(A.h, np.ascontiguousarray(dinv, np.float64), lam, lo, hi, smoother, steps,
                     1 if x0 is None else 0, np.ascontiguousarray(b, np.float64), x, d, r)
    return x


def power_start(n):
    v = np.empty(n)
    lib().orc_power_start(n, v)
    return v


class MG:
    """Galerkin geometric multigrid + CG (the oracle's FESolve)."""

    def __init__(self, K: CSR, node_dims, levels, smoother=0, nu1=2, nu2=2, lo=0.1, hi=1.1,
                 eig_its=10):
        self.K = K  # keep alive
        self.h = lib().orc_mg_create(K.h, node_dims[0], node_dims[1], node_dims[2], levels,
                                     smoother, nu1, nu2, lo, hi, eig_its)
        if not self.h:
            raise RuntimeError("orc_mg_create failed")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_mg_free(self.h)
            self.h = None

    @property
    def levels(self):
        return lib().orc_mg_levels(self.h)

    def A(self, l):
        return CSR(lib().orc_mg_A(self.h, l), owned=False)

    def P(self, l):
        return CSR(lib().orc_mg_P(self.h, l), owned=False)

    def lam(self, l):
        return lib().orc_mg_lambda(self.h, l)

    def dims(self, l):
        d = (C.c_int * 3)()
        lib().orc_mg_dims(self.h, l, d)
        return tuple(d)

    def setup_seconds(self):
        return lib().orc_mg_setup_seconds(self.h)

    def vcycle(self, r):
        z = np.empty_like(r)
        lib().orc_mg_vcycle(self.h, np.ascontiguousarray(r, np.float64), z)
        return z

    def solve(self, b, x0=None, rtol=1e-8, maxit=1000):
        x = np.zeros_like(b) if x0 is None else np.array(x0, np.float64, copy=True)
        its = C.c_int()
        rn = C.c_double()
        reason = lib().orc_pcg(self.h, np.ascontiguousarray(b, np.float64), x, rtol, maxit,
                               C.byref(its), C.byref(rn))
        return x, its.value, reason, rn.value


def num_threads():
    return lib().orc_num_threads()
