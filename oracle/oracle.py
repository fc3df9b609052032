"""ORACLE — test infrastructure only.

ctypes wrapper around liboracle.so, the plain-C restatement of the reference
ddCOSMO path (ddcosmo_oracle.c).  Only tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg may import this module.

Arrays are returned in the reference's Fortran layout as numpy arrays shaped
(fast, slow) -> e.g. sigma has shape (nsph, nylm) in C order, which is the
column-major (nylm, nsph) buffer of the reference.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


# callbacks with the reference's Fortran argument passing (jacobi_diis.f90:33-55)
MATVEC_FN = C.CFUNCTYPE(None, _ip, _dp, _dp)
NORM_FN = C.CFUNCTYPE(C.c_double, _ip, _dp)


def build():
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def _lib():
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    lib.orc_init.argtypes = [C.POINTER(C.c_void_p), C.c_int, _dp, _dp, _dp, _dp,
                             C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]
    lib.orc_free.argtypes = [C.c_void_p]
    lib.orc_get_sizes.argtypes = [C.c_void_p, _ip]
    lib.orc_copy_out.argtypes = [C.c_void_p, C.c_char_p, _dp]
    lib.orc_copy_out_int.argtypes = [C.c_void_p, C.c_char_p, _ip]
    for f in ("orc_lx", "orc_lstarx", "orc_ldm1x"):
        getattr(lib, f).argtypes = [C.c_void_p, _dp, _dp]
    lib.orc_hnorm.argtypes = [C.c_void_p, _dp]
    lib.orc_hnorm.restype = C.c_double
    lib.orc_mkrhs.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_int, _dp, _dp, C.c_int, _dp]
    lib.orc_efld.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _dp]
    lib.orc_wghpot.argtypes = [C.c_void_p, _dp, _dp]
    lib.orc_intrhs.argtypes = [C.c_void_p, _dp, _dp]
    lib.orc_jacobi_diis.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, _dp, _dp, C.c_int,
                                    C.c_void_p, C.c_void_p]
    lib.orc_jacobi_diis_ops.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp,
                                        C.c_int, MATVEC_FN, MATVEC_FN, NORM_FN, _ip, _dp]
    lib.orc_cosmo.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_void_p]
    lib.orc_forces_dd.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    lib.orc_ddmkzeta.argtypes = [C.c_void_p, _dp, _dp]
    lib.orc_energy_forces.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_void_p, C.c_void_p]
    lib.orc_fsw.argtypes = [C.c_double, C.c_double, C.c_double]
    lib.orc_fsw.restype = C.c_double
    lib.orc_dfsw.argtypes = [C.c_double, C.c_double, C.c_double]
    lib.orc_dfsw.restype = C.c_double
    lib.orc_ylmbas.argtypes = [C.c_void_p, _dp, _dp]
    lib.orc_dbasis.argtypes = [C.c_void_p, _dp, _dp, _dp]
    lib.orc_lebedev.argtypes = [C.c_int, _dp, _dp]
    lib.orc_snap_ngrid.argtypes = [C.c_int]
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _lib()
    return _LIB


def _p(a):
    return a.ctypes.data_as(_dp)


class Hist(C.Structure):
    _fields_ = [("n_iter", C.c_int), ("ok", C.c_int), ("n_singular", C.c_int),
                ("norms", C.c_double * 256)]


def fsw(t, eta, s=-1.0):
    return lib().orc_fsw(float(t), float(s), float(eta))


def dfsw(t, eta, s=-1.0):
    return lib().orc_dfsw(float(t), float(s), float(eta))


def lebedev(ngrid):
    w = np.zeros(ngrid)
    g = np.zeros((ngrid, 3))
    if lib().orc_lebedev(ngrid, _p(w), _p(g)) != 0:
        raise ValueError("no Lebedev grid with %d points" % ngrid)
    return w, g


def snap_ngrid(ngrid):
    return lib().orc_snap_ngrid(int(ngrid))


class Oracle:
    """One ddCOSMO system, set up as ddinit does (ddcosmo.f90:159-483)."""

    def __init__(self, x, y, z, r, lmax, ngrid, iconv=8, igrad=1, eps=78.39, eta=0.1):
        self.L = lib()
        x, y, z, r = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, r))
        self.h = C.c_void_p()
        rc = self.L.orc_init(C.byref(self.h), x.size, _p(x), _p(y), _p(z), _p(r),
                             lmax, ngrid, iconv, igrad, eps, eta)
        if rc != 0:
            raise RuntimeError("orc_init failed (%d)" % rc)
        s = (C.c_int * 8)()
        self.L.orc_get_sizes(self.h, s)
        (self.nsph, self.ngrid, self.lmax, self.nylm, self.ncav, self.nnl,
         self.igrad, self.iconv) = list(s)
        self.eps, self.eta = eps, eta

    @classmethod
    def from_input(cls, d, **over):
        x, y, z, r = d.bohr()
        kw = dict(lmax=d.lmax, ngrid=d.ngrid, iconv=d.iconv, igrad=d.igrad,
                  eps=d.eps, eta=d.eta)
        kw.update(over)
        return cls(x, y, z, r, **kw)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_free(self.h)
            self.h = None

    # ---- state -----------------------------------------------------------
    def get(self, name):
        shapes = {
            "grid": (self.ngrid, 3), "w": (self.ngrid,), "basis": (self.ngrid, self.nylm),
            "facs": (self.nylm,), "facl": (self.nylm,), "fi": (self.nsph, self.ngrid),
            "ui": (self.nsph, self.ngrid), "zi": (self.nsph, self.ngrid, 3),
            "ccav": (self.ncav, 3), "csph": (self.nsph, 3), "rsph": (self.nsph,),
        }
        if name in ("inl", "nl"):
            a = np.zeros(self.nsph + 1 if name == "inl" else self.nnl, dtype=np.int32)
            self.L.orc_copy_out_int(self.h, name.encode(), a.ctypes.data_as(_ip))
            return a
        a = np.zeros(shapes[name])
        if self.L.orc_copy_out(self.h, name.encode(), _p(a)) != 0:
            raise KeyError(name)
        return a

    # ---- operators -------------------------------------------------------
    def _vec(self, x):
        return np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(self.nsph, self.nylm))

    def lx(self, x):
        x = self._vec(x)
        y = np.zeros_like(x)
        self.L.orc_lx(self.h, _p(x), _p(y))
        return y

    def lstarx(self, x):
        x = self._vec(x)
        y = np.zeros_like(x)
        self.L.orc_lstarx(self.h, _p(x), _p(y))
        return y

    def ldm1x(self, x):
        x = self._vec(x)
        y = np.zeros_like(x)
        self.L.orc_ldm1x(self.h, _p(x), _p(y))
        return y

    def hnorm(self, x):
        x = self._vec(x)
        return self.L.orc_hnorm(self.h, _p(x))

    def ylmbas(self, s):
        s = np.ascontiguousarray(s, dtype=np.float64)
        out = np.zeros(self.nylm)
        self.L.orc_ylmbas(self.h, _p(s), _p(out))
        return out

    def dbasis(self, s):
        s = np.ascontiguousarray(s, dtype=np.float64)
        b = np.zeros(self.nylm)
        d = np.zeros((self.nylm, 3))
        self.L.orc_dbasis(self.h, _p(s), _p(b), _p(d))
        return b, d

    def intrhs(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(self.nylm)
        self.L.orc_intrhs(self.h, _p(x), _p(out))
        return out

    def wghpot(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        g = np.zeros((self.nsph, self.ngrid))
        self.L.orc_wghpot(self.h, _p(phi), _p(g))
        return g

    # ---- drivers ---------------------------------------------------------
    def mkrhs(self, charge):
        charge = np.ascontiguousarray(charge, dtype=np.float64)
        c = self.get("csph")
        x, y, z = (np.ascontiguousarray(c[:, k]) for k in range(3))
        ccav = np.ascontiguousarray(self.get("ccav"))
        phi = np.zeros(self.ncav)
        psi = np.zeros((self.nsph, self.nylm))
        self.L.orc_mkrhs(self.nsph, _p(charge), _p(x), _p(y), _p(z), self.ncav,
                         _p(ccav), _p(phi), self.nylm, _p(psi))
        return phi, psi

    def cosmo(self, star, phi=None, psi=None, glm=None):
        sigma = np.zeros((self.nsph, self.nylm))
        es = C.c_double(0.0)
        hist = Hist()
        z = np.zeros(1)
        phi_a = np.ascontiguousarray(phi if phi is not None else z, dtype=np.float64)
        psi_a = np.ascontiguousarray(psi, dtype=np.float64)
        cart = 1 if glm is None else 0
        glm_a = np.ascontiguousarray(glm if glm is not None else z, dtype=np.float64)
        rc = self.L.orc_cosmo(self.h, int(star), cart, _p(phi_a), _p(glm_a), _p(psi_a),
                              _p(sigma), C.cast(C.byref(es), _dp), C.byref(hist))
        return sigma, es.value, hist, rc

    def forces_dd(self, phi, sigma, s):
        fx = np.zeros((self.nsph, 3))
        self.L.orc_forces_dd(self.h, _p(np.ascontiguousarray(phi)), _p(np.ascontiguousarray(sigma)),
                             _p(np.ascontiguousarray(s)), _p(fx))
        return fx

    def ddmkzeta(self, s):
        zeta = np.zeros(self.ncav)
        self.L.orc_ddmkzeta(self.h, _p(np.ascontiguousarray(s)), _p(zeta))
        return zeta

    def energy_forces(self, charge):
        charge = np.ascontiguousarray(charge, dtype=np.float64)
        fx = np.zeros((self.nsph, 3))
        es = C.c_double(0.0)
        hd, ha = Hist(), Hist()
        rc = self.L.orc_energy_forces(self.h, _p(charge), C.cast(C.byref(es), _dp), _p(fx),
                                      C.byref(hd), C.byref(ha))
        if rc != 0:
            raise RuntimeError("oracle: solver did not converge")
        return es.value, fx, hd, ha


def efld(src, csrc, ctrg):
    src = np.ascontiguousarray(src, dtype=np.float64)
    csrc = np.ascontiguousarray(csrc, dtype=np.float64)
    ctrg = np.ascontiguousarray(ctrg, dtype=np.float64)
    ef = np.zeros((ctrg.shape[0], 3))
    lib().orc_efld(src.size, _p(src), _p(csrc), ctrg.shape[0], _p(ctrg), _p(ef))
    return ef


def dense_callbacks(a, d):
    """matvec y = A x, dm1vec y = x / d and u_norm sqrt(sum_i x_i^2 / i / n)
    for a dense system, summed in index order as a plain loop would (the
    callbacks of oracle/jd_dump.f90), as ctypes function pointers."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    inv_i = 1.0 / np.arange(1, d.size + 1, dtype=np.float64)

    def mv(n_p, x_p, y_p):
        n = n_p[0]
        x = np.ctypeslib.as_array(x_p, shape=(n,))
        np.ctypeslib.as_array(y_p, shape=(n,))[:] = np.cumsum(a * x[None, :], axis=1)[:, -1]

    def dm1(n_p, x_p, y_p):
        n = n_p[0]
        np.ctypeslib.as_array(y_p, shape=(n,))[:] = np.ctypeslib.as_array(x_p, shape=(n,)) / d

    def nrm(n_p, x_p):
        n = n_p[0]
        x = np.ctypeslib.as_array(x_p, shape=(n,))
        return float(np.sqrt(np.cumsum((x * x) * inv_i)[-1] / n))

    return MATVEC_FN(mv), MATVEC_FN(dm1), NORM_FN(nrm)


def jacobi_diis_ops(n, diis_max, norm, tol, rhs, x, n_iter, matvec, dm1vec, u_norm):
    """The reference's jacobi_diis with callbacks; returns (x, n_iter, ok,
    norms (iterations, 2): rms or u_norm, max or -1)."""
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    x = np.array(x, dtype=np.float64)
    ok = C.c_int(0)
    norms = np.zeros((256, 2))
    it = lib().orc_jacobi_diis_ops(n, diis_max, norm, tol, _p(rhs), _p(x), n_iter, matvec,
                                   dm1vec, u_norm, C.byref(ok), _p(norms))
    return x, it, ok.value, norms[:min(it if ok.value else n_iter, 256)]
