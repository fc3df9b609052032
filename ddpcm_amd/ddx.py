"""Python host mirror of the reference ddCOSMO module over libddx's C-ABI.

Procedure names and argument meaning follow the reference module
(ddcosmo.f90, cosmo.f90, forces_dd.f90, matvec.f90, jacobi_diis.f90,
mkrhs.f90, efld.f90); arrays use the reference's Fortran layout, exposed as
C-ordered numpy arrays with the slow index first, e.g. sigma(nylm, nsph)
is an array of shape (nsph, nylm).  All compute runs in libddx.so on the GPU;
this module has no CPU fallback and raises if the library cannot be loaded.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libddx.so")

DDX_OK, DDX_ERR_ARG, DDX_ERR_ALLOC, DDX_ERR_NOCONV = 0, 1, 2, 3
DDX_ERR_CAPACITY, DDX_ERR_HIP, DDX_ERR_STATE = 4, 5, 6
DDX_PREC_FP64, DDX_PREC_MIXED = 0, 1

EXPORTS = (
    "ddx_last_error", "ddx_init", "ddx_free", "ddx_get_sizes", "ddx_get_array",
    "ddx_get_neighbours", "ddx_mkrhs", "ddx_cosmo", "ddx_forces_dd", "ddx_mkzeta",
    "ddx_efld", "ddx_lx", "ddx_lstarx", "ddx_ldm1x", "ddx_hnorm", "ddx_jacobi_diis",
    "ddx_get_history", "ddx_energy_forces", "ddx_upload", "ddx_run", "ddx_result",
    "ddx_set_precision", "ddx_get_stats", "ddx_kernel_timer", "ddx_kernel_timer_read",
    "ddx_batch_init", "ddx_batch_energy_forces", "ddx_batch_result", "ddx_batch_iterations",
    "ddx_batch_sizes", "ddx_gemm_bf16", "ddx_set_guess", "ddx_guess_train", "ddx_guess_reset",
    "ddx_mlp_grad", "ddx_set_graphs", "ddx_autotune", "ddx_efld_noalloc",
    "ddx_jacobi_diis_ops", "ddx_op_lx", "ddx_op_lstarx", "ddx_op_ldm1x", "ddx_op_hnorm",
    "ddx_mkrhs_at", "ddx_set_coulomb",
)
DDX_GUESS_REFERENCE, DDX_GUESS_PREVIOUS, DDX_GUESS_LEARNED, DDX_GUESS_LEARNED_PREVIOUS = 0, 1, 2, 3
KERNELS = ("sweep_l", "sweep_lstar", "project", "expand", "jacobi", "diis", "extrap", "field",
           "potential", "forces", "sweep_pair")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class DDXError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("libddx error %d: %s" % (code, msg))
        self.code = code


class NotConverged(DDXError):
    pass


class Stats(C.Structure):
    _fields_ = [("niter_direct", C.c_int), ("niter_adjoint", C.c_int),
                ("ntriples", C.c_int), ("nnl", C.c_int), ("ncav", C.c_int),
                ("max_degree", C.c_int), ("ms_setup", C.c_double), ("ms_rhs", C.c_double),
                ("ms_direct", C.c_double), ("ms_adjoint", C.c_double),
                ("ms_forces", C.c_double)]


_LIB = None


def load_library(path: str = LIB_PATH):
    """Loads libddx.so and declares its signatures (no GPU call is made)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise ImportError("libddx.so is not built (%s); run __graft_entry__.build()" % path)
    L = C.CDLL(path)
    L.ddx_last_error.restype = C.c_char_p
    L.ddx_init.argtypes = [C.POINTER(C.c_void_p), C.c_int, _dp, _dp, _dp, _dp, C.c_int,
                           C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]
    L.ddx_free.argtypes = [C.c_void_p]
    L.ddx_get_sizes.argtypes = [C.c_void_p, _ip, _ip, _ip, _ip, _ip]
    L.ddx_get_array.argtypes = [C.c_void_p, C.c_char_p, _dp]
    L.ddx_get_neighbours.argtypes = [C.c_void_p, _ip, _ip]
    L.ddx_mkrhs.argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.ddx_cosmo.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]
    L.ddx_forces_dd.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    L.ddx_mkzeta.argtypes = [C.c_void_p, _dp, _dp]
    L.ddx_efld.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_int, _dp, _dp]
    # device pointers and a hipStream_t, as integers
    L.ddx_efld_noalloc.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                   C.c_void_p, C.c_void_p]
    for f in ("ddx_lx", "ddx_lstarx", "ddx_ldm1x"):
        getattr(L, f).argtypes = [C.c_void_p, _dp, _dp]
    L.ddx_hnorm.argtypes = [C.c_void_p, _dp, _dp]
    L.ddx_jacobi_diis.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, _dp, _dp, _ip, _ip]
    L.ddx_jacobi_diis_ops.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_double, _dp, _dp, _ip, _ip, C.c_void_p, C.c_void_p,
                                      C.c_void_p]
    L.ddx_op_hnorm.restype = C.c_double
    L.ddx_mkrhs_at.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, C.c_int, _dp, _dp,
                               C.c_int, _dp]
    L.ddx_get_history.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _ip]
    L.ddx_energy_forces.argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.ddx_upload.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
    L.ddx_run.argtypes = [C.c_void_p]
    L.ddx_result.argtypes = [C.c_void_p, _dp, _dp]
    L.ddx_set_precision.argtypes = [C.c_void_p, C.c_int]
    L.ddx_get_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
    L.ddx_kernel_timer.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.ddx_kernel_timer_read.argtypes = [C.c_void_p, C.c_int, _ip, _dp]
    L.ddx_batch_init.argtypes = [C.POINTER(C.c_void_p), C.c_int, _ip, _dp, _dp, _dp, _dp,
                                 C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]
    L.ddx_batch_energy_forces.argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.ddx_batch_result.argtypes = [C.c_void_p, _dp, _dp]
    L.ddx_batch_iterations.argtypes = [C.c_void_p, _ip, _ip]
    L.ddx_batch_sizes.argtypes = [C.c_void_p, _ip, _ip, _ip]
    _u16 = C.POINTER(C.c_uint16)
    L.ddx_gemm_bf16.argtypes = [C.c_int, C.c_int, C.c_int, _u16, _u16, C.POINTER(C.c_float)]
    L.ddx_set_guess.argtypes = [C.c_void_p, C.c_int]
    L.ddx_set_graphs.argtypes = [C.c_void_p, C.c_int]
    L.ddx_set_coulomb.argtypes = [C.c_void_p, C.c_int]
    L.ddx_autotune.argtypes = [C.c_void_p, _ip, _ip]
    L.ddx_guess_train.argtypes = [C.c_void_p, C.c_int, C.c_double, _dp]
    L.ddx_guess_reset.argtypes = [C.c_void_p, C.c_uint]
    _fp = C.POINTER(C.c_float)
    L.ddx_mlp_grad.argtypes = [C.c_int, C.c_int, C.c_int, _u16, _fp, C.POINTER(_fp), _fp,
                               C.POINTER(_fp), _dp]
    _LIB = L
    return L


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


class DDCosmo:
    """One ddCOSMO system on the GPU (ddinit, ddcosmo.f90:159).

    x, y, z, rvdw in bohr.  ngrid is snapped to the nearest Lebedev size.
    """

    def __init__(self, x, y, z, rvdw, lmax, ngrid, iconv=8, igrad=1, eps=78.39,
                 eta=0.1, iprint=0):
        self.L = load_library()
        x, y, z, r = (_f64(a).ravel() for a in (x, y, z, rvdw))
        self.h = C.c_void_p()
        self._check(self.L.ddx_init(C.byref(self.h), x.size, _p(x), _p(y), _p(z), _p(r),
                                    iprint, lmax, ngrid, iconv, igrad, eps, eta))
        s = [C.c_int() for _ in range(5)]
        self._check(self.L.ddx_get_sizes(self.h, *[C.byref(v) for v in s]))
        self.nsph, self.ncav, self.nylm, self.ngrid, self.lmax = (v.value for v in s)
        self.iconv, self.igrad, self.eps, self.eta, self.iprint = iconv, igrad, eps, eta, iprint

    @classmethod
    def from_input(cls, d, **over):
        x, y, z, r = d.bohr()
        kw = dict(lmax=d.lmax, ngrid=d.ngrid, iconv=d.iconv, igrad=d.igrad, eps=d.eps,
                  eta=d.eta, iprint=d.iprint)
        kw.update(over)
        return cls(x, y, z, r, **kw)

    def _check(self, rc, allow_noconv=False):
        if rc == DDX_OK:
            return
        msg = self.L.ddx_last_error().decode(errors="replace")
        if rc == DDX_ERR_NOCONV:
            if allow_noconv:
                return
            raise NotConverged(rc, msg)
        raise DDXError(rc, msg)

    def memfree(self):
        """memfree (ddcosmo.f90:491)."""
        if getattr(self, "h", None):
            self.L.ddx_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.memfree()
        except Exception:
            pass

    # ---- module data -------------------------------------------------------
    def get(self, name):
        shapes = {"grid": (self.ngrid, 3), "w": (self.ngrid,), "basis": (self.ngrid, self.nylm),
                  "facl": (self.nylm,), "facs": (self.nylm,), "fi": (self.nsph, self.ngrid),
                  "ui": (self.nsph, self.ngrid), "zi": (self.nsph, self.ngrid, 3),
                  "ccav": (self.ncav, 3), "csph": (self.nsph, 3), "rsph": (self.nsph,),
                  "rhs": (self.nsph, self.nylm), "xi": (self.nsph, self.ngrid),
                  "sigma": (self.nsph, self.nylm), "s": (self.nsph, self.nylm)}
        a = np.zeros(shapes[name])
        self._check(self.L.ddx_get_array(self.h, name.encode(), _p(a)))
        return a

    @property
    def ccav(self):
        return self.get("ccav")

    def neighbours(self):
        """(inl, nl), 1-based as the reference stores them."""
        inl = np.zeros(self.nsph + 1, dtype=np.int32)
        st = self.stats()
        nl = np.zeros(max(st.nnl, 1), dtype=np.int32)
        self._check(self.L.ddx_get_neighbours(self.h, inl.ctypes.data_as(_ip),
                                              nl.ctypes.data_as(_ip)))
        return inl, nl[:st.nnl]

    def stats(self):
        s = Stats()
        self._check(self.L.ddx_get_stats(self.h, C.byref(s)))
        return s

    def history(self, star):
        cnt = C.c_int()
        buf = np.zeros(256)
        self._check(self.L.ddx_get_history(self.h, int(star), 256, _p(buf), C.byref(cnt)))
        return buf[:cnt.value].copy()

    # ---- reference procedures ----------------------------------------------
    def mkrhs(self, charge):
        """mkrhs (mkrhs.f90:1): phi(ncav), psi(nylm, nsph)."""
        q = _f64(charge).ravel()
        phi = np.zeros(self.ncav)
        psi = np.zeros((self.nsph, self.nylm))
        self._check(self.L.ddx_mkrhs(self.h, _p(q), _p(phi), _p(psi)))
        return phi, psi

    def cosmo(self, star, cart=True, phi=None, glm=None, psi=None):
        """cosmo (cosmo.f90:1): returns (sigma, esolv)."""
        psi = _f64(psi, (self.nsph, self.nylm))
        sigma = np.zeros((self.nsph, self.nylm))
        es = C.c_double(0.0)
        z = np.zeros(1)
        phi_a = _f64(phi) if phi is not None else z
        glm_a = _f64(glm, (self.nsph, self.nylm)) if glm is not None else z
        rc = self.L.ddx_cosmo(self.h, int(bool(star)), int(bool(cart)), _p(phi_a), _p(glm_a),
                              _p(psi), _p(sigma), C.cast(C.byref(es), _dp))
        self._check(rc)
        return sigma, es.value

    def forces_dd(self, phi, sigma, s):
        """forces_dd (forces_dd.f90:1): fx(3, nsph)."""
        fx = np.zeros((self.nsph, 3))
        self._check(self.L.ddx_forces_dd(self.h, _p(_f64(phi)), _p(_f64(sigma)), _p(_f64(s)),
                                         _p(fx)))
        return fx

    def ddmkzeta(self, s):
        zeta = np.zeros(self.ncav)
        self._check(self.L.ddx_mkzeta(self.h, _p(_f64(s)), _p(zeta)))
        return zeta

    def efld(self, src, csrc, ctrg):
        """efld (efld.f90:1): field of point sources at targets, (ntrg, 3)."""
        src, csrc, ctrg = _f64(src).ravel(), _f64(csrc), _f64(ctrg)
        ef = np.zeros((ctrg.shape[0], 3))
        self._check(self.L.ddx_efld(self.h, src.size, _p(src), _p(csrc), ctrg.shape[0],
                                    _p(ctrg), _p(ef)))
        return ef

    def _op(self, f, x):
        x = _f64(x, (self.nsph, self.nylm))
        y = np.zeros_like(x)
        self._check(f(self.h, _p(x), _p(y)))
        return y

    def lx(self, x):
        return self._op(self.L.ddx_lx, x)

    def lstarx(self, x):
        return self._op(self.L.ddx_lstarx, x)

    def ldm1x(self, x):
        return self._op(self.L.ddx_ldm1x, x)

    def hnorm(self, x):
        x = _f64(x, (self.nsph, self.nylm))
        v = C.c_double()
        self._check(self.L.ddx_hnorm(self.h, _p(x), C.byref(v)))
        return v.value

    def jacobi_diis(self, star, rhs, x, diis_max=25, tol=None, n_iter=200):
        """jacobi_diis (jacobi_diis.f90:43) with norm 4; returns (x, n_iter, ok)."""
        tol = 10.0 ** (-self.iconv) if tol is None else tol
        rhs = _f64(rhs, (self.nsph, self.nylm))
        x = _f64(x, (self.nsph, self.nylm)).copy()
        nit, ok = C.c_int(n_iter), C.c_int(0)
        self._check(self.L.ddx_jacobi_diis(self.h, int(star), diis_max, tol, _p(rhs), _p(x),
                                           C.byref(nit), C.byref(ok)))
        return x, nit.value, bool(ok.value)

    def energy_forces(self, charge):
        """main.f90:164-265 for point charges: (esolv in Eh, fx (nsph, 3) in Eh/bohr)."""
        q = _f64(charge).ravel()
        es = C.c_double()
        fx = np.zeros((self.nsph, 3))
        self._check(self.L.ddx_energy_forces(self.h, _p(q), C.cast(C.byref(es), _dp), _p(fx)))
        return es.value, fx

    # ---- device-resident path ------------------------------------------------
    def upload(self, x, y, z, rvdw, charge):
        x, y, z, r, q = (_f64(a).ravel() for a in (x, y, z, rvdw, charge))
        self._check(self.L.ddx_upload(self.h, _p(x), _p(y), _p(z), _p(r), _p(q)))

    def run(self):
        self._check(self.L.ddx_run(self.h))

    def result(self):
        es = C.c_double()
        fx = np.zeros((self.nsph, 3))
        self._check(self.L.ddx_result(self.h, C.cast(C.byref(es), _dp), _p(fx)))
        return es.value, fx

    def kernel_timer(self, name, enable=True):
        """Brackets every launch of one kernel class with HIP events."""
        self._check(self.L.ddx_kernel_timer(self.h, KERNELS.index(name), int(enable)))

    def kernel_timer_read(self, name):
        """(launches, total ms) since the last read; resets the count."""
        n, ms = C.c_int(), C.c_double()
        self._check(self.L.ddx_kernel_timer_read(self.h, KERNELS.index(name), C.byref(n),
                                                 C.byref(ms)))
        return n.value, ms.value

    def set_precision(self, prec):
        self._check(self.L.ddx_set_precision(self.h, prec))

    # ---- initial guess (no counterpart in the reference) ----------------------
    def set_guess(self, mode):
        """DDX_GUESS_REFERENCE (D^-1 rhs), _PREVIOUS or _LEARNED."""
        self._check(self.L.ddx_set_guess(self.h, int(mode)))

    def guess_train(self, steps=4, lr=1e-3):
        """Adam steps of the guess networks towards the last solve's solutions;
        returns their losses (direct, adjoint)."""
        loss = np.zeros(2)
        self._check(self.L.ddx_guess_train(self.h, int(steps), float(lr), _p(loss)))
        return float(loss[0]), float(loss[1])

    def autotune(self):
        """Pick sweep workers per CU and the projection split by timing the
        uploaded system's pipeline; returns (workers_per_cu, proj_split)."""
        w, sp = C.c_int(), C.c_int()
        self._check(self.L.ddx_autotune(self.h, C.byref(w), C.byref(sp)))
        return w.value, sp.value

    def set_coulomb(self, mode):
        """Coulomb sums: 0 automatic, 1 direct, 2 far-field expansions."""
        self._check(self.L.ddx_set_coulomb(self.h, int(mode)))

    def set_graphs(self, enable=True):
        """Replay the solver's iteration groups as captured HIP graphs."""
        self._check(self.L.ddx_set_graphs(self.h, int(bool(enable))))

    def guess_reset(self, seed=1):
        self._check(self.L.ddx_guess_reset(self.h, int(seed)))


class DDCosmoBatch(DDCosmo):
    """A batch of independent ddCOSMO systems on one handle (no counterpart in
    the reference, which solves one system per process).  Each instance
    gets what main.f90:164-265 gives it alone; all run in the same kernels.

    xs, ys, zs, rs: per-instance coordinate and radius arrays in bohr.
    """

    def __init__(self, xs, ys, zs, rs, lmax, ngrid, iconv=8, igrad=1, eps=78.39, eta=0.1):
        self.L = load_library()
        sizes = np.array([len(np.ravel(x)) for x in xs], dtype=np.int32)
        x, y, z, r = (np.ascontiguousarray(np.concatenate([np.ravel(a) for a in v]),
                                           dtype=np.float64) for v in (xs, ys, zs, rs))
        self.h = C.c_void_p()
        self._check(self.L.ddx_batch_init(C.byref(self.h), sizes.size,
                                          sizes.ctypes.data_as(_ip), _p(x), _p(y), _p(z), _p(r),
                                          lmax, ngrid, iconv, igrad, eps, eta))
        s = [C.c_int() for _ in range(5)]
        self._check(self.L.ddx_get_sizes(self.h, *[C.byref(v) for v in s]))
        self.nsph, self.ncav, self.nylm, self.ngrid, self.lmax = (v.value for v in s)
        self.iconv, self.igrad, self.eps, self.eta, self.iprint = iconv, igrad, eps, eta, 0
        self.count = int(sizes.size)
        so = np.zeros(self.count + 1, dtype=np.int32)
        co = np.zeros(self.count + 1, dtype=np.int32)
        self._check(self.L.ddx_batch_sizes(self.h, None, so.ctypes.data_as(_ip),
                                           co.ctypes.data_as(_ip)))
        self.sph_off, self.cav_off = so, co

    @classmethod
    def from_inputs(cls, decks, **over):
        b = [d.bohr() for d in decks]
        d0 = decks[0]
        kw = dict(lmax=d0.lmax, ngrid=d0.ngrid, iconv=d0.iconv, igrad=d0.igrad, eps=d0.eps,
                  eta=d0.eta)
        kw.update(over)
        return cls([v[0] for v in b], [v[1] for v in b], [v[2] for v in b], [v[3] for v in b],
                   **kw)

    def energy_forces(self, charges):
        """(esolv[count] in Eh, fx (nsph, 3) in Eh/bohr, instance after instance)."""
        q = np.ascontiguousarray(np.concatenate([np.ravel(c) for c in charges]),
                                 dtype=np.float64)
        es = np.zeros(self.count)
        fx = np.zeros((self.nsph, 3))
        self._check(self.L.ddx_batch_energy_forces(self.h, _p(q), _p(es), _p(fx)))
        return es, fx

    def result(self):
        es = np.zeros(self.count)
        fx = np.zeros((self.nsph, 3))
        self._check(self.L.ddx_batch_result(self.h, _p(es), _p(fx)))
        return es, fx

    def iterations(self):
        """(direct, adjoint) iteration counts per instance."""
        d = np.zeros(self.count, dtype=np.int32)
        a = np.zeros(self.count, dtype=np.int32)
        self._check(self.L.ddx_batch_iterations(self.h, d.ctypes.data_as(_ip),
                                                a.ctypes.data_as(_ip)))
        return d, a

    def split(self, fx):
        """Per-instance views of a (nsph, 3) array."""
        return [fx[self.sph_off[q]:self.sph_off[q + 1]] for q in range(self.count)]


def bf16_bits(x):
    """float array -> BF16 bit patterns (round to nearest even) as uint16."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_value(b):
    """BF16 bit patterns -> float32 values."""
    return (np.ascontiguousarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def gemm_bf16(a_bits, b_bits):
    """C = A B on the device's BF16 product kernel (FP32 result)."""
    L = load_library()
    a = np.ascontiguousarray(a_bits, dtype=np.uint16)
    b = np.ascontiguousarray(b_bits, dtype=np.uint16)
    M, K = a.shape
    K2, N = b.shape
    assert K == K2
    c = np.zeros((M, N), dtype=np.float32)
    rc = L.ddx_gemm_bf16(M, N, K, a.ctypes.data_as(C.POINTER(C.c_uint16)),
                         b.ctypes.data_as(C.POINTER(C.c_uint16)),
                         c.ctypes.data_as(C.POINTER(C.c_float)))
    if rc:
        raise DDXError(rc, L.ddx_last_error().decode())
    return c


def mlp_grad(F_bits, T, weights):
    """One forward/backward pass of the guess network on the device:
    F_bits (n, in) BF16 patterns, T (n, out), weights [W1, b1, W2, b2, W3, b3]
    (FP32, W row-major (in, out)); returns (Y, grads like weights, loss)."""
    L = load_library()
    F = np.ascontiguousarray(F_bits, dtype=np.uint16)
    T = np.ascontiguousarray(T, dtype=np.float32)
    n, nin = F.shape
    nout = T.shape[1]
    ws = [np.ascontiguousarray(w, dtype=np.float32) for w in weights]
    gs = [np.zeros_like(w) for w in ws]
    fp = C.POINTER(C.c_float)
    warr = (fp * 6)(*[w.ctypes.data_as(fp) for w in ws])
    garr = (fp * 6)(*[g.ctypes.data_as(fp) for g in gs])
    Y = np.zeros((n, nout), dtype=np.float32)
    loss = C.c_double()
    rc = L.ddx_mlp_grad(n, nin, nout, F.ctypes.data_as(C.POINTER(C.c_uint16)),
                        T.ctypes.data_as(fp), warr, Y.ctypes.data_as(fp), garr, C.byref(loss))
    if rc:
        raise DDXError(rc, L.ddx_last_error().decode())
    return Y, gs, loss.value
