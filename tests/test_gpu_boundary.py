"""The drop-in boundary beyond the module's fixed calling sequence.

* jacobi_diis with the reference's plugin interface (jacobi_diis.f90:43-55):
  caller-supplied matvec / dm1vec / u_norm and every norm choice.  The dense
  systems of tests/golden/jacobi_dense.npz were solved by the reference's
  own jacobi_diis.f90 (oracle/jd_dump.f90); the C-ABI entry
  ddx_jacobi_diis_ops and the Fortran module's jacobi_diis must give the
  same iteration counts and ok flags, x within 1e-12 of the largest entry,
  and the same printed residual lines: the same text, and each norm within
  one unit of its last printed digit, or 5% once it is below 1e-6 (the Gram
  dots are summed in another order, and the late increments of a DIIS
  extrapolation carry that rounding).  The oracle's restatement
  (orc_jacobi_diis_ops) stands beside.
* With the handle's own operators the solve runs on the device and equals
  cosmo's adjoint solve bit for bit; a caller's wrapper of the same operator
  takes the host-callback path and agrees within the convergence threshold.
* mkrhs honours the coordinates it is given (mkrhs.f90:86-123): moved
  charges give the oracle's potential at the cavity points (1e-13).
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, deck, golden, hsnorm_rms

pytestmark = pytest.mark.gpu

CASES = ["slow", "norm1", "norm2", "norm3", "plain", "fail"]
JD_MODULE = os.path.join(ROOT, "tests", "fortran", "jd_module")


def _case(name):
    z = golden("jacobi_dense")
    g = {k[len(name) + 1:]: z[k] for k in z if k.startswith(name + "_")}
    n, dmax, nrm, nit = (int(v) for v in g["params"])
    return g, n, dmax, nrm, nit


def _same_lines(got, ref):
    got = [l for l in got if "iter=" in l]
    assert len(got) == len(ref)
    for a, b in zip(got, ref):
        ta, tb = a.split(), str(b).split()
        assert len(ta) == len(tb), (a, b)
        for u, v in zip(ta, tb):
            if "D" in v and v[0] in "-0123456789":
                fu, fv = float(u.replace("D", "E")), float(v.replace("D", "E"))
                unit = 1e-4 * 10.0 ** int(v.split("D")[1]) if fv != 0.0 else 0.0
                tol = unit * 1.01 if abs(fv) >= 1e-6 else max(unit * 1.01, 5e-2 * abs(fv))
                assert abs(fu - fv) <= tol, (a, b)
            else:
                assert u == v, (a, b)


def _lib():
    from ddpcm_amd import ddx
    return ddx, ddx.load_library()


@pytest.mark.parametrize("name", CASES)
def test_jacobi_diis_ops_dense_against_reference(name, oracle_mod, capfd):
    ddx, L = _lib()
    g, n, dmax, nrm, nit = _case(name)
    print("jacobi_diis_ops dense case %s n=%d" % (name, n), flush=True)
    capfd.readouterr()
    mv, dm1, nrmf = oracle_mod.dense_callbacks(g["a"], g["d"])
    x = np.array(g["x0"])
    it, ok = C.c_int(nit), C.c_int(-1)
    rc = L.ddx_jacobi_diis_ops(None, n, 1, dmax, nrm, float(g["tol"]), ddx._p(g["rhs"]),
                               ddx._p(x), C.byref(it), C.byref(ok),
                               C.cast(mv, C.c_void_p), C.cast(dm1, C.c_void_p),
                               C.cast(nrmf, C.c_void_p))
    out = capfd.readouterr().out.splitlines()
    assert rc == (ddx.DDX_OK if int(g["ok"]) else ddx.DDX_ERR_NOCONV), L.ddx_last_error()
    assert it.value == int(g["n_iter"]) and ok.value == int(g["ok"])
    assert np.max(np.abs(x - g["x"])) <= 1e-12 * np.max(np.abs(g["x"]))
    _same_lines(out, g["lines"])
    xo, ito, oko, _ = oracle_mod.jacobi_diis_ops(n, dmax, nrm, float(g["tol"]), g["rhs"],
                                                 g["x0"], nit, mv, dm1, nrmf)
    assert ito == it.value and oko == ok.value
    assert np.max(np.abs(x - xo)) <= 1e-12 * np.max(np.abs(xo))


def test_jacobi_diis_ops_rejects_norm4_without_u_norm(oracle_mod):
    ddx, L = _lib()
    g, n, dmax, nrm, nit = _case("norm1")
    mv, dm1, _ = oracle_mod.dense_callbacks(g["a"], g["d"])
    x = np.array(g["x0"])
    it, ok = C.c_int(nit), C.c_int(0)
    rc = L.ddx_jacobi_diis_ops(None, n, 0, 25, 4, 1e-8, ddx._p(g["rhs"]), ddx._p(x),
                               C.byref(it), C.byref(ok), C.cast(mv, C.c_void_p),
                               C.cast(dm1, C.c_void_p), None)
    assert rc == ddx.DDX_ERR_ARG and b"norm" in L.ddx_last_error()


def test_jacobi_diis_ops_with_the_handles_operators():
    ddx, L = _lib()
    d = deck("caffeine")
    g = golden("caffeine")
    h = ddx.DDCosmo.from_input(d)
    phi, psi = h.mkrhs(d.charge)
    s_ref, _ = h.cosmo(True, psi=psi)
    nn = psi.size
    tol = 10.0 ** (-d.iconv)
    # the library's own callbacks: the device solver
    x = h.ldm1x(psi)
    it, ok = C.c_int(200), C.c_int(0)
    rc = L.ddx_jacobi_diis_ops(h.h, nn, 0, 25, 4, tol, ddx._p(np.ascontiguousarray(psi)),
                               ddx._p(x), C.byref(it), C.byref(ok),
                               C.cast(L.ddx_op_lstarx, C.c_void_p),
                               C.cast(L.ddx_op_ldm1x, C.c_void_p),
                               C.cast(L.ddx_op_hnorm, C.c_void_p))
    assert rc == 0 and ok.value == 1, L.ddx_last_error()
    assert it.value == int(g["niter_adjoint"])
    assert np.array_equal(x, s_ref)
    hist = h.history(2)
    assert len(hist) == it.value and hist[-1] < tol
    # a caller's callbacks around the same operators: the host-callback path
    keep = []

    def wrap_mv(n_p, x_p, y_p):
        xv = np.ctypeslib.as_array(x_p, shape=psi.shape)
        np.ctypeslib.as_array(y_p, shape=psi.shape)[:] = h.lstarx(xv.copy())

    def wrap_dm1(n_p, x_p, y_p):
        xv = np.ctypeslib.as_array(x_p, shape=psi.shape)
        np.ctypeslib.as_array(y_p, shape=psi.shape)[:] = h.ldm1x(xv.copy())

    def wrap_norm(n_p, x_p):
        return h.hnorm(np.ctypeslib.as_array(x_p, shape=psi.shape).copy())

    import oracle
    cbs = (oracle.MATVEC_FN(wrap_mv), oracle.MATVEC_FN(wrap_dm1), oracle.NORM_FN(wrap_norm))
    keep.append(cbs)
    x2 = h.ldm1x(psi)
    it2, ok2 = C.c_int(200), C.c_int(0)
    rc = L.ddx_jacobi_diis_ops(h.h, nn, 0, 25, 4, tol, ddx._p(np.ascontiguousarray(psi)),
                               ddx._p(x2), C.byref(it2), C.byref(ok2),
                               *[C.cast(f, C.c_void_p) for f in cbs])
    assert rc == 0 and ok2.value == 1, L.ddx_last_error()
    assert it2.value == it.value
    assert hsnorm_rms(x2 - g["s"], d.lmax) < tol


@pytest.mark.parametrize("name", ["norm3", "slow", "fail"])
def test_fortran_module_jacobi_diis_dense(name, tmp_path):
    g, n, dmax, nrm, nit = _case(name)
    fin, fout = str(tmp_path / "jd.in"), str(tmp_path / "jd.out")
    with open(fin, "wb") as f:
        np.array([n, dmax, nrm, nit, 1], dtype="<i4").tofile(f)
        np.array([float(g["tol"])], dtype="<f8").tofile(f)
        for v in (g["a"].T, g["d"], g["rhs"], g["x0"]):
            np.ascontiguousarray(v, dtype="<f8").tofile(f)
    print("jd_module dense %s" % name, flush=True)
    r = subprocess.run([JD_MODULE, "dense", fin, fout], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(fout, "rb").read()
    it, ok = np.frombuffer(raw, dtype="<i4", count=2)
    x = np.frombuffer(raw, dtype="<f8", offset=8, count=n)
    assert it == int(g["n_iter"]) and ok == int(g["ok"])
    assert np.max(np.abs(x - g["x"])) <= 1e-12 * np.max(np.abs(g["x"]))
    _same_lines(r.stdout.splitlines(), g["lines"])


def test_fortran_module_jacobi_diis_and_mkrhs(oracle_mod, tmp_path):
    d = deck("caffeine")
    g = golden("caffeine")
    fout = str(tmp_path / "m.out")
    print("jd_module module caffeine", flush=True)
    r = subprocess.run([JD_MODULE, "module", os.path.join(GOLDEN, "inputs", "caffeine.txt"),
                        fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(fout, "rb").read()
    it1, ok1, it2, ok2, _ = np.frombuffer(raw, dtype="<i4", count=5)
    nsph, nylm = g["s"].shape
    ncav = int(g["ncav"])
    v = np.frombuffer(raw, dtype="<f8", offset=20)
    s1 = v[:nsph * nylm].reshape(nsph, nylm)
    s2 = v[nsph * nylm:2 * nsph * nylm].reshape(nsph, nylm)
    s0 = v[2 * nsph * nylm:3 * nsph * nylm].reshape(nsph, nylm)
    phim = v[3 * nsph * nylm:3 * nsph * nylm + ncav]
    assert ok1 == 1 and ok2 == 1
    assert it1 == int(g["niter_adjoint"]) and it2 == it1
    assert np.array_equal(s1, s0)
    tol = 10.0 ** (-d.iconv)
    assert hsnorm_rms(s2 - g["s"], d.lmax) < tol
    x, y, z, rr = d.bohr()
    o = oracle_mod.Oracle(x, y, z, rr, lmax=d.lmax, ngrid=d.ngrid, iconv=d.iconv, igrad=1,
                          eps=d.eps, eta=d.eta)
    ccav = o.get("ccav")
    phi_o = np.zeros(ncav)
    import ctypes
    lib = oracle_mod.lib()
    xm, ym, zm = x + 0.1, y - 0.05, np.array(z)
    psi_o = np.zeros((nsph, nylm))
    lib.orc_mkrhs(nsph, oracle_mod._p(np.ascontiguousarray(d.charge)),
                  oracle_mod._p(np.ascontiguousarray(xm)), oracle_mod._p(np.ascontiguousarray(ym)),
                  oracle_mod._p(np.ascontiguousarray(zm)), ncav,
                  oracle_mod._p(np.ascontiguousarray(ccav)), oracle_mod._p(phi_o), nylm,
                  oracle_mod._p(psi_o))
    assert np.max(np.abs(phim - phi_o)) <= 1e-13 * np.max(np.abs(phi_o))
    assert np.max(np.abs(phim - g["phi"])) > 1e-6   # the moved charges were used
    del ctypes
