"""Pins the C oracle (oracle/ddcosmo_oracle.c) before it is trusted as the checker.

Three kinds of evidence:
  * the compiled reference's own full-precision outputs (tests/golden/*.npz)
    and the six shipped logs of its test harness (tests/golden/logs);
  * closed forms and textbook routines: real spherical harmonics from
    scipy's associated Legendre functions, Lebedev quadrature exactness,
    the quintic smoothstep, the Born energy of a single ion, Coulomb sums in
    extended precision;
  * invariants that a dropped term or swapped index breaks: adjointness of
    L and L*, a dense direct solve of L, translation invariance of the
    forces, forces as central differences of the energy.
"""
import math
import os
import re

import numpy as np
import pytest
from scipy.special import lpmv

from conftest import GOLDEN, deck, golden, hsnorm_rms

MOLS = ["formaldehyde", "benzene", "caffeine", "taxol", "ala25", "crambin"]
FULL = ["formaldehyde", "benzene", "caffeine", "c1", "c1_forces", "ion"]


def make(oracle_mod, name, **kw):
    return oracle_mod.Oracle.from_input(deck(name), **kw)


# ---------------------------------------------------------------- fixtures
@pytest.mark.parametrize("name", FULL)
def test_setup_matches_reference_dump(oracle_mod, name):
    o = make(oracle_mod, name)
    g = golden(name)
    for k in ("grid", "w", "basis", "facs", "facl", "fi", "ui", "ccav"):
        np.testing.assert_array_equal(o.get(k), g[k], err_msg=k)
    if "zi" in g:
        np.testing.assert_array_equal(o.get("zi"), g["zi"])
    assert np.array_equal(o.get("inl") + 1, g["inl"]) and np.array_equal(o.get("nl") + 1, g["nl"])


@pytest.mark.parametrize("name", MOLS + ["c1", "c1_forces", "c2_small", "ion",
                                          "c3_small", "l12", "l16"])
def test_operators_and_solution_match_reference_dump(oracle_mod, name):
    d = deck(name)
    o = make(oracle_mod, name)
    g = golden(name)
    v = g["probe_v"]
    np.testing.assert_allclose(o.lx(v), g["probe_lx"], rtol=0, atol=1e-14 * np.abs(g["probe_lx"]).max())
    np.testing.assert_allclose(o.lstarx(v), g["probe_lstarx"], rtol=0,
                               atol=1e-14 * np.abs(g["probe_lstarx"]).max())
    assert o.hnorm(v) == pytest.approx(float(g["hnorm_probe"]), rel=1e-15)
    phi, psi = o.mkrhs(d.charge)
    np.testing.assert_allclose(phi, g["phi"], rtol=1e-14, atol=0)
    es, fx, hd, ha = o.energy_forces(d.charge)
    assert es == pytest.approx(float(g["esolv"]), rel=1e-13)
    assert hd.n_iter == g["niter_direct"]
    if d.igrad:
        assert ha.n_iter == g["niter_adjoint"]
        np.testing.assert_allclose(fx, g["fx"], rtol=0, atol=1e-12)


def _log_values(path):
    txt = open(path).read()
    e = float(re.search(r"solvation energy \(kcal/mol\):\s*([-0-9.]+)", txt).group(1))
    forces = []
    tail = txt.split("ddCOSMO forces (atomic units):")[-1].splitlines()[2:]
    for line in tail:
        t = line.split()
        if len(t) == 4 and t[0].isdigit():
            forces.append([float(x) for x in t[1:]])
    blocks = re.split(r"solution time of ddCOSMO", txt)
    iters = [len(re.findall(r"iter=", b)) for b in blocks[:2]]
    return e, np.array(forces), iters


@pytest.mark.parametrize("name", MOLS)
def test_shipped_logs(oracle_mod, name):
    """Energies (f14.6 kcal/mol), forces (f16.8 Eh/bohr) and iteration counts
    of the reference's six golden logs."""
    d = deck(name)
    o = make(oracle_mod, name)
    es, fx, hd, ha = o.energy_forces(d.charge)
    e_log, f_log, it_log = _log_values(os.path.join(GOLDEN, "logs", name + ".log"))
    assert round(es * 627.509469, 6) == pytest.approx(e_log, abs=1.5e-6)
    np.testing.assert_allclose(fx, f_log, rtol=0, atol=1.5e-8)
    assert [hd.n_iter, ha.n_iter] == it_log


# ---------------------------------------------------------------- closed forms
def _real_ylm(lmax, s):
    """Real harmonics without Condon-Shortley phase from scipy's P_l^m."""
    x, y, z = s
    th = math.acos(max(-1.0, min(1.0, z)))
    ph = math.atan2(y, x)
    out = np.zeros((lmax + 1) ** 2)
    for l in range(lmax + 1):
        for m in range(0, l + 1):
            p = lpmv(m, l, z) * (-1) ** m          # drop scipy's (-1)^m
            n = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - m) / math.factorial(l + m))
            if m == 0:
                out[l * l + l] = n * p
            else:
                out[l * l + l + m] = math.sqrt(2) * n * p * math.cos(m * ph)
                out[l * l + l - m] = math.sqrt(2) * n * p * math.sin(m * ph)
    return out


def test_ylm_closed_form(oracle_mod):
    o = make(oracle_mod, "c2_small")   # lmax 8
    rng = np.random.default_rng(5)
    pts = rng.standard_normal((40, 3))
    pts /= np.linalg.norm(pts, axis=1)[:, None]
    pts = np.vstack([pts, [[0, 0, 1.0], [0, 0, -1.0], [1.0, 0, 0]]])
    for s in pts:
        np.testing.assert_allclose(o.ylmbas(s), _real_ylm(o.lmax, s), rtol=0, atol=2e-14)


@pytest.mark.parametrize("ngrid,deg", [(110, 17), (302, 29), (590, 41), (1202, 59)])
def test_lebedev_quadrature(oracle_mod, ngrid, deg):
    w, g = oracle_mod.lebedev(ngrid)
    assert abs(w.sum() - 4 * math.pi) < 1e-13
    np.testing.assert_allclose(np.linalg.norm(g, axis=1), 1.0, atol=1e-15)
    lmax = deg // 2
    Y = np.array([_real_ylm(lmax, s) for s in g])
    gram = (Y * w[:, None]).T @ Y
    np.testing.assert_allclose(gram, np.eye(Y.shape[1]), atol=1e-13)


def test_lebedev_sizes(oracle_mod):
    sizes = [6, 14, 26, 38, 50, 74, 86, 110, 146, 170, 194, 230, 266, 302, 350, 434, 590, 770,
             974, 1202, 1454, 1730, 2030, 2354, 2702, 3074, 3470, 3890, 4334, 4802, 5294, 5810]
    for n in sizes:
        w, g = oracle_mod.lebedev(n)
        assert abs(w.sum() - 4 * math.pi) < 1e-12
    assert oracle_mod.snap_ngrid(300) == 302 and oracle_mod.snap_ngrid(1) == 6
    assert oracle_mod.snap_ngrid(100000) == 5810 and oracle_mod.snap_ngrid(98) == 86


def test_switching_function(oracle_mod):
    for eta in (0.1, 0.2, 0.5):
        for t in np.linspace(0.0, 1.2, 241):
            u = (1.0 - t) / eta
            ref = 1.0 if t <= 1 - eta else (0.0 if t >= 1 else u ** 3 * (10 - 15 * u + 6 * u * u))
            assert oracle_mod.fsw(t, eta) == pytest.approx(ref, abs=1e-12)
            h = 1e-6
            if abs(t - 1) > 2 * h and abs(t - 1 + eta) > 2 * h:
                fd = (oracle_mod.fsw(t + h, eta) - oracle_mod.fsw(t - h, eta)) / (2 * h)
                assert oracle_mod.dfsw(t, eta) == pytest.approx(fd, abs=1e-6)


def test_dbasis_is_surface_gradient(oracle_mod):
    o = make(oracle_mod, "c1")   # lmax 6
    rng = np.random.default_rng(11)
    for _ in range(20):
        s = rng.standard_normal(3)
        s /= np.linalg.norm(s)
        b, d = o.dbasis(s)
        np.testing.assert_allclose(b, o.ylmbas(s), atol=1e-15)
        h = 1e-6
        fd = np.zeros_like(d)
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            fp, fm = s + e, s - e
            fd[:, k] = (o.ylmbas(fp / np.linalg.norm(fp)) - o.ylmbas(fm / np.linalg.norm(fm))) / (2 * h)
        np.testing.assert_allclose(d, fd, atol=1e-7)


def test_born_ion(oracle_mod):
    """One ion: L is diagonal and E = -1/2 (eps-1)/eps q^2 / R exactly."""
    for q, r_ang in ((1.0, 1.5), (-2.0, 2.1)):
        x = y = z = np.zeros(1)
        r = np.array([r_ang / 0.52917721092])
        o = oracle_mod.Oracle(x, y, z, r, lmax=6, ngrid=302, iconv=10, igrad=1, eps=78.39)
        es, fx, hd, ha = o.energy_forces(np.array([q]))
        assert es == pytest.approx(-0.5 * (78.39 - 1) / 78.39 * q * q / r[0], rel=1e-13)
        assert hd.n_iter == 1 and np.max(np.abs(fx)) < 1e-14


def test_coulomb_extended_precision(oracle_mod):
    rng = np.random.default_rng(2)
    src = rng.uniform(-1, 1, 300)
    csrc = rng.uniform(-5, 5, (300, 3))
    ctrg = rng.uniform(-5, 5, (50, 3)) + 12.0
    ef = oracle_mod.efld(src, csrc, ctrg)
    d = ctrg[:, None, :].astype(np.longdouble) - csrc[None, :, :].astype(np.longdouble)
    r = np.sqrt((d * d).sum(-1))
    ref = (src[None, :, None] * d / (r ** 3)[..., None]).sum(1)
    np.testing.assert_allclose(ef, ref.astype(np.float64), rtol=1e-13, atol=1e-15)


# ---------------------------------------------------------------- invariants
@pytest.mark.parametrize("name", ["caffeine", "c1", "c2_small"])
def test_adjointness(oracle_mod, name):
    o = make(oracle_mod, name)
    rng = np.random.default_rng(4)
    u = rng.standard_normal((o.nsph, o.nylm))
    v = rng.standard_normal((o.nsph, o.nylm))
    a = float(np.sum(u * o.lx(v)))
    b = float(np.sum(o.lstarx(u) * v))
    assert abs(a - b) <= 1e-12 * (abs(a) + abs(b))


def test_dense_solve_of_L(oracle_mod):
    """Assemble L = D + O column by column and solve it directly: the
    Jacobi/DIIS solution must agree to its convergence threshold."""
    d = deck("benzene")
    o = oracle_mod.Oracle.from_input(d, iconv=12)
    n = o.nsph * o.nylm
    facl = o.get("facl")
    L = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        col = o.lx(e).ravel()
        col[k] += 1.0 / facl[k % o.nylm]
        L[:, k] = col
    phi, psi = o.mkrhs(d.charge)
    g = o.wghpot(phi)
    rhs = np.array([o.intrhs(g[i]) for i in range(o.nsph)]).ravel()
    x = np.linalg.solve(L, rhs).reshape(o.nsph, o.nylm)
    sigma, es, hd, rc = o.cosmo(False, phi=phi, psi=psi)
    assert rc == 0
    assert hsnorm_rms(sigma - x, o.lmax) < 1e-11
    # the adjoint operator is the transpose
    xs = np.linalg.solve(L.T, psi.ravel()).reshape(o.nsph, o.nylm)
    s, _, ha, rc = o.cosmo(True, psi=psi)
    assert hsnorm_rms(s - xs, o.lmax) < 1e-11


@pytest.mark.parametrize("name", ["formaldehyde", "c1_forces"])
def test_forces_are_energy_gradient(oracle_mod, name):
    d = deck(name)
    x, y, z, r = d.bohr()
    kw = dict(lmax=d.lmax, ngrid=d.ngrid, iconv=11, igrad=1, eps=d.eps, eta=d.eta)
    o = oracle_mod.Oracle(x, y, z, r, **kw)
    es, fx, _, _ = o.energy_forces(d.charge)
    assert np.max(np.abs(fx.sum(axis=0))) < 1e-10 * max(1.0, np.abs(fx).max())
    h = 1e-4
    for i in range(min(3, o.nsph)):
        for k in range(3):
            c = [x.copy(), y.copy(), z.copy()]
            c[k][i] += h
            ep = oracle_mod.Oracle(*c, r, **kw).energy_forces(d.charge)[0]
            c[k][i] -= 2 * h
            em = oracle_mod.Oracle(*c, r, **kw).energy_forces(d.charge)[0]
            grad = (ep - em) / (2 * h)
            # printed forces are -dE/dx (main.f90:275-282)
            assert fx[i, k] == pytest.approx(-grad, abs=2e-7 + 1e-4 * abs(grad))


def test_nonconvergence_reported(oracle_mod):
    d = deck("taxol")
    o = oracle_mod.Oracle.from_input(d)
    phi, psi = o.mkrhs(d.charge)
    g = o.wghpot(phi)
    # max 3 iterations: the reference returns n_iter + 1 and ok = false
    import ctypes as C
    from oracle import Hist, _p
    rhs = np.array([o.intrhs(g[i]) for i in range(o.nsph)])
    x = o.ldm1x(rhs)
    hist = Hist()
    ok = o.L.orc_jacobi_diis(o.h, rhs.size, 25, 1e-7, _p(rhs), _p(x), 3,
                             C.cast(o.L.orc_lx, C.c_void_p), C.byref(hist))
    assert ok == 0 and hist.n_iter == 4 and hist.ok == 0


@pytest.mark.parametrize("name", ["slow", "norm1", "norm2", "norm3", "plain", "fail"])
def test_jacobi_diis_plugin_interface_against_reference(oracle_mod, name):
    """The oracle's jacobi_diis with callbacks (every norm choice, plain
    Jacobi, the drop-10 rule past 25 iterations, a run out of iterations)
    against the reference's own jacobi_diis.f90 on dense systems
    (oracle/jd_dump.f90): bit for bit, and the printed norms."""
    z = golden("jacobi_dense")
    g = {k[len(name) + 1:]: z[k] for k in z if k.startswith(name + "_")}
    n, dmax, nrm, nit = (int(v) for v in g["params"])
    cb = oracle_mod.dense_callbacks(g["a"], g["d"])
    x, it, ok, norms = oracle_mod.jacobi_diis_ops(n, dmax, nrm, float(g["tol"]), g["rhs"],
                                                  g["x0"], nit, *cb)
    assert it == int(g["n_iter"]) and ok == int(g["ok"])
    assert np.array_equal(x, g["x"])
    lines = [l for l in g["lines"] if "iter=" in l]
    assert len(lines) == len(norms)
    for line, (rms, mx) in zip(lines, norms):
        vals = [float(t.replace("D", "E")) for t in line.split()[3 if nrm < 4 else 3:]
                if "D" in t]
        want = {1: [mx], 2: [rms], 3: [rms, mx], 4: [rms]}[nrm]
        for v, w in zip(vals, want):
            assert abs(v - w) <= 0.5001e-4 * 10.0 ** np.ceil(np.log10(abs(w))), (line, w)
