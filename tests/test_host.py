"""Host-side logic that runs without a GPU: the Input.txt format, the
synthetic generator, and that libddx.so loads and exports every entry point
include/ddx.h declares (no GPU call is made)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT


def test_read_reference_decks():
    from ddpcm_amd.inputs import read_input
    d = read_input(os.path.join(GOLDEN, "inputs", "formaldehyde.txt"))
    assert (d.iprint, d.nproc, d.lmax, d.ngrid, d.iconv, d.igrad) == (4, 1, 6, 110, 7, 1)
    assert d.eps == 78.3553 and d.eta == 0.2 and d.n == 4
    np.testing.assert_array_equal(d.charge, [0.08130, -0.20542, 0.06206, 0.06206])
    np.testing.assert_array_equal(d.xyz[0], [0.0, 0.0, -1.16691])


def test_only_declared_atoms_are_read(tmp_path):
    """The reference reads n records even when more follow (Input.txt:9)."""
    from ddpcm_amd.inputs import read_input
    lines = ["1", "2", "6", "110", "7", "1", "78.3553", "0.2", "2   # n"]
    lines += ["0.1 0 0 0 1.5", "-0.1 1 0 0 1.5", "9 9 9 9 9"]
    p = tmp_path / "Input.txt"
    p.write_text("\n".join(lines) + "\n")
    d = read_input(str(p))
    assert d.n == 2 and d.charge[1] == -0.1


def test_write_read_roundtrip(tmp_path):
    from ddpcm_amd import inputs
    d = inputs.config_input("c1")
    p = tmp_path / "deck.txt"
    inputs.write_input(str(p), d)
    e = inputs.read_input(str(p))
    np.testing.assert_array_equal(e.xyz, d.xyz)
    np.testing.assert_array_equal(e.charge, d.charge)
    np.testing.assert_array_equal(e.rvdw, d.rvdw)
    assert (e.lmax, e.ngrid, e.iconv, e.igrad, e.eps, e.eta) == (6, 302, 8, 0, 78.39, 0.1)


def test_bohr_conversion():
    from ddpcm_amd.inputs import DDInput, TOANG
    d = DDInput(0, 1, 6, 110, 7, 1, 78.0, 0.1, np.ones(1), np.array([[1.0, 2.0, 3.0]]),
                np.array([1.5]))
    x, y, z, r = d.bohr()
    assert x[0] == 1.0 * (1.0 / TOANG) and r[0] == 1.5 * (1.0 / TOANG)


def test_synthetic_generator():
    from ddpcm_amd import inputs
    q, xyz, r = inputs.synthetic_molecule(2000, seed=2000)
    assert abs(q.sum()) < 1e-12 and np.all(np.abs(q) <= 0.62)
    radius = (2000 / (0.1 * 4 * np.pi / 3)) ** (1 / 3)
    assert np.all(np.linalg.norm(xyz, axis=1) <= radius + 0.3 * np.sqrt(3) + 1e-9)
    # no two atoms share a lattice site: minimum separation >= 1.5 - 2*0.3*sqrt(3)
    d = np.linalg.norm(xyz[:200, None] - xyz[None, :200], axis=-1) + np.eye(200) * 99
    assert d.min() > 1.5 - 0.6 * np.sqrt(3)
    radii = set(np.round(r / 1.1, 6))
    assert radii <= {1.2, 1.7, 1.55, 1.52}
    frac_h = np.mean(np.isclose(r, 1.2 * 1.1))
    assert 0.45 < frac_h < 0.55
    q2, xyz2, _ = inputs.synthetic_molecule(2000, seed=2000)
    np.testing.assert_array_equal(xyz, xyz2)


def test_batch_and_trajectory():
    from ddpcm_amd import inputs
    sizes = inputs.batch_sizes()
    assert len(sizes) == 512 and sizes.min() >= 20 and sizes.max() <= 800
    frames = list(zip(range(3), inputs.trajectory(n=50, frames=3)))
    d1 = frames[1][1][1] - frames[0][1][1]
    assert 0.03 < d1.std() < 0.07


def _declared_symbols():
    txt = open(os.path.join(ROOT, "include", "ddx.h")).read()
    return sorted(set(re.findall(r"\b(ddx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_entry_point():
    lib_path = os.path.join(ROOT, "ddpcm_amd", "libddx.so")
    if not os.path.exists(lib_path):
        pytest.fail("libddx.so is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(lib_path)
    names = _declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    from ddpcm_amd import ddx
    assert set(ddx.EXPORTS) == set(names)


def test_last_error_without_gpu():
    from ddpcm_amd import ddx
    L = ddx.load_library()
    assert isinstance(L.ddx_last_error(), bytes)
    # argument validation happens before any device call
    h = ctypes.c_void_p()
    z = np.zeros(1)
    rc = L.ddx_init(ctypes.byref(h), 0, ddx._p(z), ddx._p(z), ddx._p(z), ddx._p(z), 0, 6, 110, 8,
                    1, 78.39, 0.1)
    assert rc == ddx.DDX_ERR_ARG
    rc = L.ddx_init(ctypes.byref(h), 1, ddx._p(z), ddx._p(z), ddx._p(z), ddx._p(z), 0, 40, 110, 8,
                    1, 78.39, 0.1)
    assert rc == ddx.DDX_ERR_ARG and b"lmax" in L.ddx_last_error()
