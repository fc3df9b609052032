"""The workloads at full size against the reference's own outputs.

Fixtures (tests/golden/make_golden.py, from the reference compiled from its
sources, oracle/_ref/ref_dump):
  c3.npz          20 000 synthetic atoms, lmax 10, 1202 points: E, F, both
                  iteration counts, ncav, sigma and S at every 50th sphere
  c4_sample.npz   every 32nd molecule of the 512-molecule batch (lmax 6, 302
                  points): E, F and iteration counts
  c5_frames.npz   frames 8..11 of the 1000-atom trajectory (lmax 8, 590
                  points): E, F and iteration counts at conv 1e-8, and E, F
                  converged to 1e-13 (the exact solution of the linear systems)
  protein.npz     the 24 036 atoms of the reference's own Input.txt (lmax 6,
                  110 points, conv 1e-7, eta 0.2): E, F, iteration counts
Tolerances (the reference's FP64 path): E 1e-10 relative, F 1e-7 Eh/bohr,
iteration counts equal; sigma and S within the convergence threshold in the
H-norm (hnorm, matvec.f90:281-323).  In the batch sample a few molecules
(4 of 16) end their DIIS history with near-singular Gram matrices, which
amplify the order in which sums are taken: their energies move by up to
5e-10 relative between two summation orders with the same iteration counts
(also between two batches that hold the molecule with different
neighbours).  There the energy must stay within 1e-10 relative plus a tenth
of the reference's own distance from the exact solution (0.2e-9 .. 8e-9
relative in this sample; tests/golden/c4_sample.npz holds it).  Guess modes other than D^-1 rhs stop
at another point of their own iteration; they must stay as close to the
exact solution as the convergence threshold allows: E within 1e-8 relative
and F within 1e-7 Eh/bohr of the 1e-13 solution.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden, hsnorm_rms

pytestmark = pytest.mark.gpu


def _check_ef(es, fx, g, what):
    assert abs(es - float(g["esolv"])) <= 1e-10 * abs(float(g["esolv"])), (what, es, g["esolv"])
    err = np.max(np.abs(fx - g["fx"]))
    assert err <= 1e-7, (what, err)


def test_c3_20000_atoms_against_reference():
    from ddpcm_amd import inputs
    from ddpcm_amd.ddx import DDCosmo
    g = golden("c3")
    d = inputs.config_input("c3")
    print("c3: 20000 atoms lmax 10, 1202 points, FP64", flush=True)
    h = DDCosmo.from_input(d)
    st = h.stats()
    assert st.ncav == int(g["ncav"])
    es, fx = h.energy_forces(d.charge)
    st = h.stats()
    _check_ef(es, fx, g, "c3")
    assert st.niter_direct == int(g["niter_direct"])
    assert st.niter_adjoint == int(g["niter_adjoint"])
    sub = g["sph_index"]
    tol = 10.0 ** (-d.iconv)
    assert hsnorm_rms(h.get("sigma")[sub] - g["sigma"], d.lmax) < tol
    assert hsnorm_rms(h.get("s")[sub] - g["s"], d.lmax) < tol
    # the mixed precision on the same handle: as close to the reference as
    # a solve stopped by the same threshold
    print("c3: mixed precision", flush=True)
    h.set_precision(1)
    em, fm = h.energy_forces(d.charge)
    assert abs(em - es) <= 1e-8 * abs(es), (em, es)
    assert np.max(np.abs(fm - fx)) <= 1e-7
    assert hsnorm_rms(h.get("sigma")[sub] - g["sigma"], d.lmax) < 2 * tol
    assert hsnorm_rms(h.get("s")[sub] - g["s"], d.lmax) < 2 * tol


def test_c4_full_batch_against_reference_sample():
    from ddpcm_amd import inputs
    from ddpcm_amd.ddx import DDCosmoBatch
    g = golden("c4_sample")
    decks = inputs.batch_inputs()
    print("c4: 512 molecules, %d spheres" % sum(d.n for d in decks), flush=True)
    b = DDCosmoBatch.from_inputs(decks)
    es, fx = b.energy_forces([d.charge for d in decks])
    nd, na = b.iterations()
    parts = b.split(fx)
    off = np.concatenate([[0], np.cumsum(g["nsph"])])
    for k, q in enumerate(g["q"]):
        assert decks[q].n == int(g["nsph"][k])
        eref, ex = float(g["esolv"][k]), float(g["esolv_exact"][k])
        assert abs(es[q] - eref) <= 1e-10 * abs(eref) + 0.1 * abs(eref - ex), (q, es[q], eref)
        assert np.max(np.abs(parts[q] - g["fx"][off[k]:off[k + 1]])) <= 1e-7, q
        assert nd[q] == int(g["niter_direct"][k]) and na[q] == int(g["niter_adjoint"][k]), q
        assert b.cav_off[q + 1] - b.cav_off[q] == int(g["ncav"][k]), q
    # mixed precision on the whole batch: every sampled molecule at least as
    # close to the exact solution as the reference's own result, plus 1e-9
    # of its energy terms and 1e-9 Eh/bohr (the rule of test_gpu_mixed)
    print("c4: mixed precision", flush=True)
    b.set_precision(1)
    em, fm = b.energy_forces([d.charge for d in decks])
    pm = b.split(fm)
    for k, q in enumerate(g["q"]):
        eref, ex = float(g["esolv"][k]), float(g["esolv_exact"][k])
        assert abs(em[q] - ex) <= abs(eref - ex) + 1e-9 * float(g["escale"][k]), (q, em[q], ex)
        fxx = g["fx_exact"][off[k]:off[k + 1]]
        ref_err = np.max(np.abs(g["fx"][off[k]:off[k + 1]] - fxx))
        assert np.max(np.abs(pm[q] - fxx)) <= ref_err + 1e-9, q
    assert np.max(np.abs(em - es) / np.abs(es)) <= 1e-7
    assert np.max(np.abs(fm - fx)) <= 1e-7


def _c5_frames(last):
    from ddpcm_amd import inputs
    out = []
    for k, fr in enumerate(inputs.trajectory()):
        if k > last:
            break
        out.append(fr)
    return out


def test_c5_trajectory_guess_modes():
    """All four initial-guess modes on the 1000-atom trajectory: the first
    8 frames train the network, frames 8..11 are checked against the
    reference (D^-1 rhs start: bit-level parity) and against the exact
    solution (the other starts)."""
    from ddpcm_amd import inputs
    from ddpcm_amd.ddx import (DDCosmo, DDX_GUESS_LEARNED, DDX_GUESS_LEARNED_PREVIOUS,
                               DDX_GUESS_PREVIOUS, DDX_GUESS_REFERENCE)
    g = golden("c5_frames")
    first = int(g["frames"][0])
    frames = _c5_frames(int(g["frames"][-1]))
    tb = 1.0 / inputs.TOANG
    its = {}
    for mode in (DDX_GUESS_REFERENCE, DDX_GUESS_PREVIOUS, DDX_GUESS_LEARNED,
                 DDX_GUESS_LEARNED_PREVIOUS):
        print("c5: guess mode %d" % mode, flush=True)
        q, xyz, r = frames[0]
        h = DDCosmo(xyz[:, 0] * tb, xyz[:, 1] * tb, xyz[:, 2] * tb, r * tb, lmax=8, ngrid=590)
        h.set_guess(mode)
        rows = []
        for f, (q, xyz, r) in enumerate(frames):
            h.upload(xyz[:, 0] * tb, xyz[:, 1] * tb, xyz[:, 2] * tb, r * tb, q)
            h.run()
            es, fx = h.result()
            st = h.stats()
            if f < first:
                if mode in (DDX_GUESS_LEARNED, DDX_GUESS_LEARNED_PREVIOUS):
                    h.guess_train(steps=8, lr=2e-3)
                continue
            k = f - first
            rows.append(st.niter_direct + st.niter_adjoint)
            if mode == DDX_GUESS_REFERENCE:
                _check_ef(es, fx, {"esolv": g["esolv"][k], "fx": g["fx"][k]}, "frame %d" % f)
                assert st.niter_direct == int(g["niter_direct"][k])
                assert st.niter_adjoint == int(g["niter_adjoint"][k])
            ex = float(g["esolv_exact"][k])
            assert abs(es - ex) <= 1e-8 * abs(ex), (mode, f, es, ex)
            assert np.max(np.abs(fx - g["fx_exact"][k])) <= 1e-7, (mode, f)
        its[mode] = float(np.mean(rows))
    print("c5 iterations (direct + adjoint) per mode:", its, flush=True)
    assert its[DDX_GUESS_PREVIOUS] < its[DDX_GUESS_REFERENCE]
    assert its[DDX_GUESS_LEARNED] < its[DDX_GUESS_REFERENCE]


def test_reference_input_protein_24036_atoms():
    from ddpcm_amd import inputs
    from ddpcm_amd.ddx import DDCosmo
    g = golden("protein")
    z = np.load(os.path.join(GOLDEN, "inputs", "protein.npz"), allow_pickle=False)
    hd = [int(v) for v in z["header"]]
    d = inputs.DDInput(hd[0], hd[1], hd[2], hd[3], hd[4], hd[5], float(z["eps"]),
                       float(z["eta"]), z["charge"], z["xyz"], z["rvdw"])
    print("protein: %d atoms lmax %d, %d points" % (d.n, d.lmax, d.ngrid), flush=True)
    h = DDCosmo.from_input(d)
    st = h.stats()
    assert st.max_degree <= 64 and st.ncav == int(g["ncav"])
    es, fx = h.energy_forces(d.charge)
    st = h.stats()
    _check_ef(es, fx, g, "protein")
    assert st.niter_direct == int(g["niter_direct"])
    assert st.niter_adjoint == int(g["niter_adjoint"])
