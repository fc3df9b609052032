"""Batched solves: many molecules on one handle, each as if it ran alone.

Every instance of a batch must give what the reference gives for that
molecule on its own (main.f90:164-265): the energy, the forces and both
iteration counts.  Checked against
  * the reference's own full-precision outputs for its six test molecules,
    which share one set of solver settings and so form a batch;
  * the oracle, run on each instance separately, for an uneven synthetic
    batch (the batched workload's generator at small sizes);
  * the same GPU library on a plain handle per instance.
"""
import numpy as np
import pytest

from conftest import deck, golden

pytestmark = pytest.mark.gpu

MOLS = ["formaldehyde", "benzene", "caffeine", "taxol", "ala25", "crambin"]


def test_reference_molecules_as_one_batch():
    from ddpcm_amd.ddx import DDCosmoBatch
    # largest first, so that instance offsets are exercised in both directions
    names = ["crambin", "formaldehyde", "ala25", "benzene", "taxol", "caffeine"]
    decks = [deck(n) for n in names]
    b = DDCosmoBatch.from_inputs(decks)
    es, fx = b.energy_forces([d.charge for d in decks])
    nd, na = b.iterations()
    for q, (n, f) in enumerate(zip(names, b.split(fx))):
        g = golden(n)
        assert abs(es[q] - g["esolv"]) <= 1e-10 * abs(g["esolv"]), n
        np.testing.assert_allclose(f, g["fx"], rtol=0, atol=1e-7, err_msg=n)
        assert nd[q] == g["niter_direct"] and na[q] == g["niter_adjoint"], n
        assert b.cav_off[q + 1] - b.cav_off[q] == g["phi"].size, n


def test_synthetic_batch_against_oracle_and_single_handles(oracle_mod):
    from ddpcm_amd import inputs
    from ddpcm_amd.ddx import DDCosmo, DDCosmoBatch
    decks = inputs.batch_inputs(count=12, lo=6, hi=90, seed=4100)
    b = DDCosmoBatch.from_inputs(decks)
    es, fx = b.energy_forces([d.charge for d in decks])
    nd, na = b.iterations()
    for q, (d, f) in enumerate(zip(decks, b.split(fx))):
        o = oracle_mod.Oracle.from_input(d)
        eo, fo, hd, ha = o.energy_forces(d.charge)
        assert abs(es[q] - eo) <= 1e-10 * abs(eo), q
        np.testing.assert_allclose(f, fo, rtol=0, atol=1e-7, err_msg=str(q))
        assert nd[q] == hd.n_iter and na[q] == ha.n_iter, q
        h = DDCosmo.from_input(d)
        e1, f1 = h.energy_forces(d.charge)
        assert abs(es[q] - e1) <= 1e-12 * abs(e1), q
        np.testing.assert_allclose(f, f1, rtol=0, atol=1e-11, err_msg=str(q))
        # translation invariance per molecule
        assert np.max(np.abs(f.sum(axis=0))) < 1e-8 * max(1.0, np.max(np.abs(f)))


def test_batch_resident_path_and_single_calls_refused():
    from ddpcm_amd import inputs
    from ddpcm_amd.ddx import DDCosmoBatch, DDXError
    decks = inputs.batch_inputs(count=5, lo=8, hi=40, seed=4200)
    b = DDCosmoBatch.from_inputs(decks)
    es, fx = b.energy_forces([d.charge for d in decks])
    xyz = [dk.bohr() for dk in decks]
    cat = [np.concatenate([v[k] for v in xyz]) for k in range(4)]
    b.upload(*cat, np.concatenate([dk.charge for dk in decks]))
    b.run()
    es2, fx2 = b.result()
    assert np.array_equal(es, es2) and np.array_equal(fx, fx2)
    with pytest.raises(DDXError):
        DDCosmoBatch.__mro__[1].energy_forces(b, np.concatenate([dk.charge for dk in decks]))


def test_large_batch_matches_single_handles():
    """A batch large enough (about 30 000 spheres) for the expansion's wide
    launch shape (four sphere tiles per wave): each molecule as on its own
    handle, whose small grid keeps one tile per wave.  At this size the
    all-pairs source chunks and the extrapolation blocks are wider than a
    single molecule's, so partial sums group differently; late DIIS steps
    amplify that to about 1e-10 of the energy (1e-9 and 1e-8 Eh/bohr here,
    iteration counts within one)."""
    from ddpcm_amd import inputs
    from ddpcm_amd.ddx import DDCosmo, DDCosmoBatch
    decks = inputs.batch_inputs(count=200, lo=100, hi=200, seed=4600)
    b = DDCosmoBatch.from_inputs(decks)
    assert ((b.ngrid + 63) // 64) * ((b.nsph + 63) // 64) >= 2048
    es, fx = b.energy_forces([d.charge for d in decks])
    nd, na = b.iterations()
    for q in range(0, 200, 7):
        d = decks[q]
        h = DDCosmo.from_input(d)
        e1, f1 = h.energy_forces(d.charge)
        st = h.stats()
        assert abs(es[q] - e1) <= 1e-9 * abs(e1), q
        np.testing.assert_allclose(b.split(fx)[q], f1, rtol=0, atol=1e-8, err_msg=str(q))
        assert abs(nd[q] - st.niter_direct) <= 1 and abs(na[q] - st.niter_adjoint) <= 1, q
    # the FP32 expansion of the mixed precision in the same wide shape
    b.set_precision(1)
    em, fm = b.energy_forces([d.charge for d in decks])
    assert np.max(np.abs(em - es) / np.abs(es)) < 1e-7
    assert np.max(np.abs(fm - fx)) < 1e-7
