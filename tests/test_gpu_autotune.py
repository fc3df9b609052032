"""The autotuner (ddx_autotune): it times the pipeline for each sweep worker
count and projection split, keeps the fastest, and the tuned handle must
still give the reference's results (energy 1e-10 relative, forces 1e-7
Eh/bohr, both iteration counts), and a tuned batch those of an untuned one.
The tuned split of the L X projection adds its partial sums in another
order; the DIIS solve of a molecule whose history ends with a near-singular
Gram matrix amplifies such last-bit changes (up to a few 1e-10 relative in
the energy), so the batch is compared at 2e-9 relative in E and 1e-8
Eh/bohr in F, with equal iteration counts.
"""
import numpy as np
import pytest

from conftest import deck, golden

pytestmark = pytest.mark.gpu


def test_autotune_single_system():
    from ddpcm_amd.ddx import DDCosmo
    d = deck("c2_small")
    g = golden("c2_small")
    h = DDCosmo.from_input(d)
    h.upload(*d.bohr(), d.charge)
    w, sp = h.autotune()
    assert w in (4, 8, 16, 32, 64) and sp in (1, 2, 4)
    for _ in range(2):
        es, fx = h.result()
        st = h.stats()
        assert abs(es - g["esolv"]) <= 1e-10 * abs(g["esolv"])
        np.testing.assert_allclose(fx, g["fx"], rtol=0, atol=1e-7)
        assert (st.niter_direct, st.niter_adjoint) == (g["niter_direct"], g["niter_adjoint"])
        h.run()


def test_autotune_batch():
    from ddpcm_amd import inputs
    from ddpcm_amd.ddx import DDCosmoBatch
    decks = inputs.batch_inputs(count=12, lo=10, hi=200, seed=4500)
    qs = [dk.charge for dk in decks]
    ref = DDCosmoBatch.from_inputs(decks)
    e0, f0 = ref.energy_forces(qs)
    b = DDCosmoBatch.from_inputs(decks)
    xyz = [dk.bohr() for dk in decks]
    b.upload(*[np.concatenate([v[k] for v in xyz]) for k in range(4)], np.concatenate(qs))
    w, sp = b.autotune()
    assert w in (4, 8, 16, 32, 64) and sp in (1, 2, 4)
    e1, f1 = b.result()
    assert np.max(np.abs(e1 - e0) / np.abs(e0)) < 2e-9
    assert np.max(np.abs(f1 - f0)) < 1e-8
    for a, c in zip(b.iterations(), ref.iterations()):
        assert np.array_equal(a, c)
