"""Graph replay of the solver's iteration groups (ddx_set_graphs).

With graphs on, the solver's rounds after the first eager group are launches
of a HIP graph of 32 rounds captured on first use (one launch covers the rest
of a solve of up to 35 iterations).  The launches inside are the eager ones,
so every result must be bitwise equal to the eager run.  Each test replays
the captured graphs many times, with eager solves, new charges and new
geometries in between, and checks every replay against the eager result.
"""
import numpy as np
import pytest

from conftest import deck, golden

pytestmark = pytest.mark.gpu


def _same(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b))


def test_graph_replays_match_eager_energy_forces():
    from ddpcm_amd.ddx import DDCosmo
    d = deck("caffeine")
    h = DDCosmo.from_input(d)
    q0 = d.charge
    q1 = 0.5 * d.charge + 0.01 * np.arange(d.charge.size) / d.charge.size
    ref = {}
    for k, q in (("a", q0), ("b", q1)):
        ref[k] = h.energy_forces(q) + (h.stats().niter_direct, h.stats().niter_adjoint)
    g = golden("caffeine")
    assert abs(ref["a"][0] - g["esolv"]) <= 1e-10 * abs(g["esolv"])
    h.set_graphs(True)
    for k in ("a", "b", "a", "b", "a"):
        q = q0 if k == "a" else q1
        e, f = h.energy_forces(q)
        st = h.stats()
        assert e == ref[k][0], k
        assert _same(f, ref[k][1]), k
        assert (st.niter_direct, st.niter_adjoint) == ref[k][2:], k
        # eager work between the replays
        h.set_graphs(False)
        e2, _ = h.energy_forces(q1 if k == "a" else q0)
        assert e2 == ref["b" if k == "a" else "a"][0]
        h.set_graphs(True)


def test_graph_replays_single_solves_and_new_geometry():
    from ddpcm_amd.ddx import DDCosmo
    d = deck("taxol")
    h = DDCosmo.from_input(d)
    phi, psi = h.mkrhs(d.charge)
    sd, ed = h.cosmo(0, phi=phi, psi=psi)
    sa, _ = h.cosmo(1, psi=psi)
    h.set_graphs(True)
    for _ in range(3):
        s, e = h.cosmo(0, phi=phi, psi=psi)
        assert _same(s, sd) and e == ed
        assert _same(h.cosmo(1, psi=psi)[0], sa)
    # moved atoms: a new setup drops the captured graphs
    x, y, z, r = d.bohr()
    rng = np.random.default_rng(7)
    dx = [v + 0.05 * rng.standard_normal(v.size) for v in (x, y, z)]
    fresh = DDCosmo(*dx, r, d.lmax, d.ngrid, iconv=d.iconv, igrad=1, eps=d.eps, eta=d.eta)
    ef, ff = fresh.energy_forces(d.charge)
    for _ in range(3):
        h.upload(*dx, r, d.charge)
        h.run()
        e, f = h.result()
        assert e == ef and _same(f, ff)
        h.set_graphs(False)
        h.upload(x, y, z, r, d.charge)
        h.run()
        h.set_graphs(True)


def test_graph_replays_batch():
    from ddpcm_amd import inputs
    from ddpcm_amd.ddx import DDCosmoBatch
    decks = inputs.batch_inputs(count=10, lo=6, hi=80, seed=4300)
    b = DDCosmoBatch.from_inputs(decks)
    qs = [dk.charge for dk in decks]
    es, fx = b.energy_forces(qs)
    nd, na = b.iterations()
    b.set_graphs(True)
    for rep in range(3):
        es2, fx2 = b.energy_forces(qs)
        nd2, na2 = b.iterations()
        assert _same(es2, es) and _same(fx2, fx), rep
        assert _same(nd2, nd) and _same(na2, na), rep
        b.set_graphs(False)
        b.energy_forces([0.5 * q for q in qs])
        b.set_graphs(True)
