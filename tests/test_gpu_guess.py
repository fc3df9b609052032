"""The learned initial guess: its BF16 MFMA product kernel against numpy."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(16, 16, 16), (1000, 128, 97), (128, 81, 1000), (37, 5, 3)])
def test_gemm_bf16_against_numpy(shape):
    from ddpcm_amd.ddx import bf16_bits, bf16_value, gemm_bf16
    M, N, K = shape
    rng = np.random.default_rng(M * 7 + N)
    a = bf16_bits(rng.standard_normal((M, K)))
    b = bf16_bits(rng.standard_normal((K, N)))
    ref = bf16_value(a).astype(np.float64) @ bf16_value(b).astype(np.float64)
    c = gemm_bf16(a, b)
    scale = np.sqrt(K) + 1.0
    assert np.max(np.abs(c - ref)) <= 1e-5 * scale * max(1.0, np.max(np.abs(ref)))


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    sg = 1.0 / (1.0 + np.exp(-z))
    return sg * (1.0 + z * (1.0 - sg))


def _mlp_reference(F, T, W, bf):
    """CPU restatement of the network's forward and backward pass, with the
    same BF16 roundings of the product operands (the device's FP32 sums are
    taken in float64 here)."""
    W1, b1, W2, b2, W3, b3 = [w.astype(np.float64) for w in W]
    r = lambda x: bf(x).astype(np.float64)
    Fv = r(F)
    Z1 = Fv @ r(W1) + b1
    H1 = r(_silu(Z1))
    Z2 = H1 @ r(W2) + b2
    H2 = r(_silu(Z2))
    Y = H2 @ r(W3) + b3
    n, out = T.shape
    dY = 2.0 * (Y - T) / (n * out)
    dYb = r(dY)
    g3 = H2.T @ dYb
    gb3 = dY.sum(0)
    dZ2 = (dYb @ r(W3).T) * _dsilu(Z2)
    dZ2b = r(dZ2)
    g2 = H1.T @ dZ2b
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2b @ r(W2).T) * _dsilu(Z1)
    g1 = Fv.T @ r(dZ1)
    gb1 = dZ1.sum(0)
    loss = float(np.mean((Y - T) ** 2))
    return Y, [g1, gb1, g2, gb2, g3, gb3], loss


@pytest.mark.parametrize("n,nin,nout", [(200, 97, 81), (64, 65, 49)])
def test_mlp_gradients_against_cpu(n, nin, nout):
    """Forward output, loss and all six gradients of the BF16 MFMA network
    against the CPU restatement."""
    from ddpcm_amd.ddx import bf16_bits, bf16_value, mlp_grad
    rng = np.random.default_rng(n + nin)
    bf = lambda x: bf16_value(bf16_bits(x))
    F = rng.standard_normal((n, nin)).astype(np.float32)
    T = (0.3 * rng.standard_normal((n, nout))).astype(np.float32)
    W = [rng.standard_normal((nin, 128)) / np.sqrt(nin), 0.1 * rng.standard_normal(128),
         rng.standard_normal((128, 128)) / np.sqrt(128), 0.1 * rng.standard_normal(128),
         rng.standard_normal((128, nout)) / np.sqrt(128), 0.1 * rng.standard_normal(nout)]
    W = [w.astype(np.float32) for w in W]
    Y, G, loss = mlp_grad(bf16_bits(F), T, W)
    Yr, Gr, lr = _mlp_reference(F, T, W, bf)
    assert np.max(np.abs(Y - Yr)) <= 1e-3 * np.max(np.abs(Yr))
    assert abs(loss - lr) <= 1e-3 * lr
    for g, gr in zip(G, Gr):
        # BF16 roundings of intermediate values may land on the other side
        # of a rounding boundary: a few elements differ by one BF16 ulp
        err = np.abs(g - gr)
        assert np.max(err) <= 2e-2 * np.max(np.abs(gr))
        assert np.mean(err) <= 2e-3 * np.mean(np.abs(gr))


def _trajectory(n, frames, seed):
    from ddpcm_amd import inputs
    return [(q, xyz, r) for q, xyz, r in inputs.trajectory(n=n, frames=frames, seed=seed)]


def _run_modes(frames, lmax, ngrid, mode, train=0):
    from ddpcm_amd import inputs
    from ddpcm_amd.ddx import DDCosmo
    q, xyz, r = frames[0]
    tb = 1.0 / inputs.TOANG
    h = DDCosmo(xyz[:, 0] * tb, xyz[:, 1] * tb, xyz[:, 2] * tb, r * tb, lmax=lmax, ngrid=ngrid)
    h.set_guess(mode)
    out = []
    for f, (q, xyz, r) in enumerate(frames):
        h.upload(xyz[:, 0] * tb, xyz[:, 1] * tb, xyz[:, 2] * tb, r * tb, q)
        h.run()
        es, fx = h.result()
        st = h.stats()
        out.append((es, fx, st.niter_direct, st.niter_adjoint))
        if f < train:
            h.guess_train(steps=8, lr=2e-3)
    return out


def test_guess_modes_on_a_trajectory():
    """Over the frames of a trajectory the previous-frame and the learned
    guess (trained on the first frames' solutions) take fewer iterations than
    D^-1 rhs, and every frame's energy and forces stay those of the
    reference's start to well inside the convergence threshold."""
    from ddpcm_amd.ddx import DDX_GUESS_LEARNED, DDX_GUESS_PREVIOUS, DDX_GUESS_REFERENCE
    frames = _trajectory(60, 36, 77)
    ref = _run_modes(frames, 6, 302, DDX_GUESS_REFERENCE)
    prev = _run_modes(frames, 6, 302, DDX_GUESS_PREVIOUS)
    lrn = _run_modes(frames, 6, 302, DDX_GUESS_LEARNED, train=24)
    for a, b, c in zip(ref, prev, lrn):
        for o in (b, c):
            assert abs(o[0] - a[0]) <= 1e-7 * abs(a[0])
            assert np.max(np.abs(o[1] - a[1])) <= 1e-6
    it = lambda rows: float(np.mean([x[2] + x[3] for x in rows[24:]]))
    assert it(prev) < it(ref)
    # what the trained network gains over the method without it
    assert it(lrn) < it(ref)
