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
