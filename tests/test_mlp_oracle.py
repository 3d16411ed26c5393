"""Pins the CPU restatement of the surrogate network (oracle/mlp_oracle.py): gradients against
central finite differences of the FP64 loss, GELU against its closed forms, BF16 rounding against
torch's own float32 -> bfloat16 conversion (round to nearest even)."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def small_problem(seed=1, rows=16):
    rng = np.random.default_rng(seed)
    flat = M.init_params(seed).astype(np.float64)
    flat += rng.normal(0, 0.05, flat.size) * (flat != 0)     # non-zero biases too
    lay = M.unpack(flat)
    for l, (W, b) in enumerate(lay):
        b[:] = rng.normal(0, 0.1, b.size)
        if l == 4:
            b[M.OUT:] = 0
    flat = M.pack(lay)
    X = np.zeros((rows, M.INP))
    X[:, :M.IN] = rng.normal(size=(rows, M.IN))
    T = rng.normal(size=(rows, M.OUTP))
    return flat, X, T


def test_gradients_match_finite_differences():
    flat, X, T = small_problem()
    loss, g = M.loss_and_grads(flat, X, T)
    rng = np.random.default_rng(7)
    # probe entries of every layer's weights and biases
    offs, o = [], 0
    for k, n in zip(M.K, M.N):
        w_idx = o + rng.integers(0, M.IN if o == 0 else k, 6) * n + rng.integers(0, min(n, M.OUT), 6)
        b_idx = o + k * n + rng.integers(0, min(n, M.OUT), 3)
        offs += list(w_idx) + list(b_idx)
        o += k * n + n
    h = 1e-6
    for i in offs:
        fp, fm = flat.copy(), flat.copy()
        fp[i] += h
        fm[i] -= h
        num = (M.loss_and_grads(fp, X, T)[0] - M.loss_and_grads(fm, X, T)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, num, g[i])


def test_padding_gets_no_gradient():
    flat, X, T = small_problem()
    _, g = M.loss_and_grads(flat, X, T)
    gl = M.unpack(g)
    assert np.all(gl[0][0][M.IN:, :] == 0)          # zero input columns
    assert np.all(gl[4][0][:, M.OUT:] == 0)         # outputs outside the loss
    assert np.all(gl[4][1][M.OUT:] == 0)


def test_gelu_closed_forms():
    assert M.gelu(0.0) == 0.0 and M.gelu_d(0.0) == 0.5
    z = np.linspace(-6, 6, 101)
    # gelu(z) - gelu(-z) = z  and  gelu'(z) + gelu'(-z) = 1
    np.testing.assert_allclose(M.gelu(z) - M.gelu(-z), z, atol=1e-14)
    np.testing.assert_allclose(M.gelu_d(z) + M.gelu_d(-z), 1.0, atol=1e-14)
    # derivative by central differences
    np.testing.assert_allclose((M.gelu(z + 1e-6) - M.gelu(z - 1e-6)) / 2e-6, M.gelu_d(z), atol=1e-8)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(size=4096) * 10.0 ** rng.integers(-30, 30, 4096),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e38, -3.0e38, 1e-40],
                                 np.float32)]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    np.testing.assert_array_equal(M.bf16(x), ref)
    bits = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    np.testing.assert_array_equal(M.bf16_bits(x), bits)


def test_forward_shapes_and_bf16_path_close_to_fp64():
    flat, X, T = small_problem(rows=64)
    y64, _ = M.forward(flat, X)
    y16, _ = M.forward(flat, X, bf16_acts=True)
    assert y64.shape == (64, M.OUTP)
    rel = np.linalg.norm(y16 - y64) / np.linalg.norm(y64)
    assert 1e-5 < rel < 3e-2
