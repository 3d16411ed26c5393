"""CPU FP64 restatement of the chemistry surrogate network -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the CPU leg of the benchmark may import this module; the product
path (libfcfd_hip.so) never does.

The surrogate has no counterpart in the reference (it stands in for the macks call at
SRC/source.f90:58), so this file restates the project's own definition:
  x (64, BF16)  -> h_l = gelu(h_{l-1} W_l + b_l), l = 0..3 (256 wide) -> y = h_3 W_4 + b_4 (64 wide)
  loss = mean over rows and the first 54 output columns of (y - t)^2
with exact (erf) GELU.  `forward(..., bf16=True)` additionally rounds every activation to BF16
(round to nearest even) where the device kernels store BF16, so it follows the device path up to
FP32-vs-FP64 accumulation.  Pinned in tests/test_mlp_oracle.py by central finite differences of
the FP64 loss, the GELU closed forms, and BF16 rounding against torch's own conversion.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

IN, OUT, HID, INP, OUTP = 56, 54, 256, 64, 64
K = (INP, HID, HID, HID, HID)
N = (HID, HID, HID, HID, OUTP)


def nparams() -> int:
    return sum(k * n + n for k, n in zip(K, N))


def unpack(flat):
    """Flat parameter vector -> [(W (K x N), b (N,)), ...] in float64."""
    flat = np.asarray(flat, np.float64)
    out, o = [], 0
    for k, n in zip(K, N):
        W = flat[o:o + k * n].reshape(k, n)
        o += k * n
        b = flat[o:o + n]
        o += n
        out.append((W, b))
    assert o == flat.size
    return out


def pack(layers) -> np.ndarray:
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    nan = (u & 0x7F800000) == 0x7F800000
    r = np.where(nan, u & 0xFFFF0000, r)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(x) -> np.ndarray:
    """BF16 bit patterns (uint16) of x."""
    return (bf16(x).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(flat, X, bf16_acts=False):
    """Returns (y, cache).  With bf16_acts the weights, inputs and stored activations are rounded
    to BF16 like the device path."""
    layers = unpack(flat)
    rnd = bf16 if bf16_acts else (lambda a: np.asarray(a, np.float64))
    h = rnd(X)
    hs, zs = [h], []
    for l, (W, b) in enumerate(layers):
        z = h @ rnd(W) + b
        if l < 4:
            zs.append(z)
            h = rnd(gelu(z))
            hs.append(h)
        else:
            y = z
    return y, (hs, zs)


def loss_and_grads(flat, X, T, bf16_acts=False):
    """MSE over the first 54 columns and its gradient (flat, same order as the parameters)."""
    layers = unpack(flat)
    rnd = bf16 if bf16_acts else (lambda a: np.asarray(a, np.float64))
    y, (hs, zs) = forward(flat, X, bf16_acts)
    R = X.shape[0]
    d = np.zeros_like(y)
    d[:, :OUT] = y[:, :OUT] - T[:, :OUT]
    loss = float((d[:, :OUT] ** 2).sum() / (R * OUT))
    dz = rnd(2.0 * d / (R * OUT))
    grads = [None] * 5
    for l in range(4, -1, -1):
        W, _ = layers[l]
        gW = hs[l].T @ dz
        gb = dz.sum(axis=0)
        grads[l] = (gW, gb)
        if l > 0:
            dz = rnd((dz @ rnd(W).T) * gelu_d(zs[l - 1]))
    return loss, pack(grads)


def init_params(seed=0):
    """He-uniform initial weights (zero biases); padding rows / columns zero."""
    rng = np.random.default_rng(seed)
    layers = []
    for l, (k, n) in enumerate(zip(K, N)):
        fan_in = IN if l == 0 else k
        lim = np.sqrt(6.0 / fan_in) if l < 4 else np.sqrt(3.0 / fan_in)
        W = rng.uniform(-lim, lim, (k, n))
        if l == 0:
            W[IN:, :] = 0.0
        if l == 4:
            W[:, OUT:] = 0.0
        layers.append((W, np.zeros(n)))
    return pack(layers).astype(np.float32)
