"""Pins of the FP64 surrogate-network oracle: the truncated transforms against
numpy.fft, every gradient against central finite differences, the zero-spectral
limit against a per-voxel MLP, and the parameter budget."""

import numpy as np
import pytest

from oracle import fno_oracle as F

SMALL = F.Spec(T=8, Kx=4, Ky=4, Kz=3, Np=10, C=4, L=2, H=5, Cin=3)


def test_parameter_budget_default():
    s = F.Spec()
    n = F.n_params(s)
    assert 1.9e6 < n < 2.5e6
    # spectral weights dominate: L * Kx*Ky*Kz * C*C * 2 reals
    assert n - s.L * s.Kx * s.Ky * s.Kz * s.C * s.C * 2 < 10_000


def test_forward_transform_equals_numpy_fft():
    s = SMALL
    rng = np.random.default_rng(0)
    v = rng.standard_normal((2, s.T, s.T, s.T, s.C))
    Ex, Ey, Ez, _ = F.tables(s)
    Z = np.einsum('kz,bxyzc->bxykc', Ez, v)
    Yf = np.einsum('qy,bxykc->bxqkc', Ey, Z)
    Xf = np.einsum('px,bxqkc->bpqkc', Ex, Yf)
    ref = np.fft.fftn(v, s=(s.Np, s.Np, s.Np), axes=(1, 2, 3))
    kx = F.signed_modes(s.Kx) % s.Np
    ky = F.signed_modes(s.Ky) % s.Np
    kz = np.arange(s.Kz)
    want = ref[:, kx][:, :, ky][:, :, :, kz]
    np.testing.assert_allclose(Xf, want, rtol=1e-12, atol=1e-10)


def test_identity_mixing_is_a_hermitian_low_pass():
    """With W = I per mode, S is the real low-pass of the zero-padded spectrum:
    kept modes plus their Hermitian mirrors, the kz = 0 plane averaged with its mirror."""
    s = SMALL
    rng = np.random.default_rng(1)
    v = rng.standard_normal((1, s.T, s.T, s.T, s.C))
    spec = np.zeros((s.Kx, s.Ky, s.Kz, s.C, s.C, 2))
    spec[..., 0] = np.eye(s.C)
    S, _ = F._spectral(s, F.tables(s), spec, v)
    kx = F.signed_modes(s.Kx) % s.Np
    ky = F.signed_modes(s.Ky) % s.Np
    mask = np.zeros((s.Np, s.Np, s.Np))
    for a in kx:
        for b in ky:
            for c in range(s.Kz):
                if c == 0:
                    mask[a, b, 0] += 0.5
                    mask[(-a) % s.Np, (-b) % s.Np, 0] += 0.5
                else:
                    mask[a, b, c] += 1.0
                    mask[(-a) % s.Np, (-b) % s.Np, (-c) % s.Np] += 1.0
    sp = np.fft.fftn(v, s=(s.Np, s.Np, s.Np), axes=(1, 2, 3))
    low = np.real(np.fft.ifftn(sp * mask[None, :, :, :, None], axes=(1, 2, 3)))[:, :s.T, :s.T, :s.T]
    np.testing.assert_allclose(S, low, rtol=1e-10, atol=1e-12)


def test_zero_spectral_weights_reduce_to_pointwise_mlp():
    s = SMALL
    rng = np.random.default_rng(2)
    flat = F.init_params(s, seed=3)
    p = F.unflatten(s, flat)
    for l in range(s.L):
        p[f"spec{l}"][:] = 0.0
    flat = F.flatten(s, p)
    x = rng.standard_normal((1, s.T, s.T, s.T, s.Cin))
    y = F.forward(s, flat, x)
    # one voxel at a time, plain products
    for idx in [(0, 0, 0, 0), (0, 3, 5, 7), (0, 7, 7, 7)]:
        h = x[idx] @ p["lift_w"] + p["lift_b"]
        for l in range(s.L):
            pre = h @ p[f"w{l}"] + p[f"b{l}"]
            h = F.gelu(pre) if l < s.L - 1 else pre
        want = F.gelu(h @ p["p1"] + p["c1"]) @ p["p2"] + p["c2"][0]
        assert y[idx] == pytest.approx(want, rel=1e-12, abs=1e-12)


def test_gradients_match_central_differences():
    s = SMALL
    rng = np.random.default_rng(4)
    flat = F.init_params(s, seed=5)
    flat = flat + 0.05 * rng.standard_normal(flat.shape)   # non-zero biases and spectra
    x = rng.standard_normal((2, s.T, s.T, s.T, s.Cin))
    gy = rng.standard_normal((2, s.T, s.T, s.T))
    y, cache = F.forward(s, flat, x, keep=True)
    g = F.backward(s, flat, cache, gy)
    # sample indices from every parameter group
    idx, o = [], 0
    for name, sh in F.param_shapes(s):
        n = int(np.prod(sh))
        idx += list(o + rng.choice(n, size=min(n, 6), replace=False))
        o += n
    h = 1e-6
    for i in idx:
        fp = flat.copy()
        fp[i] += h
        fm = flat.copy()
        fm[i] -= h
        fd = (np.sum(gy * F.forward(s, fp, x)) - np.sum(gy * F.forward(s, fm, x))) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=2e-6, abs=1e-7), i


def test_loss_gradient_matches_differences():
    rng = np.random.default_rng(6)
    y = rng.standard_normal((2, 4, 4, 4))
    t = rng.standard_normal((2, 4, 4, 4))
    w = rng.uniform(0.1, 1.0, (2, 4, 4, 4))
    L, gy = F.loss_and_grad(y, t, w)
    for i in [(0, 0, 0, 0), (1, 2, 3, 1)]:
        yp = y.copy()
        yp[i] += 1e-6
        ym = y.copy()
        ym[i] -= 1e-6
        fd = (F.loss_and_grad(yp, t, w)[0] - F.loss_and_grad(ym, t, w)[0]) / 2e-6
        assert gy[i] == pytest.approx(fd, rel=1e-6)


def test_bf16_rounding_matches_torch():
    import torch
    rng = np.random.default_rng(7)
    a = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-6, 6, 10000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.005859375, 65504.0]])
    want = torch.tensor(a, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(F.bf16(a), want)
