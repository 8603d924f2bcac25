"""Device surrogate network against the FP64 oracle (oracle/fno_oracle.py).

The oracle rounds to bfloat16 where the device does (inputs and weights of the
per-voxel channel mixes, stored activations, the staged gradients of the
weight products), so the remaining differences are FP32 accumulation order and
rare one-ulp bf16 rounding flips: tolerances below are 1e-3 relative in L2."""

import numpy as np
import pytest

from oracle import fno_oracle as F

pytestmark = pytest.mark.gpu

SPEC = F.Spec()


def _inputs(seed, nb=1):
    """Smooth layered-looking fields plus noise, so both the spectral and the
    pointwise paths carry signal."""
    rng = np.random.default_rng(seed)
    g = np.linspace(0.0, 1.0, 64)
    X, Y, Z = np.meshgrid(g, g, g, indexing='ij')
    x = np.zeros((nb, 64, 64, 64, 8), np.float32)
    for b in range(nb):
        for c in range(8):
            a1, a2, a3 = rng.uniform(0.5, 3.0, 3)
            x[b, ..., c] = (np.sin(2 * np.pi * a1 * X + c) * np.cos(2 * np.pi * a2 * Y)
                            + np.sin(2 * np.pi * a3 * Z) + 0.3 * rng.standard_normal((64, 64, 64)))
    return x


def _params(seed):
    flat = F.init_params(SPEC, seed=seed)
    p = F.unflatten(SPEC, flat)
    rng = np.random.default_rng(seed + 100)
    for l in range(SPEC.L):
        p[f"spec{l}"] *= 10.0          # spectral path comparable to the pointwise path
        p[f"b{l}"] = 0.1 * rng.standard_normal(SPEC.C)
    p["c1"] = 0.1 * rng.standard_normal(SPEC.H)
    p["c2"] = np.array([0.3])
    return F.flatten(SPEC, p).astype(np.float32).astype(np.float64)


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-30))


@pytest.fixture(scope="module")
def net():
    from surrogate.network import FourierSurrogate
    n = FourierSurrogate(modes=(SPEC.Kx, SPEC.Ky, SPEC.Kz), np_len=SPEC.Np, max_batch=2, train=True)
    assert n.n_params == F.n_params(SPEC)
    yield n
    n.close()


def _dev(a):
    from simulation._bdlib import DeviceArray
    return DeviceArray(a.shape).upload(a)


def test_forward_matches_oracle(net):
    flat = _params(1)
    x = _inputs(2)
    net.set_params(flat)
    y = net.forward(_dev(x)).download()
    yo = F.forward(SPEC, flat, x.astype(np.float64), emulate_bf16=True)
    err = _rel(y, yo)
    print(f"forward rel L2 {err:.2e}, max abs {np.abs(y - yo).max():.2e}, |y| rms {np.sqrt((yo ** 2).mean()):.3f}")
    assert err < 1e-3
    # and the unrounded FP64 network is within bf16 distance
    yf = F.forward(SPEC, flat, x.astype(np.float64))
    assert _rel(y, yf) < 3e-2


def test_gradients_match_oracle(net):
    flat = _params(3)
    x = _inputs(4)
    rng = np.random.default_rng(5)
    gy = rng.standard_normal((1, 64, 64, 64)).astype(np.float32)
    net.set_params(flat)
    xd = _dev(x)
    net.forward(xd)
    net.backward(xd, _dev(gy))
    g = net.get_grads().astype(np.float64)
    yo, cache = F.forward(SPEC, flat, x.astype(np.float64), keep=True, emulate_bf16=True)
    go = F.backward(SPEC, flat, cache, gy.astype(np.float64))
    o = 0
    for name, sh in F.param_shapes(SPEC):
        n = int(np.prod(sh))
        e = _rel(g[o:o + n], go[o:o + n])
        print(f"{name:8s} rel L2 {e:.2e}")
        assert e < 1e-3, name
        o += n


def test_backward_is_bit_reproducible(net):
    flat = _params(6)
    x = _dev(_inputs(7, nb=2))
    gy = _dev(np.random.default_rng(8).standard_normal((2, 64, 64, 64)).astype(np.float32))
    net.set_params(flat)
    out = []
    for _ in range(2):
        net.forward(x)
        net.backward(x, gy)
        out.append(net.get_grads())
    np.testing.assert_array_equal(out[0], out[1])


def test_loss_kernel_matches_numpy(net):
    rng = np.random.default_rng(9)
    y = rng.standard_normal((2, 64, 64, 64)).astype(np.float32)
    t = rng.standard_normal((2, 64, 64, 64)).astype(np.float32)
    w = rng.uniform(0.0, 1.0, (2, 64, 64, 64)).astype(np.float32)
    w /= w.reshape(2, -1).sum(1)[:, None, None, None]
    L, gy = net.loss(*(_dev(a) for a in (y, t, w)))
    d = y.astype(np.float64) - t
    assert L == pytest.approx(float((w * d * d).sum() / 2), rel=1e-6)
    np.testing.assert_allclose(gy.download(), 2 * w * (y - t) / 2, rtol=1e-6, atol=1e-12)


def test_training_fits_a_fixed_target(net):
    """AdamW on the device reduces the loss of a fixed two-tile regression by 10x
    (target: a smooth function of the inputs)."""
    x = _inputs(10, nb=2)
    tgt = (0.5 * np.tanh(x[..., 0]) + 0.3 * x[..., 4] - 0.2).astype(np.float32)
    w = np.full((2, 64, 64, 64), 1.0 / 64 ** 3, np.float32)
    net.set_params(F.init_params(SPEC, seed=11).astype(np.float32))
    xd, td, wd = (_dev(a) for a in (x, tgt, w))
    losses = [net.train_step(xd, td, wd, lr=3e-3) for _ in range(60)]
    print("loss", losses[0], "->", losses[-1])
    assert losses[-1] < 0.1 * losses[0]
