"""Host handle of the device Fourier neural operator (include/bugdef_fno.h).

Tensors are library-owned device buffers (simulation._bdlib.DeviceArray): the
network, its data and the engine share one GPU runtime, the library's.  Every
call checks shapes; uploads are synchronous, and calls that hand results back
to the host wait for the network's stream.
"""

import ctypes

import numpy as np

from simulation import _bdlib as B

T = 64      # tile edge
CIN = 8     # input channels
C = 32      # width
L = 4       # Fourier layers
H = 64      # projection width

DEFAULT_MODES = (6, 6, 8)
DEFAULT_NP = 80


def param_layout(modes=DEFAULT_MODES):
    """Ordered (name, shape) of the flat parameter vector (include/bugdef_fno.h)."""
    kx, ky, kz = modes
    out = [("lift_w", (CIN, C)), ("lift_b", (C,))]
    for l in range(L):
        out += [(f"spec{l}", (kx, ky, kz, C, C, 2)), (f"w{l}", (C, C)), (f"b{l}", (C,))]
    out += [("p1", (C, H)), ("c1", (H,)), ("p2", (H,)), ("c2", (1,))]
    return out


def init_params(modes=DEFAULT_MODES, seed=0, spectral_scale=2.0):
    """Fan-in scaled normal weights, zero biases; spectral weights
    spectral_scale / (2 C) per real component."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, sh in param_layout(modes):
        if name.startswith("spec"):
            parts.append(rng.standard_normal(sh) * spectral_scale / (2.0 * C))
        elif name in ("lift_w", "p1", "p2") or (name[0] == "w" and name[1:].isdigit()):
            parts.append(rng.standard_normal(sh) / np.sqrt(sh[0]))
        else:
            parts.append(np.zeros(sh))
    return np.concatenate([p.ravel() for p in parts]).astype(np.float32)


def _dev_ptr(t, shape, what):
    if not (isinstance(t, B.DeviceArray) and t.dtype == np.float32):
        raise ValueError(f"{what}: expected a float32 DeviceArray")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{what}: expected shape {tuple(shape)}, got {tuple(t.shape)}")
    return t.ptr


class FourierSurrogate:
    """A network context on one device: parameters, optional training state."""

    def __init__(self, modes=DEFAULT_MODES, np_len=DEFAULT_NP, max_batch=8, train=False, device=0):
        self.modes = tuple(int(m) for m in modes)
        self.np_len = int(np_len)
        self.max_batch = int(max_batch)
        self.train = bool(train)
        self.device = int(device)
        self.mc_photons = 0   # > 0: trained on features whose channel 5 is a run of this many photons
        spec = B.FnoSpec(self.modes[0], self.modes[1], self.modes[2], self.np_len)
        h = ctypes.c_void_p()
        B.check(B.lib().bd_fno_create(self.device, ctypes.byref(spec), self.max_batch, 1 if train else 0,
                                      ctypes.byref(h)), "bd_fno_create")
        self._h = h
        n = ctypes.c_int64(0)
        B.check(B.lib().bd_fno_num_params(h, ctypes.byref(n)), "bd_fno_num_params")
        self.n_params = int(n.value)

    # -- parameters -----------------------------------------------------------
    def save(self, path, **meta):
        np.savez(path, params=self.get_params(), modes=np.array(self.modes), np_len=self.np_len,
                 mc_photons=self.mc_photons, **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path, max_batch=8, train=False, device=0):
        with np.load(path, allow_pickle=False) as z:
            net = cls(modes=tuple(int(v) for v in z["modes"]), np_len=int(z["np_len"]), max_batch=max_batch,
                      train=train, device=device)
            net.set_params(z["params"])
            net.mc_photons = int(z["mc_photons"]) if "mc_photons" in z.files else 0
        return net

    def set_params(self, flat):
        a = np.ascontiguousarray(flat, dtype=np.float32)
        if a.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {a.size}")
        B.check(B.lib().bd_fno_set_params(self._h, B._ptr(a), a.size), "bd_fno_set_params")

    def get_params(self):
        a = np.empty(self.n_params, np.float32)
        B.check(B.lib().bd_fno_get_params(self._h, B._ptr(a), a.size), "bd_fno_get_params")
        return a

    def get_grads(self):
        a = np.empty(self.n_params, np.float32)
        B.check(B.lib().bd_fno_get_grads(self._h, B._ptr(a), a.size), "bd_fno_get_grads")
        return a

    # -- compute ----------------------------------------------------------------
    def forward(self, x, y=None, wait=True):
        """x: float32 DeviceArray [B,64,64,64,8] -> y: float32 DeviceArray [B,64,64,64]."""
        nb = int(x.shape[0])
        if y is None:
            y = B.DeviceArray((nb, T, T, T), device=self.device)
        px = _dev_ptr(x, (nb, T, T, T, CIN), "forward x")
        py = _dev_ptr(y, (nb, T, T, T), "forward y")
        B.check(B.lib().bd_fno_forward(self._h, px, nb, py), "bd_fno_forward")
        if wait:
            self.sync()
        return y

    def loss(self, y, target, weight, gy=None):
        nb = int(y.shape[0])
        shp = (nb, T, T, T)
        if gy is None:
            gy = B.DeviceArray(shp, device=self.device)
        out = ctypes.c_double(0.0)
        args = [_dev_ptr(y, shp, "loss y"), _dev_ptr(target, shp, "loss target"), _dev_ptr(weight, shp, "loss w")]
        B.check(B.lib().bd_fno_loss(self._h, *args, nb, _dev_ptr(gy, shp, "loss gy"), ctypes.byref(out)),
                "bd_fno_loss")
        return float(out.value), gy

    def backward(self, x, gy):
        nb = int(x.shape[0])
        px = _dev_ptr(x, (nb, T, T, T, CIN), "backward x")
        pg = _dev_ptr(gy, (nb, T, T, T), "backward gy")
        B.check(B.lib().bd_fno_backward(self._h, px, pg, nb), "bd_fno_backward")

    def adamw(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=1e-4):
        B.check(B.lib().bd_fno_adamw(self._h, lr, beta1, beta2, eps, weight_decay), "bd_fno_adamw")

    def train_step(self, x, target, weight, lr, y=None, gy=None, **adam):
        y = self.forward(x, y, wait=False)
        loss, gy = self.loss(y, target, weight, gy)
        self.backward(x, gy)
        self.adamw(lr, **adam)
        return loss

    def sync(self):
        B.check(B.lib().bd_fno_sync(self._h), "bd_fno_sync")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            B.lib().bd_fno_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
