"""CPU oracle of the warm-start network (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use this module; the
product never imports it.

The network has no counterpart in the reference (the reference solves the capacitance system
directly, solver/pressure_schur.py:92-111): it is the periodic 1D convolution network that warm-starts
GMRES on S y = c.  This restatement mirrors the device arithmetic:

* every layer is out[n, t, o] = b[o] + sum_{k, i} W[o, i, k] a[n, (t + (k - 2) dil) mod L, i];
* BF16 rounding (round to nearest, ties to even) of the layer inputs, of the weights and of the
  back-propagated errors, exactly where the device rounds them, with sums carried in float64
  (the device accumulates in FP32, so the two agree to FP32 round-off plus rare one-ulp BF16 flips);
* loss = mean over positions of (z_out - y)^2, or, given the capacitance matrix S, of the residual
  (S z_out - c)^2 per sample (FP32 matrix products on the device).

``exact=True`` drops the BF16 rounding and runs in float64; in that mode the gradients of both
losses are pinned by central finite differences (tests/test_warmstart_cpu.py), and the convolution
by a direct loop and the adjoint identities.
"""
import numpy as np

KS = 5


def bf16(x):
    """Round float32 values to the nearest BF16 value (ties to even); returns float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def conv(a, W, dil):
    """out[n, t, o] = sum_k sum_i W[o, i, k] a[n, (t + (k-2) dil) mod L, i]  (float64)."""
    out = 0.
    for k in range(KS):
        shifted = np.roll(a, -(k - 2) * dil, axis=1)
        out = out + shifted.astype(np.float64) @ W[:, :, k].T.astype(np.float64)
    return out


def conv_adjoint(d, W, dil):
    """da[n, s, i] = sum_k sum_o W[o, i, k] d[n, (s - (k-2) dil) mod L, o]  (float64)."""
    out = 0.
    for k in range(KS):
        shifted = np.roll(d, (k - 2) * dil, axis=1)
        out = out + shifted.astype(np.float64) @ W[:, :, k].astype(np.float64)
    return out


def conv_weight_grad(d, a, dil):
    """dW[o, i, k] = sum_{n, t} d[n, t, o] a[n, (t + (k-2) dil) mod L, i]  (float64)."""
    co, ci = d.shape[2], a.shape[2]
    g = np.zeros((co, ci, KS))
    d2 = d.reshape(-1, co).astype(np.float64)
    for k in range(KS):
        shifted = np.roll(a, -(k - 2) * dil, axis=1).reshape(-1, ci).astype(np.float64)
        g[:, :, k] = d2.T @ shifted
    return g


class ConvNet:
    def __init__(self, layers, cin, channels, dilations, params, exact=False):
        self.layers, self.cin, self.ch = layers, cin, channels
        self.dil = list(dilations)
        self.exact = exact
        self.shapes = []
        for l in range(layers):
            ci = cin if l == 0 else channels
            co = 1 if l == layers - 1 else channels
            self.shapes.append((co, ci))
        self.nparams = sum(co * ci * KS + co for co, ci in self.shapes)
        self.set_params(params)

    def set_params(self, params):
        params = np.asarray(params, dtype=np.float64 if self.exact else np.float32)
        assert params.size == self.nparams
        self.W, self.b = [], []
        off = 0
        for co, ci in self.shapes:
            self.W.append(params[off:off + co * ci * KS].reshape(co, ci, KS))
            off += co * ci * KS
            self.b.append(params[off:off + co])
            off += co
        self.params = params

    def _r(self, x):
        return x if self.exact else bf16(x)

    def _f(self, x):
        return np.asarray(x, dtype=np.float64 if self.exact else np.float32)

    def forward(self, x):
        """x: (n, L, cin) -> output (n, L); keeps the layer inputs and pre-activations."""
        a = self._r(self._f(x))
        self.acts, self.zs = [a], []
        for l in range(self.layers):
            z = self._f(conv(a, self._r(self.W[l]), self.dil[l]) + self.b[l])
            self.zs.append(z)
            if l < self.layers - 1:
                a = self._r(np.maximum(z, 0))
                self.acts.append(a)
        return self.zs[-1][:, :, 0]

    def loss_grad(self, x, y, S=None):
        """Loss and flat gradient.  Without S: mean (out - y)^2.  With S (L x L): mean over positions
        of (S out - y)^2 per sample, where y is then the scaled right-hand side."""
        out = self.forward(x)
        y = self._f(y)
        P = out.size
        if S is None:
            d = self._f(out - y)
            loss = float(np.sum(d.astype(np.float64) ** 2) / P)
            dz = self._r(self._f(2 * d / P))[:, :, None]
        else:
            S = self._f(S)
            q = self._f(out.astype(np.float64) @ S.T.astype(np.float64) - y)
            loss = float(np.sum(q.astype(np.float64) ** 2) / P)
            g = self._f(q.astype(np.float64) @ S.astype(np.float64))
            dz = self._r(self._f(self._f(2.0 / P) * g))[:, :, None]
        grads = [None] * self.layers
        for l in range(self.layers - 1, -1, -1):
            gW = self._f(conv_weight_grad(dz, self.acts[l], self.dil[l]))
            gb = self._f(np.sum(dz.astype(np.float64), axis=(0, 1)))
            grads[l] = (gW, gb)
            if l > 0:
                da = self._f(conv_adjoint(dz, self._r(self.W[l]), self.dil[l]))
                dz = self._r(self._f(da * (self.zs[l - 1] > 0)))
        flat = np.concatenate([np.concatenate((gW.ravel(), gb.ravel())) for gW, gb in grads])
        return loss, flat
