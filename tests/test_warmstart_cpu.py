"""CPU pins of the warm-start network oracle: BF16 rounding, the periodic convolution against a
direct loop, the adjoint identity, and the gradient against central finite differences."""
import numpy as np
import pytest

from oracle import ghostnet_oracle as N


def test_bf16_round_to_nearest_even():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)                  # BF16 spacing at 1
    cases = [(1.0, 1.0), (1 + 2.0 ** -8, 1.0),     # tie -> even (1.0)
             (1 + 3 * 2.0 ** -8, 1 + 2.0 ** -6),   # tie -> even (1 + 2 ulp)
             (1 + 2.0 ** -8 + 2.0 ** -20, 1 + 2.0 ** -7), (-(1 + 2.0 ** -8 + 2.0 ** -20), -(1 + 2.0 ** -7)),
             (3.140625, 3.140625), (0.0, 0.0)]
    for x, want in cases:
        assert N.bf16(np.float32(x)) == np.float32(want), x
    x = np.random.default_rng(1).normal(size=1000).astype(np.float32)
    r = N.bf16(x)
    assert np.all(abs(r - x) <= abs(x) * float(ulp) / 2 * 1.0000001)
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)
    assert one == N.bf16(one)


def test_periodic_conv_matches_direct_loop():
    rng = np.random.default_rng(2)
    n, L, ci, co, dil = 2, 7, 3, 2, 2
    a = rng.normal(size=(n, L, ci))
    W = rng.normal(size=(co, ci, N.KS))
    out = N.conv(a, W, dil)
    ref = np.zeros((n, L, co))
    for s in range(n):
        for t in range(L):
            for o in range(co):
                for k in range(N.KS):
                    for i in range(ci):
                        ref[s, t, o] += W[o, i, k] * a[s, (t + (k - 2) * dil) % L, i]
    assert np.allclose(out, ref, rtol=1e-13, atol=1e-13)


def test_conv_adjoint_and_weight_gradient_identities():
    rng = np.random.default_rng(3)
    n, L, ci, co, dil = 3, 11, 4, 5, 3
    a = rng.normal(size=(n, L, ci))
    d = rng.normal(size=(n, L, co))
    W = rng.normal(size=(co, ci, N.KS))
    lhs = np.sum(N.conv(a, W, dil) * d)
    assert np.isclose(lhs, np.sum(a * N.conv_adjoint(d, W, dil)), rtol=1e-12)
    assert np.isclose(lhs, np.sum(W * N.conv_weight_grad(d, a, dil)), rtol=1e-12)


def test_exact_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    layers, cin, ch, dil = 3, 4, 5, (1, 2, 1)
    shapes = [(ch, cin), (ch, ch), (1, ch)]
    nparams = sum(co * ci * N.KS + co for co, ci in shapes)
    p = rng.normal(scale=0.4, size=nparams)
    x = rng.normal(size=(2, 9, cin))
    y = rng.normal(size=(2, 9))
    net = N.ConvNet(layers, cin, ch, dil, p, exact=True)
    loss, g = net.loss_grad(x, y)
    eps = 1e-6
    fd = np.empty(nparams)
    for k in range(nparams):
        q = p.copy(); q[k] += eps
        net.set_params(q); lp, _ = net.loss_grad(x, y)
        q[k] -= 2 * eps
        net.set_params(q); lm, _ = net.loss_grad(x, y)
        fd[k] = (lp - lm) / (2 * eps)
    assert np.max(abs(fd - g)) <= 1e-6 * np.max(abs(g)), np.max(abs(fd - g))


def test_exact_residual_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(6)
    layers, cin, ch, dil = 2, 4, 3, (1, 3)
    shapes = [(ch, cin), (1, ch)]
    nparams = sum(co * ci * N.KS + co for co, ci in shapes)
    p = rng.normal(scale=0.4, size=nparams)
    L = 8
    x = rng.normal(size=(2, L, cin))
    S = np.eye(L) + 0.3 * rng.normal(size=(L, L))
    c = x[:, :, 0]
    net = N.ConvNet(layers, cin, ch, dil, p, exact=True)
    loss, g = net.loss_grad(x, c, S)
    out = net.forward(x)
    assert np.isclose(loss, np.sum((out @ S.T - c) ** 2) / out.size, rtol=1e-13)
    eps = 1e-6
    fd = np.empty(nparams)
    for k in range(nparams):
        q = p.copy(); q[k] += eps
        net.set_params(q); lp, _ = net.loss_grad(x, c, S)
        q[k] -= 2 * eps
        net.set_params(q); lm, _ = net.loss_grad(x, c, S)
        fd[k] = (lp - lm) / (2 * eps)
    assert np.max(abs(fd - g)) <= 1e-6 * np.max(abs(g)), np.max(abs(fd - g))


def test_bf16_gradient_close_to_exact():
    rng = np.random.default_rng(5)
    layers, cin, ch, dil = 4, 4, 16, (1, 2, 4, 8)
    shapes = [(ch, cin), (ch, ch), (ch, ch), (1, ch)]
    nparams = sum(co * ci * N.KS + co for co, ci in shapes)
    p = rng.normal(scale=0.2, size=nparams)
    x = rng.normal(size=(3, 40, cin))
    y = rng.normal(size=(3, 40))
    le, ge = N.ConvNet(layers, cin, ch, dil, p, exact=True).loss_grad(x, y)
    lb, gb = N.ConvNet(layers, cin, ch, dil, p.astype(np.float32)).loss_grad(x, y)
    assert abs(lb - le) <= 0.02 * le
    assert np.linalg.norm(gb - ge) <= 0.05 * np.linalg.norm(ge)


@pytest.mark.parametrize('nsteps', [3])
def test_previous_index(pkg, nsteps):
    import importlib
    ws = importlib.import_module(pkg.__name__ + '.warmstart')
    tags = np.array([2, 0, 1, 4, 5, 8, 9])
    assert list(ws.previous_index(tags)) == [-1, -1, -1, 1, 2, 3, 4]
