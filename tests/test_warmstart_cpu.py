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


def _gmres_steps(S, c, x0=None, rtol=1e-12, maxit=400):
    """Arnoldi steps of full (unrestarted) GMRES with CGS2 until |c - S x| <= rtol |c|."""
    x = np.zeros_like(c) if x0 is None else x0
    r0 = c - S @ x
    beta, target = np.linalg.norm(r0), rtol * np.linalg.norm(c)
    if beta <= target:
        return 0
    V = [r0 / beta]
    H = np.zeros((maxit + 1, maxit))
    g = np.zeros(maxit + 1)
    g[0] = beta
    cs, sn = np.zeros(maxit), np.zeros(maxit)
    for j in range(maxit):
        w = S @ V[j]
        for _ in range(2):
            d = np.array([q @ w for q in V])
            H[:j + 1, j] += d
            w = w - np.array(V).T @ d
        H[j + 1, j] = np.linalg.norm(w)
        V.append(w / H[j + 1, j])
        for i in range(j):
            t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
            H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
            H[i, j] = t
        rho = np.hypot(H[j, j], H[j + 1, j])
        cs[j], sn[j] = H[j, j] / rho, H[j + 1, j] / rho
        g[j + 1], g[j] = -sn[j] * g[j], cs[j] * g[j]
        if abs(g[j + 1]) <= target:
            return j + 1
    return maxit


def test_guess_accuracy_bounds_the_warm_start_saving(pkg):
    """Why a learned warm start cannot pay off on the capacitance system: on the 440x82 channel's S
    (rank 116, cond ~3e5; pressure_schur.py:92-111) GMRES to 1e-12 needs about as many Arnoldi steps
    from the exact multiplier rounded to BF16 (the resolution of the network's operands) as from
    zero, while the FP32 rounding saves far more: the steps go into resolving the bulk of the
    spectrum to 1e-12, and a guess accurate to ~1e-3 removes only the first few digits."""
    import importlib
    tables = importlib.import_module(pkg.__name__ + '.tables')
    device = importlib.import_module(pkg.__name__ + '.device')
    ws = importlib.import_module(pkg.__name__ + '.warmstart')
    nx, ny, Ly = 440, 82, 4.1
    yc = (np.arange(ny) + 0.5) * (Ly / ny)
    inlet = 4 * 1.5 * yc * (Ly - yc) / Ly ** 2
    ops = tables.GhostOperators(nx, ny, 22., Ly, 1., 2., 2., 0.05, 1., inlet)
    S = device.pressure_pieces(ops)[1].S
    r = S.shape[0]
    # the network-side rounding is the oracle's BF16 rounding
    probe = np.random.default_rng(3).normal(size=257)
    assert np.array_equal(ws.bf16_round(probe), N.bf16(probe.astype(np.float32)).astype(np.float64))
    rng = np.random.default_rng(4)
    t = 2 * np.pi * np.arange(r) / r
    steps = {'zero': [], 'bf16': [], 'fp32': []}
    for _ in range(3):
        y = np.cos(t + rng.uniform(0, 2 * np.pi)) + 0.5 * np.sin(2 * t) + 0.1 * rng.normal(size=r)
        c = S @ y
        steps['zero'].append(_gmres_steps(S, c))
        steps['bf16'].append(_gmres_steps(S, c, ws.bf16_round(y)))
        steps['fp32'].append(_gmres_steps(S, c, y.astype(np.float32).astype(np.float64)))
    mean = {k: float(np.mean(v)) for k, v in steps.items()}
    print('rank', r, 'mean Arnoldi steps', mean)
    assert 40 <= mean['zero'] <= 80
    assert mean['bf16'] >= 0.8 * mean['zero']          # a BF16-accurate guess saves under 20%
    assert mean['fp32'] <= mean['bf16'] - 5            # accuracy, not the guess as such, is what saves
