"""Learned warm start for GMRES on the capacitance (Schur) system of the pressure solve.

The pressure solve applies S^{-1} (S = I + C R^{-1} E, pressure_schur.py:92-111) directly; this
module adds the iterative alternative and its learned initial guess:

* ``dump_pairs`` records the (c, y = S^{-1} c) pairs of the solver's own first pressure solves
  (``mcf_set_dump``), in boundary-angle order;
* ``WarmStartNet`` is the periodic 1D convolution network over the boundary angle
  (``mcf_net_*``: BF16 matrix-core forward / backward, FP32 accumulation and Adam), trained on those
  pairs and asked for y / |c| given c / |c| and per-row geometry features;
* ``gmres_study`` runs device GMRES to a relative residual of 1e-12 from the zero guess, the
  previous step's multiplier, the network's guess and the previous multiplier corrected by the
  network, and reports Arnoldi steps and the error against the direct FP64 solve.
"""
import ctypes

import numpy as np

from . import _native as nat

DEFAULT_DILATIONS = (1, 2, 4, 8)


def dump_pairs(flow, nsteps, dt, capacity=None):
    """Run nsteps on the device recording every projection's first pressure solve.
    Returns (c, y, tags) with tags = 4 * step + projection (step counted from this call)."""
    dev = flow.gg.device
    cap = capacity or 2 * nsteps + 4
    dev.set_dump(cap)
    try:
        flow._push()
        diag = dev.run(nsteps, dt)
        flow._pull()
        flow.time = float(diag[-1, 0])
        c, y, tags = dev.get_dump()
    finally:
        dev.set_dump(0)
    return c, y, tags


def net_inputs(c, features):
    """(n, L, 1 + nf) float32: c scaled to unit 2-norm per sample, then the features."""
    c = np.atleast_2d(c)
    s = np.linalg.norm(c, axis=1, keepdims=True)
    s[s == 0] = 1.
    x = np.empty((c.shape[0], c.shape[1], 1 + features.shape[1]), dtype=np.float32)
    x[:, :, 0] = c / s
    x[:, :, 1:] = features[None].astype(np.float32)
    return x, s


class WarmStartNet:
    """Owner of one ``mcf_net`` (periodic conv net, 5 taps, ReLU, BF16 matrix cores)."""

    def __init__(self, layers=4, channels=64, dilations=DEFAULT_DILATIONS, cin=4, seed=5000):
        self.lib = nat.library()
        self.layers, self.channels, self.cin = layers, channels, cin
        self.dilations = np.ascontiguousarray(dilations[:layers], dtype=np.int32)
        h = ctypes.c_void_p()
        rc = self.lib.mcf_net_create(layers, cin, channels, nat.iptr(self.dilations), seed, ctypes.byref(h))
        if rc:
            msg = self.lib.mcf_net_last_error(h).decode() if h.value else 'mcf_net_create failed'
            if h.value:
                self.lib.mcf_net_destroy(h)
            raise nat.STATUS_EXCEPTIONS.get(rc, RuntimeError)(msg)
        self.h = h
        self.nparams = self.lib.mcf_net_num_params(h)

    def _check(self, rc):
        if rc:
            raise nat.STATUS_EXCEPTIONS.get(rc, RuntimeError)(self.lib.mcf_net_last_error(self.h).decode())

    @property
    def params(self):
        out = np.empty(self.nparams, dtype=np.float32)
        self._check(self.lib.mcf_net_get_params(self.h, nat.as_arg(out)))
        return out

    @params.setter
    def params(self, value):
        value = nat.f32(value)
        if value.size != self.nparams:
            raise ValueError(f'expected {self.nparams} parameters')
        self._check(self.lib.mcf_net_set_params(self.h, nat.as_arg(value)))

    def clear(self):
        self._check(self.lib.mcf_net_clear_groups(self.h))

    def add_pairs(self, c, y, features, S=None):
        """One training group: pairs of one boundary (angle order), scaled by |c| per sample.  With S
        (angle order) the group trains the residual |S yhat - c|^2, otherwise the error |yhat - y|^2."""
        x, s = net_inputs(c, features)
        cn = nat.f32(x[:, :, 0])
        yn = None if S is not None else nat.f32(np.atleast_2d(y) / s)
        self._check(self.lib.mcf_net_add_group(self.h, cn.shape[0], cn.shape[1], nat.as_arg(cn),
                                               None if yn is None else nat.as_arg(yn), nat.as_arg(nat.f32(features)),
                                               None if S is None else nat.as_arg(nat.f32(S))))

    def train(self, steps, positions=16384, lr=1e-3, seed=5000):
        losses = np.empty(steps, dtype=np.float32)
        self._check(self.lib.mcf_net_train(self.h, int(steps), int(positions), float(lr), int(seed),
                                           nat.as_arg(losses)))
        return losses

    def forward(self, x):
        x = nat.f32(x)
        n, L = x.shape[0], x.shape[1]
        y = np.empty((n, L), dtype=np.float32)
        self._check(self.lib.mcf_net_predict(self.h, n, L, nat.as_arg(x), nat.as_arg(y)))
        return y

    def predict(self, c, features):
        """Guess of S^{-1} c for the rows of c (angle order)."""
        x, s = net_inputs(c, features)
        return self.forward(x).astype(np.float64) * s

    def loss_grad(self, x, y, S=None):
        x, y = nat.f32(x), nat.f32(y)
        loss = np.zeros(1, dtype=np.float32)
        g = np.empty(self.nparams, dtype=np.float32)
        self._check(self.lib.mcf_net_loss_grad(self.h, x.shape[0], x.shape[1], nat.as_arg(x), nat.as_arg(y),
                                               None if S is None else nat.as_arg(nat.f32(S)), nat.as_arg(loss),
                                               nat.as_arg(g)))
        return float(loss[0]), g

    def profile(self, nsamp, L, reps=50):
        ms = np.zeros(1, dtype=np.float32)
        self._check(self.lib.mcf_net_profile(self.h, nsamp, L, reps, nat.as_arg(ms)))
        return float(ms[0])

    def close(self):
        if getattr(self, 'h', None) is not None and self.h.value:
            self.lib.mcf_net_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def previous_index(tags):
    """Index of the same projection's pair one step earlier, or -1."""
    pos = {int(t): k for k, t in enumerate(tags)}
    return np.array([pos.get(int(t) - 4, -1) for t in tags])


def min_residual_guess(S, c, basis):
    """x0 = sum_j a_j basis_j minimising |c - S x0| (per system; basis: list of (n, r) arrays).
    Including the previous multiplier in the basis means the start is never worse than it."""
    B = np.stack(basis, axis=2)                                  # (n, r, m)
    SB = np.einsum('ij,njm->nim', S, B)
    x0 = np.empty_like(c)
    for k in range(len(c)):
        a, *_ = np.linalg.lstsq(SB[k], c[k], rcond=None)
        x0[k] = B[k] @ a
    return x0


def bf16_round(x):
    """x rounded to the nearest BF16 value (ties to even), returned as float64."""
    b = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = ((b + 0x7fff + ((b >> 16) & 1)) >> 16) << 16
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def gmres_study(dev, c, y, tags, S, net=None, rtol=1e-12, max_iter=400, ceilings=False):
    """Arnoldi steps of device GMRES (to |c - S y| <= rtol |c|) and the error against the direct
    FP64 multiplier y, per initial guess: zero; the previous step's multiplier; the network's guess
    (raw, and scaled to minimum residual); and the previous multiplier combined with the network's
    correction of its residual (raw sum, and the minimum-residual combination).  With ``ceilings``
    also the direct multiplier itself rounded to BF16 (at most 2^-8 relative per entry, the resolution of the
    network's BF16 operands) and to FP32 (its output format): they bound what a guess at that
    accuracy, however it is obtained, can save.  S is the
    Schur matrix in angle order.  Pairs without a previous step are skipped for every guess, so all
    guesses see the same systems."""
    prev = previous_index(tags)
    keep = np.flatnonzero(prev >= 0)
    c, y, yp = c[keep], y[keep], y[prev[keep]]
    feats = dev.schur_geometry['features']
    guesses = {'zero': None, 'previous': yp}
    if ceilings:
        guesses['direct multiplier in BF16 (ceiling)'] = bf16_round(y)
        guesses['direct multiplier in FP32'] = y.astype(np.float32).astype(np.float64)
    if net is not None:
        g = net.predict(c, feats)
        guesses['network'] = g
        guesses['network, min-residual scaled'] = min_residual_guess(S, c, [g])
        r0 = c - yp @ S.T
        corr = net.predict(r0, feats)
        guesses['previous+network'] = yp + corr
        guesses['previous+network, min-residual'] = min_residual_guess(S, c, [yp, corr])
    out = {}
    ynorm = np.linalg.norm(y, axis=1)
    cnorm = np.linalg.norm(c, axis=1)
    for name, x0 in guesses.items():
        init_res = np.ones(len(c)) if x0 is None else np.linalg.norm(c - x0 @ S.T, axis=1) / cnorm
        yg, its, res = dev.schur_gmres(c, x0, rtol=rtol, max_iter=max_iter)
        err = np.linalg.norm(yg - y, axis=1) / ynorm
        out[name] = dict(systems=int(len(c)), mean_iterations=float(its.mean()), max_iterations=int(its.max()),
                         max_relres=float(res.max()), max_error_vs_direct=float(err.max()),
                         median_initial_residual=float(np.median(init_res)))
    return out


TRAINING_CASES = (
    ('440x82 Re 20 zero start', dict(nx=440, ny=82, Lx=22., Ly=4.1, xc=2., yc=2., nu=0.05, inflow='parabolic',
                                     U_m=1.5, initial='zero'), 2e-3),
    ('1320x246 Re 100 perturbed', dict(nx=1320, ny=246, Lx=22., Ly=4.1, xc=2., yc=2., nu=0.01, inflow='parabolic',
                                       U_m=1.5, initial='profile', v_perturbation=dict(seed=5000)), 5e-4),
    ('800x820 uniform Re 100', dict(nx=800, ny=820, Lx=40., Ly=41., xc=10., yc=20.5, nu=0.01, inflow='uniform',
                                    U_inf=1.0, initial='profile'), 2e-3),
)


def angle_schur(dev):
    """The pressure Schur matrix in angle order (host copy, kept with the boundary geometry)."""
    return dev.schur_geometry['S']


def step_differences(c, y, tags):
    """(c_t - c_{t-1}, y_t - y_{t-1}) of the same projection: the system the previous-step guess leaves."""
    prev = previous_index(tags)
    k = np.flatnonzero(prev >= 0)
    return c[k] - c[prev[k]], y[k] - y[prev[k]]


def study(steps=(2000, 4000, 2000), train_steps=3000, transient=200, dump_steps=60, seed=5000, log=None):
    """Warm-start study.  Dump (c, y) pairs from the three channel / box cases, train the network on
    the device (pairs and step differences, error loss), then compare GMRES to 1e-12 from each
    initial guess on held-out flows: the 440x82 and 1320x246 channels at Re ~ U[200,250] (training
    geometry, unseen Reynolds number) and two new channels (Re ~ U[200,250], h = D/30..D/60, offset
    cylinder, zero start, seeds 9000+k).  An untrained network of the same architecture and
    initialisation is the control.  The training-geometry flows run `transient` steps before
    `dump_steps` steps are recorded; a held-out flow that fails a gate of the reference step is
    listed in `failed`."""
    import time
    from .ensemble import channel_case, solver_kwargs
    from .flow import MAC_CylinderFlow

    def say(msg):
        if log:
            log(msg)

    def flow(**kw):
        return MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, **kw)

    t0 = time.perf_counter()
    net = WarmStartNet(seed=seed)
    control = WarmStartNet(seed=seed)
    pairs = 0
    for (name, kw, dt), n in zip(TRAINING_CASES, steps):
        f = flow(**kw)
        c, y, tags = dump_pairs(f, n, dt)
        fe = f.gg.device.schur_geometry['features']
        net.add_pairs(c, y, fe)
        dc, dy = step_differences(c, y, tags)
        net.add_pairs(dc, dy, fe)
        pairs += len(c) + len(dc)
        say(f'{name}: {len(c)} pairs, r = {c.shape[1]}')
        f.gg.device.h.close()
    t1 = time.perf_counter()
    losses = net.train(train_steps, positions=16384, lr=1e-3, seed=seed)
    t2 = time.perf_counter()
    say(f'trained {train_steps} steps in {t2 - t1:.2f} s; loss {losses[:5].mean():.4g} -> {losses[-50:].mean():.4g}')
    rng = np.random.default_rng(seed)
    heldout = []
    for (name, kw, dt) in TRAINING_CASES[:2]:
        re = rng.uniform(200, 250)
        heldout.append(('training geometry', f'{name.split()[0]} Re {re:.0f}', dict(kw, nu=1.0 / re), dt))
    for k in range(2):
        r2 = np.random.default_rng(9000 + k)
        case = channel_case(int(r2.integers(30, 61)), r2.uniform(0.15, 0.3), r2.uniform(-0.1, 0.1), r2.uniform(200, 250))
        heldout.append(('new geometry', f"{case['nx']}x{case['ny']} Re {case['Re']:.0f}",
                        dict(initial='zero', **solver_kwargs(case)), case['dt']))
    cases, totals, failed = [], {}, []
    for group, label, kw, dt in heldout:
        f = flow(**kw)
        try:
            # the new channels start from rest: they are recorded from the first step (their coarse,
            # high-Re members can fail the reference's CFL gate within a few hundred steps)
            if transient and group == 'training geometry':
                f.solve(n_steps=transient, dt=dt, perturb_transient=False, save_artifacts=False, verbose=False)
            c, y, tags = dump_pairs(f, dump_steps, dt)
        except RuntimeError as e:
            failed.append(dict(case=label, group=group, error=str(e)))
            say(f'held-out {label}: {e}')
            f.gg.device.h.close()
            continue
        dev = f.gg.device
        S = angle_schur(dev)
        st = gmres_study(dev, c, y, tags, S, net, ceilings=True)
        ctl = gmres_study(dev, c, y, tags, S, control)
        for g in ('network', 'network, min-residual scaled', 'previous+network, min-residual'):
            st[g.replace('network', 'untrained network', 1)] = ctl[g]
        cases.append(dict(group=group, case=label, rank=int(c.shape[1]), guesses=st))
        for g, v in st.items():
            tot = totals.setdefault(group, {}).setdefault(
                g, dict(systems=0, iterations=0., initial=[], max_relres=0., max_error=0.))
            tot['systems'] += v['systems']
            tot['iterations'] += v['mean_iterations'] * v['systems']
            tot['initial'].append(v['median_initial_residual'])
            tot['max_relres'] = max(tot['max_relres'], v['max_relres'])
            tot['max_error'] = max(tot['max_error'], v['max_error_vs_direct'])
        say(f'held-out {label} ({group}): ' + ', '.join(f"{g} {v['mean_iterations']:.1f}" for g, v in st.items()))
        dev.h.close()
    summary = {grp: {g: dict(systems=v['systems'], mean_iterations=v['iterations'] / v['systems'],
                             median_initial_residual=float(np.median(v['initial'])), max_relres=v['max_relres'],
                             max_error_vs_direct=v['max_error']) for g, v in gs.items()} for grp, gs in totals.items()}
    out = dict(training_pairs=pairs, training_steps=train_steps, training_seconds=t2 - t1,
               loss_first=float(losses[:5].mean()), loss_last=float(losses[-50:].mean()),
               conv_ms_per_launch=net.profile(64, 340), seconds=time.perf_counter() - t0, heldout=cases,
               failed=failed, summary=summary)
    net.close()
    control.close()
    return out


def _accumulate(totals, group, st):
    """Fold one case's per-guess GMRES figures into the group totals."""
    for g, v in st.items():
        tot = totals.setdefault(group, {}).setdefault(
            g, dict(systems=0, iterations=0., max_iterations=0, initial=[], max_relres=0., max_error=0.))
        tot['systems'] += v['systems']
        tot['iterations'] += v['mean_iterations'] * v['systems']
        tot['max_iterations'] = max(tot['max_iterations'], v['max_iterations'])
        tot['initial'].append(v['median_initial_residual'])
        tot['max_relres'] = max(tot['max_relres'], v['max_relres'])
        tot['max_error'] = max(tot['max_error'], v['max_error_vs_direct'])


def _summary(totals):
    return {grp: {g: dict(systems=v['systems'], mean_iterations=v['iterations'] / v['systems'],
                          max_iterations=v['max_iterations'], median_initial_residual=float(np.median(v['initial'])),
                          max_relres=v['max_relres'], max_error_vs_direct=v['max_error'])
                  for g, v in gs.items()} for grp, gs in totals.items()}


def config5_study(c1_steps=2000, c2_seeds=(5000, 5001, 5002, 5003), c2_steps=22000, c3_steps=10000, ncases=64,
                  re_range=(200., 250.), transient=20, dump_steps=30, train_steps=4000, seed=5000, log=None):
    """The learned warm start on its benchmark workload.  Training pairs (about 200k with the default
    counts) are dumped from the 440x82 Re 20 channel (zero start, dt 2e-3, c1_steps), the 1320x246
    Re 100 channel started from the inlet profile with one seeded v-perturbation per seed in c2_seeds
    (dt 5e-4, c2_steps each) and the 800x820 uniform-inflow box (its 1250 perturbed start-up steps
    run first, then c3_steps are dumped, dt 2e-3); every group also trains on its step differences.
    The evaluation systems are those of the 64 seeded ensemble geometries (``ensemble_case(k,
    re_range)``: the ensemble's geometries, none trained on, with Reynolds numbers drawn from
    U[200, 250]) from the zero start: `transient` steps run, then `dump_steps` steps are recorded,
    and every system is solved by device GMRES to 1e-12 from each initial guess, the trained network
    and an untrained one of the same initialisation (the control), with the BF16 / FP32 roundings
    of the direct multiplier as ceilings.  Members that fail a gate of the reference step before
    their systems are recorded are listed in `failed`."""
    import time
    from .ensemble import build_flows, ensemble_case
    from .flow import MAC_CylinderFlow

    def say(msg):
        if log:
            log(msg)

    def flow(**kw):
        return MAC_CylinderFlow(boundary='ghost_cell', time_scheme='ab2_cn', verbose=False, **kw)

    t0 = time.perf_counter()
    net = WarmStartNet(seed=seed)
    control = WarmStartNet(seed=seed)
    c1, c2, c3 = (kw for _, kw, _ in TRAINING_CASES)
    runs = [('440x82 Re 20', c1, TRAINING_CASES[0][2], c1_steps, 0)]
    runs += [(f'1320x246 Re 100 seed {s}', dict(c2, v_perturbation=dict(seed=s)), TRAINING_CASES[1][2], c2_steps, 0)
             for s in c2_seeds]
    runs.append(('800x820 box Re 100', c3, TRAINING_CASES[2][2], c3_steps, 1250))
    pairs, groups = 0, []
    for name, kw, dt, n, lead in (run for run in runs if run[3] > 0):
        f = flow(**kw)
        if lead:
            f.solve(n_steps=lead, dt=dt, perturb_transient=True, save_artifacts=False, verbose=False)
        c, y, tags = dump_pairs(f, n, dt)
        fe = f.gg.device.schur_geometry['features']
        net.add_pairs(c, y, fe)
        net.add_pairs(*step_differences(c, y, tags), fe)
        pairs += len(c)
        groups.append(dict(case=name, pairs=int(len(c)), rank=int(c.shape[1]), steps=n))
        say(f'{name}: {len(c)} pairs, r = {c.shape[1]} ({time.perf_counter() - t0:.0f} s)')
        del c, y, tags
        f.gg.device.h.close()
    t1 = time.perf_counter()
    losses = net.train(train_steps, positions=16384, lr=1e-3, seed=seed)
    t2 = time.perf_counter()
    say(f'trained {train_steps} steps on {pairs} pairs in {t2 - t1:.2f} s; '
        f'loss {losses[:5].mean():.4g} -> {losses[-50:].mean():.4g}')
    cases = [ensemble_case(k, re_range=re_range) for k in range(ncases)]
    starts = {}
    flows = build_flows(cases, failures=starts)
    failed = [dict(case=k, error=v['message']) for k, v in starts.items()]
    group = f'ensemble geometries, Re U[{re_range[0]:g},{re_range[1]:g}]'
    totals, per_case = {}, []
    for k, f in enumerate(flows):
        if f is None:
            continue
        case = cases[k]
        dev = f.gg.device
        try:
            if transient:
                f.solve(n_steps=transient, dt=case['dt'], perturb_transient=False, save_artifacts=False, verbose=False)
            c, y, tags = dump_pairs(f, dump_steps, case['dt'])
        except RuntimeError as e:
            failed.append(dict(case=k, error=str(e)))
            dev.h.close()
            continue
        S = angle_schur(dev)
        st = gmres_study(dev, c, y, tags, S, net, ceilings=True)
        ctl = gmres_study(dev, c, y, tags, S, control)
        for g in ('network', 'network, min-residual scaled', 'previous+network, min-residual'):
            st[g.replace('network', 'untrained network', 1)] = ctl[g]
        _accumulate(totals, group, st)
        per_case.append(dict(case=k, grid=f"{case['nx']}x{case['ny']}", Re=round(case['Re'], 2), rank=int(c.shape[1]),
                             steps={g: round(v['mean_iterations'], 2) for g, v in st.items()}))
        say(f"ensemble case {k} ({case['nx']}x{case['ny']}, Re {case['Re']:.0f}, r {c.shape[1]}): "
            + ', '.join(f"{g} {v['mean_iterations']:.1f}" for g, v in st.items()))
        dev.h.close()
        flows[k] = None
    out = dict(training_pairs=pairs, training_groups=groups, training_steps=train_steps, training_seconds=t2 - t1,
               loss_first=float(losses[:5].mean()), loss_last=float(losses[-50:].mean()), cases=per_case,
               failed=failed, summary=_summary(totals), seconds=time.perf_counter() - t0)
    net.close()
    control.close()
    return out
