"""``MAC_CylinderFlow`` for the ghost-cell AB2/CN path, stepping on the GPU.

Mirrors the reference facade (solver/main.py:19-617) for
``boundary='ghost_cell', time_scheme='ab2_cn'``: same constructor keywords and
validation, same fields (``u (ny, nx+1)``, ``v (ny+1, nx)``, ``p (ny, nx)``),
``solve()`` returning ``(t, cd, cl)``, ``last_step_diagnostics``, the ``gg``
integrator with the reference's operator attributes, and npz checkpoints with
the same keys.  Every step runs in libghostflow.so (HIP, gfx950).
"""
import hashlib
import json
import os
import platform
import tempfile
import time
from pathlib import Path

import numpy as np
import scipy

from . import _native as nat
from .device import GhostDevice, PreparedCase, solve_record
from .tables import GhostOperators

SOURCE_FILES = ('flow.py', 'device.py', 'tables.py', '_native.py', 'csrc/ghostflow.hip')


def gaussian_v_perturbation(xv, yv, xc, yc, R, D, seed, amplitude=1e-3, width=0.5):
    """Seeded transverse perturbation v += a xi exp(-r^2/(2 w^2)) on faces with 0 < r-R <= D.

    xi ~ N(0, 1) is drawn for the whole (ny+1, nx) v array in row-major order from
    ``numpy.random.default_rng(seed)``; used by the 1320x246 Re=100 workload.
    """
    X, Y = np.meshgrid(xv, yv)
    r = np.hypot(X - xc, Y - yc)
    xi = np.random.default_rng(seed).standard_normal(X.shape)
    band = (r - R > 0) & (r - R <= D)
    return np.where(band, amplitude * xi * np.exp(-r * r / (2 * width * width)), 0.)


class GhostCellIntegrator:
    """The reference's ``GhostCellIntegrator`` surface (optimized_ghost_cell.py:61-286) over the device."""

    def __init__(self, solver, prebuilt=None):
        s = self.solver = solver
        if isinstance(prebuilt, PreparedCase):
            # recorded uploads: the operators themselves stay on the host that prepared them
            if not np.array_equal(prebuilt.u_inlet, s.u_inlet) or prebuilt.u_ref != s.U_ref:
                raise ValueError('prepared case does not match this configuration')
            self.device = GhostDevice.from_prepared(prebuilt)
            ops = self.device.ops
        else:
            if prebuilt is not None:
                ops, pressure = prebuilt
            else:
                ops = GhostOperators(s.nx, s.ny, s.Lx, s.Ly, s.D, s.xc, s.yc, s.nu, s.rho, s.u_inlet,
                                     interpolation=s.ghost_velocity_interpolation,
                                     pressure_extension=s.ghost_pressure_extension)
                pressure = None
            self.device = GhostDevice(ops, s.u_inlet, pressure=pressure)
            self.device.set_reference_velocity(s.U_ref)
        # optimized_ghost_cell.py:109-111: GMRES preconditioned by the rectangle inverse when iterative
        self.device.set_pressure_backend(s.pressure_solver == 'iterative')
        self.ops = ops
        for name in ('ui', 'vi', 'pi', 'qsize', 'E', 'P', 'G', 'A', 'L', 'Dfull', 'offset', 'div_offset',
                     'diff_offset', 'outlet_u', 'du_dn', 'dv_dn', 'pressure_surface', 'normals',
                     'normal_weights', 'surface_points', 'usize'):
            setattr(self, name, getattr(ops, name))
        self.D = ops.Dop
        self.pressure_inverse_info = self.device.pressure_info
        self.helmholtz_inverse_info = []
        self.previous_advection = self.previous_dt = self.half_pressure = None
        self.last_diagnostics = {}
        self.reconstruction_info = dict(velocity=s.ghost_velocity_interpolation, pressure=s.ghost_pressure_extension)
        if s.ghost_pressure_extension == 'quadratic':   # optimized_ghost_cell.py:76-77
            self.reconstruction_info.update({k: v for k, v in ops.ghost_p.items() if k != 'ghost_nodes'})

    # --- packing (optimized_ghost_cell.py:174-184)
    def pack(self, u, v, out=None):
        q = np.concatenate((np.asarray(u).ravel()[self.ui], np.asarray(v).ravel()[self.vi]))
        if out is not None:
            out[:] = q
            return out
        return q

    def unpack(self, q):
        return self.device.extend(q)

    # --- operators (optimized_ghost_cell.py:186-216, 231-234), evaluated by the device
    def advection(self, q, compiled=True, out=None):
        return self._apply(nat.APPLY_ADVECTION, self.device.q_to_vel(q), out, vel_out=True)

    def diffusion(self, q, out=None):
        return self._apply(nat.APPLY_DIFFUSION, self.device.q_to_vel(q), out, vel_out=True)

    def gradient(self, p, out=None):
        p = np.asarray(p)
        full = p.ravel() if p.ndim > 1 else self.device.q_to_p(p)
        if p.ndim > 1:
            full = np.where(self.device.body['p'].ravel(), 0., full)
        return self._apply(nat.APPLY_GRADIENT, full, out, vel_out=True)

    def divergence(self, q, out=None):
        return self._apply(nat.APPLY_DIVERGENCE, self.device.q_to_vel(q), out, vel_out=False)

    def H_matvec(self, q, dt, out=None):
        return self._apply(nat.APPLY_HMATVEC, self.device.q_to_vel(q), out, vel_out=True, dt=dt)

    def _apply(self, op, arr, out, vel_out, dt=0.):
        o = self.ops
        res = np.empty(o.usize + o.vsize if vel_out else o.nx * o.ny)
        self.device.h.call('mcf_apply_operator', op, nat.dptr(nat.f64(arr)), nat.dptr(res), float(dt))
        r = self.device.vel_to_q(res) if vel_out else self.device.p_to_q(res)
        if out is not None:
            out[:] = r
            return out
        return r

    def pressure_solve(self, values):
        """Monitored pressure solve of values on the active cells (pressure_schur.py:33-71); the
        record lands in ``pressure_inverse_info['last_solve']``."""
        return self.device.pressure_solve(values)

    def project(self, q, dt, out=None):
        q_out, phi, info = self.device.project(q, dt)
        if out is not None:
            out[:] = q_out
            q_out = out
        return q_out, phi

    def surface_forces(self, u, v, p, mu):
        return self.device.forces(u, v, p, mu)

    def advance(self, u, v, p, dt):
        """One step from (u, v, p) with the stored history (optimized_ghost_cell.py:236-278)."""
        s = self.solver
        self.device.set_state(u, v, p, self.previous_advection, self.half_pressure, self.previous_dt, s.time)
        diag = self.device.run(1, dt)[0]
        st = self.device.get_state()
        self.previous_advection, self.half_pressure, self.previous_dt = (st['previous_advection'],
                                                                         st['half_pressure'], st['previous_dt'])
        self.last_diagnostics = diagnostics_from_row(diag, self.pressure_inverse_info['backend'])
        self.pressure_inverse_info['last_solve'] = self.last_diagnostics['pressure_solves'][-1]
        self.helmholtz_inverse_info = self.device.helmholtz_inverse_info
        return st['u'], st['v'], st['p']


def diagnostics_from_row(row, backend='fft_boundary_schur'):
    """The reference's per-step diagnostics dict (optimized_ghost_cell.py:266-276; main.py:409-413):
    one ``last_solve`` record per pressure solve of the step, in call order, and their aggregates.
    Per-solve seconds are wall times of the host-driven GMRES solves; solves replayed inside a step
    graph report 0."""
    d = {n: float(row[k]) for k, n in enumerate(nat.DIAG_NAMES) if not n.startswith('unused')}
    count = int(d['pressure_solve_count'])
    solves = [solve_record(row[nat.DIAG_SOLVE0 + nat.SOLVE_LEN * q:nat.DIAG_SOLVE0 + nat.SOLVE_LEN * (q + 1)])
              for q in range(min(count, nat.MAX_SOLVES))]
    return dict(momentum_iterations=0, momentum_residual_linf=d['momentum_residual_linf'],
                coupled_momentum_residual_linf=d['coupled_momentum_residual_linf'],
                pressure_backend=backend, pressure_solve_count=count,
                pressure_iterations=sum(c['iterations'] for c in solves),
                pressure_refinement_iterations=sum(c['refinement_iterations'] for c in solves),
                pressure_residual_linf=max((c['residual_linf'] for c in solves), default=0.0),
                pressure_residual_relative_l2=max((c['residual_relative_l2'] for c in solves), default=0.0),
                pressure_seconds=sum(c['seconds'] for c in solves), pressure_solves=solves,
                helmholtz_refinement_iterations=int(d['helmholtz_refinement_iterations']),
                time=d['time'], dt=d['dt'], max_divergence=d['max_divergence'], cfl=d['cfl'], cd=d['cd'], cl=d['cl'],
                cd_pressure=d['cd_pressure'], cd_viscous=d['cd_viscous'], cl_pressure=d['cl_pressure'],
                cl_viscous=d['cl_viscous'])


class MAC_CylinderFlow:
    """Ghost-cell AB2/CN ``MAC_CylinderFlow`` (solver/main.py:19-617) stepping on the GPU."""

    def __init__(self, nx=400, ny=100, Lx=20.0, Ly=5.0, D=1.0, xc=5.0, yc=2.5, nu=0.01, rho=1.0, U_inf=1.0,
                 inflow='uniform', U_m=None, slip_walls=False, verbose=True, boundary='immersed',
                 transport='centered', pressure_solver='direct', marker_spacing_factor=1.5, time_scheme='legacy',
                 ghost_pressure_extension='neumann', ghost_velocity_interpolation='bilinear', initial='profile',
                 v_perturbation=None, prebuilt=None):
        # argument checks (main.py:45-80)
        if boundary not in ('immersed', 'staircase', 'ghost_cell', 'cut_cell') or transport not in ('centered', 'upwind'):
            raise ValueError('Choose boundary=immersed/staircase/ghost_cell/cut_cell and transport=centered/upwind')
        if pressure_solver not in ('direct', 'iterative', 'cg_dct'):
            raise ValueError("pressure_solver must be direct, iterative or cg_dct")
        if time_scheme not in ('legacy', 'ab2_cn'):
            raise ValueError('time_scheme must be legacy or ab2_cn')
        if time_scheme == 'ab2_cn' and (boundary not in ('cut_cell', 'ghost_cell') or transport != 'centered'):
            raise ValueError('ab2_cn requires cut_cell/ghost_cell and centered transport')
        if ghost_pressure_extension not in ('neumann', 'quadratic') or ghost_velocity_interpolation not in ('bilinear', 'biquadratic'):
            raise ValueError('Unknown Ghost pressure extension or velocity interpolation')
        if boundary != 'ghost_cell' or time_scheme != 'ab2_cn':
            raise NotImplementedError('This GPU build provides the ghost_cell / ab2_cn path only')
        if pressure_solver not in ('direct', 'iterative') or slip_walls:
            raise ValueError('Ghost AB2/CN requires direct/iterative pressure and no-slip channel walls')
        if any(isinstance(n, bool) or not isinstance(n, (int, np.integer)) or n < 4 for n in (nx, ny)):
            raise ValueError('nx and ny must be integers >= 4')
        if not all(np.isfinite(x) and x > 0 for x in (Lx, Ly, D, nu, rho, U_inf)):
            raise ValueError('Domain lengths, D, nu, rho and U_inf must be finite and positive')
        if inflow not in ('uniform', 'parabolic'):
            raise ValueError('inflow must be uniform or parabolic')
        if U_m is not None and (not np.isfinite(U_m) or U_m <= 0):
            raise ValueError('U_m must be finite and positive')
        if not (D / 2 < xc < Lx - D / 2 and D / 2 < yc < Ly - D / 2):
            raise ValueError('Cylinder must lie strictly inside the channel')
        self.time_scheme, self.boundary, self.transport, self.pressure_solver = time_scheme, boundary, transport, pressure_solver
        self.ghost_pressure_extension = ghost_pressure_extension
        self.ghost_velocity_interpolation = ghost_velocity_interpolation
        self.marker_spacing_factor = marker_spacing_factor
        self.configuration = self.configuration_of(
            nx=nx, ny=ny, Lx=Lx, Ly=Ly, D=D, xc=xc, yc=yc, nu=nu, rho=rho, U_inf=U_inf, inflow=inflow, U_m=U_m,
            slip_walls=slip_walls, boundary=boundary, transport=transport, pressure_solver=pressure_solver,
            marker_spacing_factor=marker_spacing_factor, time_scheme=time_scheme,
            ghost_pressure_extension=ghost_pressure_extension, ghost_velocity_interpolation=ghost_velocity_interpolation)
        self.run_options = {}
        self.last_step_diagnostics = {}
        self.time = 0.0
        self.nx, self.ny, self.Lx, self.Ly = nx, ny, Lx, Ly
        self.dx, self.dy = Lx / nx, Ly / ny
        self.verbose = verbose
        self.nu, self.rho, self.U_inf, self.inflow, self.slip_walls = nu, rho, U_inf, inflow, slip_walls
        self.D, self.xc, self.yc = D, xc, yc
        if inflow == 'parabolic':
            self.U_m = 1.5 * U_inf if U_m is None else U_m
            self.U_ref = (2.0 / 3.0) * self.U_m
        else:
            self.U_m = U_inf
            self.U_ref = U_inf
        self.Re = self.U_ref * self.D / self.nu
        self.y_centers = (np.arange(ny) + 0.5) * self.dy
        self.x_centers = (np.arange(nx) + 0.5) * self.dx
        r_norm_sq = (((self.x_centers - xc) / (D / 2.0))**2)[None, :] + (((self.y_centers - yc) / (D / 2.0))**2)[:, None]
        self.mask = r_norm_sq <= (1.0 + 1e-13)
        if not self.mask.any() or self.mask[[0, -1], :].any() or self.mask[:, [0, -1]].any():
            raise ValueError('Grid must resolve the cylinder and a fluid layer at every outer boundary')
        self.u = np.zeros((ny, nx + 1))
        self.v = np.zeros((ny + 1, nx))
        self.p = np.zeros((ny, nx))
        self.phi = np.zeros_like(self.p)
        y = self.y_centers
        if inflow == 'parabolic':
            self.u_inlet = 4.0 * self.U_m * y * (Ly - y) / (Ly**2)
        else:
            self.u_inlet = np.full(ny, float(U_inf))
        self.u_inlet_base = self.u_inlet.copy()
        if initial == 'profile':   # main.py:166-174
            self.u[:, :] = self.u_inlet[:, None]
            jm, im = np.nonzero(self.mask)
            self.u[jm, im] = 0.; self.u[jm, im + 1] = 0.
            self.v[jm, im] = 0.; self.v[jm + 1, im] = 0.
        elif initial == 'zero':
            self.u[:, 0] = self.u_inlet
        else:
            raise ValueError("initial must be 'profile' or 'zero'")
        t0 = time.time()
        self.gg = GhostCellIntegrator(self, prebuilt=prebuilt)
        self.gc = self.gg.ops.gc
        if v_perturbation is not None:
            self.v += gaussian_v_perturbation(self.gc.xv, self.gc.yv, xc, yc, D / 2, D, **v_perturbation)
        self.A = self.gg.A
        initial_q, _ = self.gg.project(self.gg.pack(self.u, self.v), 1.)   # main.py:229-230
        self.u[:], self.v[:] = self.gg.unpack(initial_q)
        self._device_current = False
        self.setup_seconds = time.time() - t0
        if verbose:
            print(f"Ghost AB2/CN device ready in {self.setup_seconds:.3f}s (Re={self.Re:.1f}).")

    CONFIGURATION_KEYS = ('nx', 'ny', 'Lx', 'Ly', 'D', 'xc', 'yc', 'nu', 'rho', 'U_inf', 'inflow', 'U_m', 'slip_walls',
                          'boundary', 'transport', 'pressure_solver', 'marker_spacing_factor', 'time_scheme', 'verbose',
                          'ghost_pressure_extension', 'ghost_velocity_interpolation')

    @classmethod
    def configuration_of(cls, **options):
        """The ``configuration`` a solver built from these keyword options records (defaults filled,
        verbose False), as the stage-JSON runner compares it (validation/run_ghost_validation.py:106-111)."""
        import inspect
        bound = inspect.signature(cls.__init__).bind(None, **options)
        bound.apply_defaults()
        args = dict(bound.arguments)
        args['verbose'] = False
        return {k: args[k] for k in cls.CONFIGURATION_KEYS}

    # ------------------------------------------------------------------ stepping
    def _push(self):
        g = self.gg
        g.device.set_state(self.u, self.v, self.p, g.previous_advection, g.half_pressure, g.previous_dt, self.time)
        self._device_current = True

    def _pull(self):
        g = self.gg
        st = g.device.get_state()
        self.u[:], self.v[:], self.p[:] = st['u'], st['v'], st['p']
        g.previous_advection, g.half_pressure, g.previous_dt = st['previous_advection'], st['half_pressure'], st['previous_dt']
        g.helmholtz_inverse_info = g.device.helmholtz_inverse_info

    def _perturbation_shape(self, perturb_amp):
        """amp sin(2 pi y / Ly) at the inlet cell centres (main.py:283)."""
        return perturb_amp * np.sin(2.0 * np.pi * self.y_centers / self.Ly)

    def _run_options(self, n_steps, dt, perturb_transient, perturb_duration, perturb_amp):
        """Argument checks of main.py:245-256; returns the perturbation duration."""
        if isinstance(n_steps, bool) or not isinstance(n_steps, (int, np.integer)) or n_steps <= 0:
            raise ValueError('n_steps must be a positive integer')
        if not np.isfinite(dt) or dt <= 0:
            raise ValueError('dt must be finite and positive')
        if perturb_duration is None:
            perturb_duration = 2.5 * (self.D / self.U_ref)
        if not np.isfinite(perturb_duration) or perturb_duration <= 0 or not np.isfinite(perturb_amp) or abs(perturb_amp) >= 1:
            raise ValueError('Require perturb_duration > 0 and finite |perturb_amp| < 1')
        self.run_options = dict(perturb_transient=perturb_transient, perturb_duration=perturb_duration, perturb_amp=perturb_amp)
        return perturb_duration

    def advance(self, n_steps, dt, *, segment=None, perturb_transient=True, perturb_duration=None, perturb_amp=0.15):
        """Run n_steps on the device and yield each segment's diagnostics rows ((k, DIAG_LEN) arrays).

        One iteration of main.py:271-434 per step.  The device runs the steps in graph-replayed chunks
        with no host round trip; ``u, v, p``, ``time`` and ``last_step_diagnostics`` are brought up to
        date after every segment of at most ``segment`` steps (all steps at once by default).  For a
        uniform inflow with ``perturb_transient`` the device sets each step's inlet from the step's end
        time (main.py:280-286) and ``u_inlet`` holds the last step's inlet afterwards.  ``segment`` is a
        step count or a list of segment lengths adding up to n_steps.  A failed gate
        raises after the segment that holds the failing step: the exception carries ``diagnostics``
        (the segment's rows) and ``failed_step`` (its index in them), ``last_step_diagnostics`` is the
        failing step's, and the state is the one at the end of that segment (the device completes the
        chunk it was replaying).
        """
        T = self._run_options(n_steps, dt, perturb_transient, perturb_duration, perturb_amp)
        rate = np.max(np.abs(self.u)) / self.dx + np.max(np.abs(self.v)) / self.dy
        if not np.isfinite(rate) or dt * rate > .5:
            raise RuntimeError(f'Explicit timestep unsafe at t={self.time:g}: dt*rate={dt * rate:g}; reduce dt')
        dev = self.gg.device
        backend = dev.pressure_info['backend']
        start = self.time
        shape = self._perturbation_shape(perturb_amp)
        # steps s whose end time start + (s+1) dt lies in the perturbation window come first
        perturbed = 0
        if perturb_transient and self.inflow == 'uniform':
            perturbed = int(np.sum(start + (np.arange(n_steps) + 1) * dt <= T))

        def inlet_after(steps):
            """The inlet of the last of `steps` steps (main.py:280-286)."""
            t = start + steps * dt
            if steps and steps <= perturbed:
                return self.u_inlet_base * (1.0 + shape * (1.0 - (t / T)))
            return self.u_inlet_base

        self._push()
        dev.set_dt(dt)
        if perturbed:
            dev.set_inlet_perturbation(self.u_inlet_base, shape, T)
        elif not np.array_equal(self.u_inlet, self.u_inlet_base):
            self.u_inlet[:] = self.u_inlet_base
            dev.set_inlet(self.u_inlet)
        perturbing = perturbed > 0
        if segment is None:
            plan = [n_steps]
        elif isinstance(segment, (int, np.integer)):
            k = max(1, int(segment))
            plan = [k] * (n_steps // k) + ([n_steps % k] if n_steps % k else [])
        else:
            plan = [int(k) for k in segment]
            if sum(plan) != n_steps or min(plan) <= 0:
                raise ValueError('segment lengths must be positive and add up to n_steps')
        done = 0
        try:
            for length in plan:
                pieces = []
                while sum(len(r) for r in pieces) < length:
                    if perturbing and done >= perturbed:
                        # the window is over: the remaining steps keep the base profile (no k_inlet launch)
                        dev.set_inlet_perturbation(None, None, 0.0)
                        perturbing = False
                        self.u_inlet[:] = self.u_inlet_base
                        dev.set_inlet(self.u_inlet)
                    k = length - sum(len(r) for r in pieces)
                    if perturbing:
                        k = min(k, perturbed - done)
                    try:
                        pieces.append(dev.run(k, dt))
                    except (RuntimeError, FloatingPointError) as error:
                        self._pull()
                        rows = error.diagnostics
                        ran = np.flatnonzero(rows[:, 0])
                        if len(ran):
                            self.time = float(rows[ran[-1], 0])
                            self.u_inlet[:] = inlet_after(done + len(ran))
                        failed = error.failed_step if error.failed_step is not None else (ran[-1] if len(ran) else 0)
                        self.last_step_diagnostics = diagnostics_from_row(rows[failed], backend)
                        # the exception reports the whole planned segment's rows
                        error.diagnostics = np.concatenate(pieces + [rows])
                        error.failed_step = failed + sum(len(r) for r in pieces)
                        raise
                    done += k
                rows = pieces[0] if len(pieces) == 1 else np.concatenate(pieces)
                self._pull()
                self.time = float(rows[-1, 0])
                self.u_inlet[:] = inlet_after(done)
                self.last_step_diagnostics = diagnostics_from_row(rows[-1], backend)
                yield rows
        finally:
            if perturbing:
                dev.set_inlet_perturbation(None, None, 0.0)
                dev.set_inlet(self.u_inlet)
            self.gg.last_diagnostics = self.last_step_diagnostics
            if self.last_step_diagnostics.get('pressure_solves'):
                self.gg.pressure_inverse_info['last_solve'] = self.last_step_diagnostics['pressure_solves'][-1]

    def solve(self, n_steps=25000, dt=0.002, perturb_transient=True, perturb_duration=None, perturb_amp=0.15,
              save_artifacts=True, verbose=None, step_callback=None):
        """Step loop of main.py:245-501; returns (t, cd, cl).  No plot or npz artefacts are written.
        With a ``step_callback`` the steps run one per segment so that the callback sees each step's
        state, as in the reference; without one they run in one device segment."""
        rows = []
        for seg in self.advance(n_steps, dt, segment=1 if step_callback is not None else None,
                                perturb_transient=perturb_transient, perturb_duration=perturb_duration,
                                perturb_amp=perturb_amp):
            rows.append(seg)
            if step_callback is not None:
                step_callback(self, self.last_step_diagnostics.copy())
        diag = np.concatenate(rows)
        self.diagnostics_history = diag
        return diag[:, 0].copy(), diag[:, 2].copy(), diag[:, 3].copy()

    def step(self, dt, *, perturb_transient=True, perturb_duration=None, perturb_amp=0.15):
        """One step (one iteration of main.py:271-434); returns the step's diagnostics dict."""
        for _ in self.advance(1, dt, perturb_transient=perturb_transient, perturb_duration=perturb_duration,
                              perturb_amp=perturb_amp):
            pass
        return dict(self.last_step_diagnostics)

    # ------------------------------------------------------------------ checkpoint (main.py:503-585)
    def save_checkpoint(self, path, *, overwrite=False):
        path = Path(path)
        if path.exists() and not overwrite:
            raise FileExistsError(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        arrays = {k: getattr(self, k) for k in ('u', 'v', 'p', 'phi', 'u_inlet')}
        for k in ('previous_advection', 'half_pressure'):
            if getattr(self.gg, k) is not None:
                arrays[k] = getattr(self.gg, k)
        launch = self.gg.device.get_launch()
        meta = dict(schema=1, configuration=self.configuration, time=self.time, run_options=self.run_options,
                    prev_dt=self.gg.previous_dt, cn_dt=self.gg.previous_dt, diagnostics=self.last_step_diagnostics,
                    source_sha256=self.metadata(1., 0)['source_sha256'],
                    launch=launch if launch['chosen'] else None)
        temporary = None
        try:
            with tempfile.NamedTemporaryFile(dir=path.parent, suffix='.npz', delete=False) as stream:
                temporary = Path(stream.name)
                np.savez(stream, checkpoint_metadata=json.dumps(meta, default=float), **arrays)
                stream.flush()
                os.fsync(stream.fileno())
            os.replace(temporary, path)
        finally:
            if temporary is not None and temporary.exists():
                temporary.unlink()

    @classmethod
    def from_checkpoint(cls, path):
        with np.load(path, allow_pickle=False) as data:
            meta = json.loads(str(data['checkpoint_metadata']))
            if meta.get('schema') != 1 or meta['configuration'].get('boundary') != 'ghost_cell':
                raise ValueError('Unsupported checkpoint')
            solver = cls(**meta['configuration'])
            if meta['source_sha256'] != solver.metadata(1., 0)['source_sha256']:
                raise ValueError('Checkpoint source mismatch')
            for name in ('u', 'v', 'p', 'phi', 'u_inlet'):
                value = data[name]
                target = getattr(solver, name)
                if value.shape != target.shape or not np.isfinite(value).all():
                    raise ValueError(f'Invalid checkpoint field: {name}')
                target[:] = value
            solver.time = float(meta['time'])
            if not np.isfinite(solver.time) or solver.time < 0:
                raise ValueError('Invalid checkpoint time')
            solver.run_options = meta['run_options']
            solver.last_step_diagnostics = meta['diagnostics']
            for name, shape in (('previous_advection', (solver.gg.qsize,)), ('half_pressure', (len(solver.gg.pi),))):
                value = data[name].copy() if name in data else None
                if solver.time > 0 and (value is None or value.shape != shape or not np.isfinite(value).all()):
                    raise ValueError(f'Incomplete Ghost AB2/CN history: {name}')
                setattr(solver.gg, name, value)
            solver.gg.previous_dt = meta['prev_dt']
            if solver.time > 0 and (solver.gg.previous_dt is None or not np.isfinite(solver.gg.previous_dt)
                                    or solver.gg.previous_dt <= 0):
                raise ValueError('Invalid Ghost timestep history')
            solver.gg.device.set_inlet(solver.u_inlet)
            if meta.get('launch'):
                # the launch shapes the run was tuned to, so that the resumed run rounds as it would have
                solver.gg.device.set_launch_classes(meta['launch'])
            return solver

    def metadata(self, dt, n_steps, perturb_transient=True, perturb_duration=None, perturb_amp=.15):
        """Inputs plus SHA-256 of this package's sources and library (main.py:587-617)."""
        root = Path(__file__).resolve().parent
        hashes = {name: hashlib.sha256((root / name).read_bytes()).hexdigest() for name in SOURCE_FILES}
        if perturb_duration is None:
            perturb_duration = 2.5 * self.D / self.U_ref
        return dict(nx=self.nx, ny=self.ny, Lx=self.Lx, Ly=self.Ly, D=self.D, xc=self.xc, yc=self.yc, nu=self.nu,
                    rho=self.rho, U_ref=self.U_ref, U_inf=self.U_inf, U_m=self.U_m, inflow=self.inflow,
                    slip_walls=self.slip_walls, dt=dt, n_steps=n_steps, end_time=self.time, boundary=self.boundary,
                    transport=self.transport, time_integrator='AB2_CN_discrete_incremental',
                    time_scheme=self.time_scheme, pressure_solver=self.pressure_solver,
                    pressure_backend='ghost_' + self.gg.pressure_inverse_info['backend'],
                    ghost_pressure_extension=self.ghost_pressure_extension,
                    ghost_velocity_interpolation=self.ghost_velocity_interpolation,
                    pressure_preconditioner={k: v for k, v in self.gg.pressure_inverse_info.items()},
                    device=nat.device_name(), perturb_transient=perturb_transient, perturb_duration=perturb_duration,
                    perturb_amp=perturb_amp, python=platform.python_version(), numpy=np.__version__,
                    scipy=scipy.__version__, platform=platform.platform(), source_sha256=hashes)
