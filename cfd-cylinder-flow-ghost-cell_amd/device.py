"""Upload of one case's tables to the HIP library and thin wrappers of its entry points.

The host tables come from ``tables.GhostOperators`` (assembled like the
reference integrator, solver/optimized_ghost_cell.py:62-150).  The device
keeps the reference's staggered arrays and applies every operator as a
structured stencil plus "override" rows: the rows of the assembled sparse
operator that differ from the cylinder-free stencil (those touching ghost
nodes).  Those rows are passed verbatim, so the device operator is the
reference's operator.
"""
import numpy as np
import scipy.sparse as sp

from . import _native as nat
from .tables import RectangleFactor, SchurFactor, boundary_rows


def solid_intervals(body):
    """Per-row [lo, hi] of the solid columns of a boolean array (hi < lo: none); rows must be convex."""
    rows, cols = body.shape
    lo = np.ones(rows, dtype=np.int32)
    hi = np.zeros(rows, dtype=np.int32)
    for j in range(rows):
        idx = np.flatnonzero(body[j])
        if len(idx):
            if idx[-1] - idx[0] + 1 != len(idx):
                raise ValueError('Solid region is not convex along a grid row')
            lo[j], hi[j] = idx[0], idx[-1]
    return lo, hi


def csr_rows(matrix, rows):
    """CSR (ptr, col, val) of the selected rows."""
    sub = matrix.tocsr()[rows]
    sub.sort_indices()
    return nat.i32(sub.indptr), sub.indices, nat.f64(sub.data)


def changed_rows(actual, regular, rel=1e-13):
    """Rows where two sparse operators differ by more than rel * max|regular|."""
    diff = (actual.tocsr() - regular.tocsr()).tocsr()
    scale = max(abs(regular).max(), 1e-300)
    diff.data[abs(diff.data) <= rel * scale] = 0.
    diff.eliminate_zeros()
    return np.flatnonzero(np.diff(diff.indptr))


def helmholtz_pieces(ops, dt):
    """Per-dt u/v Helmholtz factors, capacitance data and override rows (optimized_ghost_cell.py:243-255)."""
    pieces = []
    for kind, selected, actual in ops.helmholtz_actual(dt):
        factor = RectangleFactor(kind, ops.nx, ops.ny, ops.dx, ops.dy, nu=ops.nu, dt=dt, rho=ops.rho)
        schur = SchurFactor(actual, selected, factor)
        pieces.append((kind, selected, actual, factor, schur))
    return pieces


def pressure_pieces(ops):
    factor = RectangleFactor('pressure', ops.nx, ops.ny, ops.dx, ops.dy, rho=ops.rho)
    return factor, SchurFactor(ops.A, ops.pi, factor)


def solve_record(rec):
    """The reference's ``last_solve`` dict (pressure_schur.py:64-68) from one MCF_SOLVE_LEN record."""
    d = {name: float(rec[k]) for k, name in enumerate(nat.SOLVE_NAMES)}
    for key in ('iterations', 'refinement_iterations', 'status'):
        d[key] = int(d[key])
    return d


class PreparedCase:
    """A case's recorded table uploads, prepared on the host without a device (picklable, compact):
    the entry point calls with their arrays, the index layout and the solver info."""

    def __init__(self, layout, calls, body, u_inlet, u_ref, dt, pressure_info, helmholtz_info, geometry):
        self.layout, self.calls, self.body = layout, calls, body
        self.schur_geometry = geometry
        self.u_inlet, self.u_ref, self.dt = u_inlet, u_ref, dt
        self.pressure_info, self.helmholtz_info = pressure_info, helmholtz_info


def schur_geometry(ops, rows):
    """Boundary-angle order of the capacitance rows (pressure cells, row-major rectangle index) and
    per-row features for the warm-start network: signed distance of the cell centre to the
    cylinder surface in cells, cos and sin of its angle about the centre."""
    j, i = np.divmod(np.asarray(rows), ops.nx)
    x, y = (i + 0.5) * ops.dx, (j + 0.5) * ops.dy
    ang = np.arctan2(y - ops.yc, x - ops.xc)
    dist = np.hypot(x - ops.xc, y - ops.yc)
    order = np.lexsort((dist, ang)).astype(np.int32)
    h = np.sqrt(ops.dx * ops.dy)
    feats = np.stack(((dist - ops.R) / h, np.cos(ang), np.sin(ang)), axis=1)[order]
    return order, np.ascontiguousarray(feats)


def prepare_case(ops, u_inlet, u_ref, dt, pressure=None):
    """Build every table upload of one case for time step dt against a ``Recorder``."""
    rec = nat.Recorder()
    dev = GhostDevice(ops, u_inlet, pressure=pressure, handle=rec)
    dev.set_reference_velocity(u_ref)
    dev.set_dt(dt)
    return PreparedCase(ops.layout(), rec.calls, dev.body, dev.u_inlet, u_ref, dt, dev.pressure_info,
                        dev.helmholtz_inverse_info, dev.schur_geometry)


class GhostDevice:
    """Device state and tables of one case (one ``mcf_handle``)."""

    def __init__(self, ops, u_inlet, pressure=None, handle=None):
        self.ops = o = ops
        self.h = handle if handle is not None else nat.Handle(o.nx, o.ny, o.dx, o.dy, o.nu, o.rho)
        self.prepared = False
        self.uref = None
        self.set_inlet(u_inlet)
        gc = o.gc
        self.body = dict(u=gc.dist_u <= 1e-13, v=gc.dist_v <= 1e-13, p=gc.dist_p <= 1e-13)
        for field, key in ((nat.FIELD_U, 'u'), (nat.FIELD_V, 'v'), (nat.FIELD_P, 'p')):
            lo, hi = solid_intervals(self.body[key])
            self.h.call('mcf_set_solid_intervals', field, lo, hi)
        self._upload_ghosts()
        self._upload_overrides()
        self.pressure_factor, self.pressure_schur = pressure if pressure is not None else pressure_pieces(o)
        self._upload_rectangle(nat.PRESSURE, self.pressure_factor, self.pressure_schur)
        ps = self.pressure_schur
        order, feats = schur_geometry(o, ps.rows)
        Cp = ps.Cp.tocsr()
        Cp.sort_indices()
        self.h.call('mcf_set_schur_system', nat.PRESSURE, nat.f64(ps.S), nat.i32(Cp.indptr), nat.i32(Cp.indices),
                    nat.f64(Cp.data), order)
        self.schur_geometry = dict(order=order, features=feats, S=np.ascontiguousarray(ps.S[np.ix_(order, order)]))
        self.pressure_info = dict(backend='fft_boundary_schur', rank=ps.rank, dense_bytes=int(ps.S.nbytes),
                                  roundoff_drop_tolerance=ps.tolerance)
        self._upload_surface()
        self.dt = None
        self.helmholtz_factors = None

    @classmethod
    def from_prepared(cls, prep):
        """A device case from recorded uploads; it keeps the operators' index layout only."""
        self = cls.__new__(cls)
        self.ops = o = prep.layout
        self.h = nat.Handle(o.nx, o.ny, o.dx, o.dy, o.nu, o.rho)
        for name, args in prep.calls:
            self.h.call(name, *args)
        self.prepared = True
        self.body = prep.body
        self.u_inlet = prep.u_inlet
        self.uref = prep.u_ref
        self.dt = prep.dt
        self.pressure_info = prep.pressure_info
        self.helmholtz_inverse_info = prep.helmholtz_info
        self.schur_geometry = prep.schur_geometry
        self.pressure_factor = self.pressure_schur = self.helmholtz_factors = None
        return self

    # ---------------------------------------------------------------- tables
    def set_inlet(self, u_inlet):
        self.u_inlet = nat.f64(u_inlet).copy()
        self.h.call('mcf_set_inlet', self.u_inlet)

    def _upload_ghosts(self):
        o = self.ops
        for field, E, ids, info in ((nat.FIELD_U, o.Eu, o.ui, o.ghost_u), (nat.FIELD_V, o.Ev, o.vi, o.ghost_v),
                                    (nat.FIELD_P, o.P, o.pi, o.ghost_p)):
            nodes = np.asarray(info['ghost_nodes'])
            ptr, col, val = csr_rows(E, nodes)
            self.h.call('mcf_set_ghost', field, len(nodes), nat.i32(nodes), ptr, nat.i32(ids[col]), val)

    def _vel_index(self, qcols):
        """Packed velocity index -> vel-layout index."""
        o = self.ops
        qcols = np.asarray(qcols)
        return np.where(qcols < o.nu_act, o.ui[np.minimum(qcols, o.nu_act - 1)],
                        o.usize + o.vi[np.clip(qcols - o.nu_act, 0, len(o.vi) - 1)])

    def _set_override(self, op, out_index, matrix, rows, col_map):
        ptr, col, val = csr_rows(matrix, rows)
        self.h.call('mcf_set_override', op, len(rows), nat.i32(out_index[rows]), ptr, nat.i32(col_map(col)), val)
        return len(rows)

    def _upload_overrides(self):
        o = self.ops
        nu = o.nu_act
        # gradient: regular stencil = raw_g / rho over the full p array
        Gfull = sp.csr_matrix((o.G.data, o.pi[o.G.indices], o.G.indptr), shape=(o.qsize, o.nx * o.ny))
        rows = changed_rows(Gfull, o.raw_g / o.rho)
        ru, rv = rows[rows < nu], rows[rows >= nu]
        self.n_override = {}
        self.n_override['grad_u'] = self._set_override(nat.OP_GRAD_U, o.ui, Gfull[:nu], ru, lambda c: c)
        self.n_override['grad_v'] = self._set_override(nat.OP_GRAD_V, o.vi, Gfull[nu:], rv - nu, lambda c: c)
        # divergence: regular stencil = Dfull without the inlet column and the v wall rows
        Dmapped = sp.csr_matrix((o.Dop.data, self._vel_index(o.Dop.indices), o.Dop.indptr), shape=o.Dfull.shape)
        keep = np.ones(o.Dfull.shape[1], bool)
        keep[np.arange(o.ny) * (o.nx + 1)] = False
        keep[o.usize + np.arange(o.nx)] = False
        keep[o.usize + o.ny * o.nx + np.arange(o.nx)] = False
        Dreg = o.Dfull @ sp.diags(keep.astype(float))
        rows = changed_rows(Dmapped, Dreg)
        self.n_override['div'] = self._set_override(nat.OP_DIV, o.pi, Dmapped, rows, lambda c: c)
        # pressure operator: rows where A differs from the rectangle (pressure_schur.py:78-92)
        rect = self.pressure_factor_matrix()
        brows, _, _ = boundary_rows(o.A, o.pi, rect)
        pos = np.full(o.nx * o.ny, -1)
        pos[o.pi] = np.arange(len(o.pi))
        arows = pos[brows]
        arows = arows[arows >= 0]
        self.n_override['pressure'] = self._set_override(nat.OP_PRESSURE, o.pi, o.A.tocsr(), arows,
                                                         lambda c: o.pi[c])

    def pressure_factor_matrix(self):
        from .tables import rectangle_matrix
        o = self.ops
        return rectangle_matrix('pressure', o.nx, o.ny, o.dx, o.dy, rho=o.rho)

    def _upload_rectangle(self, kind, factor, schur):
        yy, xx = factor.shape
        self.h.call('mcf_set_rectangle', kind, yy, xx, float(factor.a), float(factor.b), nat.f64(factor.V),
                    nat.f64(factor.lam), nat.f64(factor.lower), nat.f64(factor.diag), nat.f64(factor.upper))
        self.h.call('mcf_set_schur', kind, schur.rank, nat.i32(schur.rows), len(schur.points), nat.i32(schur.points),
                    nat.f64(schur.SC))

    def _upload_surface(self):
        o = self.ops
        mats = []
        for M in (o.pressure_surface, o.du_dn, o.dv_dn):
            M = M.tocsr(); M.sort_indices()
            mats.append((nat.i32(M.indptr), nat.i32(M.indices), nat.f64(M.data)))
        (pp, pc, pv), (up, uc, uv), (vp, vc, vv) = mats
        self._surface = mats
        self._normals = nat.f64(o.normals)
        self._weights = nat.f64(o.normal_weights)

    def set_reference_velocity(self, u_ref):
        if self.uref == u_ref:
            return
        if self.prepared:
            raise ValueError('a prepared case keeps the reference velocity it was prepared with')
        (pp, pc, pv), (up, uc, uv), (vp, vc, vv) = self._surface
        self.h.call('mcf_set_surface', len(self._normals), pp, pc, pv, up, uc, uv, vp, vc, vv, self._normals,
                    self._weights, float(self.ops.D), float(u_ref))
        self.uref = u_ref

    def set_dt(self, dt, pieces=None):
        """Build (or take prebuilt) and upload the u/v Helmholtz factors for dt (optimized_ghost_cell.py:243-255)."""
        if self.dt == dt:
            return
        if self.prepared:
            raise ValueError(f'a prepared case keeps the time step it was prepared for ({self.dt!r}, asked {dt!r})')
        o = self.ops
        if pieces is None:
            pieces = helmholtz_pieces(o, dt)
        infos, factors = [], []
        for (kind, selected, actual, factor, schur), code, op in zip(pieces, (nat.U, nat.V), (nat.OP_HELM_U, nat.OP_HELM_V)):
            self._upload_rectangle(code, factor, schur)
            pos = np.full(factor.shape[0] * factor.shape[1], -1)
            pos[selected] = np.arange(len(selected))
            arows = pos[schur.rows]
            arows = arows[arows >= 0]
            if kind == 'u':
                out_index, col_map = o.ui, (lambda c: o.ui[c])
            else:
                out_index, col_map = o.vi, (lambda c: o.usize + o.vi[c])
            self._set_override(op, out_index, actual.tocsr(), arows, col_map)
            infos.append(dict(rank=schur.rank, dense_bytes=int(schur.S.nbytes),
                              roundoff_drop_tolerance=schur.tolerance, override_rows=len(arows)))
            factors.append((factor, schur))
        self.helmholtz_factors = factors
        self.helmholtz_inverse_info = infos
        self.dt = dt

    # ---------------------------------------------------------------- state
    def vel_from(self, u, v):
        return np.concatenate((np.ascontiguousarray(u, dtype=float).ravel(), np.ascontiguousarray(v, dtype=float).ravel()))

    def split_vel(self, vel):
        o = self.ops
        return vel[:o.usize].reshape(o.ushape), vel[o.usize:].reshape(o.vshape)

    def q_to_vel(self, q):
        o = self.ops
        vel = np.zeros(o.usize + o.vsize)
        vel[o.ui] = q[:o.nu_act]
        vel[o.usize + o.vi] = q[o.nu_act:]
        return vel

    def vel_to_q(self, vel):
        o = self.ops
        return np.concatenate((vel[o.ui], vel[o.usize + o.vi]))

    def p_to_q(self, p):
        return np.ascontiguousarray(p).ravel()[self.ops.pi]

    def q_to_p(self, q):
        o = self.ops
        p = np.zeros(o.nx * o.ny)
        p[o.pi] = q
        return p

    def set_state(self, u, v, p, previous_advection=None, half_pressure=None, previous_dt=None, time=0.):
        vel = nat.f64(self.vel_from(u, v))
        pp = nat.f64(np.ascontiguousarray(p, dtype=float).ravel())
        prev = nat.f64(self.q_to_vel(previous_advection)) if previous_advection is not None else None
        half = nat.f64(self.q_to_p(half_pressure)) if half_pressure is not None else None
        self.h.call('mcf_set_state', nat.dptr(vel), nat.dptr(pp), nat.dptr(prev), nat.dptr(half),
                    float(previous_dt or 0.), float(time))

    def get_state(self):
        o = self.ops
        vel = np.empty(o.usize + o.vsize)
        p = np.empty(o.nx * o.ny)
        prev = np.zeros(o.usize + o.vsize)
        half = np.zeros(o.nx * o.ny)
        prev_dt = np.zeros(1); t = np.zeros(1); hist = np.zeros(1, dtype=np.int32)
        self.h.call('mcf_get_state', nat.dptr(vel), nat.dptr(p), nat.dptr(prev), nat.dptr(half), nat.dptr(prev_dt),
                    nat.dptr(t), nat.iptr(hist))
        u, v = self.split_vel(vel)
        has = bool(hist[0])
        return dict(u=u.copy(), v=v.copy(), p=p.reshape(o.pshape),
                    previous_advection=self.vel_to_q(prev) if has else None,
                    half_pressure=self.p_to_q(half.reshape(o.pshape)) if has else None,
                    previous_dt=float(prev_dt[0]) if has else None, time=float(t[0]))

    # ---------------------------------------------------------------- operations
    def run(self, n, dt):
        """n steps at dt; returns the (n, DIAG_LEN) diagnostics rows.  When a step fails a gate the
        library stops after the chunk that holds it: the exception carries ``diagnostics`` (the rows
        computed, zeros after them) and ``failed_step`` (the first row whose status is set)."""
        self.set_dt(dt)
        diag = np.zeros((n, nat.DIAG_LEN))
        try:
            self.h.call('mcf_run', int(n), float(dt), nat.dptr(diag))
        except (RuntimeError, FloatingPointError) as error:
            status = diag[:, nat.DIAG_NAMES.index('status')]
            bad = np.flatnonzero(status != 0)
            error.diagnostics = diag
            error.failed_step = int(bad[0]) if len(bad) else None
            raise
        return diag

    def set_probes(self, matrix):
        """Pressure probes (at most two CSR rows over p; None removes them), evaluated by the device on
        every step's output pressure and recorded in the diagnostics columns probe0 / probe1."""
        if matrix is None:
            self.h.call('mcf_set_probes', 0, None, None, None)
            return
        m = sp.csr_matrix(matrix)
        m.sort_indices()
        self.h.call('mcf_set_probes', m.shape[0], nat.i32(m.indptr), nat.i32(m.indices), nat.f64(m.data))

    def set_inlet_perturbation(self, base, shape, duration):
        """Steps that follow set their own inlet base (1 + shape (1 - t/duration)) from their end time t
        while t <= duration (main.py:280-286); duration <= 0 removes the perturbation."""
        if duration > 0:
            self.h.call('mcf_set_inlet_perturbation', nat.f64(base), nat.f64(shape), float(duration))
        else:
            self.h.call('mcf_set_inlet_perturbation', None, None, 0.0)

    def project(self, q, dt):
        """q (packed) -> (projected q, phi (packed)); optimized_ghost_cell.py:218-229."""
        o = self.ops
        vel = nat.f64(self.q_to_vel(q))
        out = np.empty_like(vel)
        phi = np.empty(o.nx * o.ny)
        info = np.zeros(4)
        self.h.call('mcf_project', nat.dptr(vel), float(dt), nat.dptr(out), nat.dptr(phi), nat.dptr(info))
        return self.vel_to_q(out), self.p_to_q(phi), info

    def extend(self, q):
        vel = nat.f64(self.q_to_vel(q))
        out = np.empty_like(vel)
        self.h.call('mcf_extend', nat.dptr(vel), nat.dptr(out))
        return self.split_vel(out)

    def forces(self, u, v, p, mu):
        vel = nat.f64(self.vel_from(u, v))
        pp = nat.f64(np.ascontiguousarray(p, dtype=float).ravel())
        out = np.zeros(6)
        self.h.call('mcf_forces', nat.dptr(vel), nat.dptr(pp), float(mu), nat.dptr(out))
        return tuple(out)

    def rect_inverse(self, kind, rhs):
        rhs = nat.f64(rhs)
        out = np.empty_like(rhs)
        self.h.call('mcf_rect_inverse', kind, nat.dptr(rhs), nat.dptr(out))
        return out

    def schur_solve(self, kind, rhs):
        rhs = nat.f64(rhs)
        out = np.empty_like(rhs)
        info = np.zeros(4)
        self.h.call('mcf_schur_solve', kind, nat.dptr(rhs), nat.dptr(out), nat.dptr(info))
        return out, info

    def divergence_stats(self, u, v, mask):
        vel = nat.f64(self.vel_from(u, v))
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        out = np.zeros(5)
        self.h.call('mcf_divergence_stats', nat.dptr(vel), m.ctypes.data_as(nat._U8), nat.dptr(out))
        return out[0], out[1], int(round(out[2])), out[3], int(round(out[4]))

    def profile(self, n, dt):
        self.set_dt(dt)
        out = np.zeros(5)
        self.h.call('mcf_profile', int(n), float(dt), nat.dptr(out))
        return out

    # ---------------------------------------------------------------- launch shapes
    # ---------------------------------------------------------------- pressure backend
    def set_pressure_backend(self, iterative, restart=50, max_cycles=100, rtol=1e-12, atol=1e-12):
        """Select the pressure backend (pressure_schur.py:16-31): the boundary-row corrected rectangle
        inverse, or restarted GMRES on the pressure operator left-preconditioned by the rectangle
        inverse (the reference's arguments: restart 50, 100 cycles, rtol = atol = 1e-12)."""
        if not hasattr(self, '_direct_info'):
            self._direct_info = dict(self.pressure_info)
        self.h.call('mcf_set_pressure_backend', nat.PB_GMRES if iterative else nat.PB_DIRECT, int(restart),
                    int(max_cycles), float(rtol), float(atol))
        # updated in place: the integrator's pressure_inverse_info is this same dict
        self.pressure_info.clear()
        if iterative:
            self.pressure_info.update(backend='gmres_rectangle', preconditioner='separable_rectangle',
                                      krylov_rtol=float(rtol), krylov_atol=float(atol), restart=int(restart),
                                      max_cycles=int(max_cycles))
        else:
            self.pressure_info.update(self._direct_info)

    @property
    def pressure_backend(self):
        out = np.zeros(1, np.int32)
        self.h.call('mcf_get_pressure_backend', out)
        return int(out[0])

    def pressure_solve(self, values):
        """One monitored pressure solve (the reference's ``pressure_solve(values)``,
        pressure_schur.py:33-71) of the values on the active cells, in ``pi`` order.  Returns the
        solution; the solve's record is ``pressure_info['last_solve']``, also when the solve fails."""
        o = self.ops
        values = np.asarray(values, dtype=float)
        if values.shape != (len(o.pi),):
            raise ValueError('Pressure RHS must hold one value per active cell')
        out = np.zeros(o.nx * o.ny)
        rec = np.zeros(nat.SOLVE_LEN)
        try:
            self.h.call('mcf_pressure_solve', self.q_to_p(values), out, rec)
        finally:
            if rec.any():
                self.pressure_info['last_solve'] = solve_record(rec)
        return self.p_to_q(out)

    def set_launch(self, gemm=-1, tri=-1):
        """Pin the transform tile / sweep width variants (-1, -1: autotune at the next run)."""
        self.h.call('mcf_set_launch', int(gemm), int(tri))

    def set_launch_classes(self, launch):
        """Restore the per-class shapes of a ``get_launch`` dict (a run's autotuned choice)."""
        self.h.call('mcf_set_launch_classes', np.array([launch['gemm_pressure'], launch['gemm_velocity'],
                                                          launch['tri_pressure'], launch['tri_velocity']], dtype=np.int32))

    def get_launch(self):
        out = np.zeros(5, dtype=np.int32)
        self.h.call('mcf_get_launch', out)
        return dict(gemm_pressure=int(out[0]), gemm_velocity=int(out[1]), tri_pressure=int(out[2]),
                    tri_velocity=int(out[3]), chosen=bool(out[4]))

    def set_speculation(self, mode=1):
        """Steady steps without the refinement launches, replayed in full when one would refine
        (1, default); always launch them (0); flag every lean step to exercise the replay (2)."""
        self.h.call('mcf_set_speculation', int(mode))

    def get_speculation(self):
        out = np.zeros(2, dtype=np.int32)
        self.h.call('mcf_get_speculation', out)
        return dict(mode=int(out[0]), replays=int(out[1]))

    # ---------------------------------------------------------------- Schur system
    def set_dump(self, capacity):
        """Record the (c, y) pair of each projection's first pressure solve (angle order)."""
        self.dump_capacity = int(capacity)
        self.h.call('mcf_set_dump', self.dump_capacity)

    def get_dump(self):
        r = len(self.schur_geometry['order'])
        cap = getattr(self, 'dump_capacity', 0)
        c = np.empty((max(cap, 1), r)); y = np.empty_like(c)
        tags = np.empty(max(cap, 1), dtype=np.int32); n = np.zeros(1, dtype=np.int32)
        self.h.call('mcf_get_dump', cap, c, y, tags, n)
        n = int(n[0])
        return c[:n].copy(), y[:n].copy(), tags[:n].copy()

    def schur_gmres(self, rhs, x0=None, rtol=1e-12, max_iter=400, restart=0):
        """GMRES on the pressure Schur system for the rows of rhs (angle order)."""
        rhs = nat.f64(np.atleast_2d(rhs))
        n = rhs.shape[0]
        y = np.empty_like(rhs)
        its = np.empty(n, dtype=np.int32)
        res = np.empty(n)
        self.h.call('mcf_schur_gmres', nat.PRESSURE, n, rhs, None if x0 is None else nat.f64(np.atleast_2d(x0)),
                    float(rtol), int(max_iter), int(restart), y, its, res)
        return y, its, res
