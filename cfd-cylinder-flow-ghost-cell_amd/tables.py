"""Host-side geometry, ghost-stencil and operator tables for the ghost-cell cylinder step.

Everything here runs once per case (and once per distinct dt for the Helmholtz
Schur factors); nothing here runs per step.  The tables feed the HIP library
through plain arrays.  Integer decisions (ghost nodes, probe cells, stencil
growth, neighbour sets, Schur rows) are evaluated with the same IEEE
expressions as the reference so that the index tables are the reference's:

  * ghost extension        solver/optimized_ghost_cell.py:14-58 (vectorised over ghost nodes)
  * operator assembly      solver/optimized_ghost_cell.py:62-128
  * surface quadrature     solver/cut_cell_solver.py:25-77
  * wall-gradient stencils solver/optimized_ghost_cell.py:152-172
  * rectangle operators    solver/rectangular_inverse.py:8-46
  * boundary-row rows      solver/pressure_schur.py:78-92
"""
from types import SimpleNamespace

import numpy as np
import scipy.sparse as sp
import scipy.linalg as sla

EPS = np.finfo(float).eps


class CircleSDF:
    """Exact circle signed distance, negative inside (solver/ghost_cell.py:12-30)."""

    def __init__(self, xc, yc, radius):
        self.xc, self.yc, self.radius = xc, yc, radius

    def distance(self, x, y):
        r = np.sqrt((x - self.xc)**2 + (y - self.yc)**2)
        return r - self.radius

    def normal(self, x, y):
        r = np.sqrt((x - self.xc)**2 + (y - self.yc)**2)
        r_safe = np.maximum(r, 1e-14)
        return (x - self.xc) / r_safe, (y - self.yc) / r_safe


class GhostCellGeometry:
    """Staggered node coordinates and SDF fields (solver/ghost_cell.py:41-62)."""

    def __init__(self, nx, ny, dx, dy, sdf):
        self.nx, self.ny, self.dx, self.dy, self.sdf = nx, ny, dx, dy, sdf
        self.xu = np.arange(nx + 1) * dx
        self.yu = (np.arange(ny) + 0.5) * dy
        self.xv = (np.arange(nx) + 0.5) * dx
        self.yv = np.arange(ny + 1) * dy
        self.xp, self.yp = self.xv, self.yu
        self.dist_u = sdf.distance(*np.meshgrid(self.xu, self.yu))
        self.dist_v = sdf.distance(*np.meshgrid(self.xv, self.yv))
        self.dist_p = sdf.distance(*np.meshgrid(self.xp, self.yp))


# ------------------------------------------------------------------------------------------
# Ghost extension, vectorised over the ghost nodes (optimized_ghost_cell.py:14-58)
# ------------------------------------------------------------------------------------------
def extension(x, y, sdf, active, *, neumann=False, interpolation='bilinear'):
    """Return ``(E, ids, info)``; ``E`` maps active values to the full node array.

    All ghost nodes are advanced together: every pending node evaluates its two
    probes at s and 2s; nodes whose 2x2 (or 3x3) probe cells are all active are
    finalised, the rest grow s by 1.25 (at most 8 tries, as the reference).
    """
    if interpolation not in ('bilinear', 'biquadratic'):
        raise ValueError('Unknown ghost probe interpolation')
    xx, yy = np.meshgrid(x, y)
    distance = sdf.distance(xx, yy)
    ids = np.flatnonzero(active)
    mapping = np.full(active.shape, -1, dtype=np.int64)
    mapping.ravel()[ids] = np.arange(len(ids))
    h = max(x[1] - x[0], y[1] - y[0])
    jj, ii = np.nonzero((distance <= 1e-13) & (distance >= -2 * h))
    ng = len(jj)
    quadratic = interpolation == 'biquadratic'
    size = 3 if quadratic else 2
    d = np.maximum(0., -distance[jj, ii])
    nxg, nyg = sdf.normal(x[ii], y[jj])
    s = np.maximum(d, 1.5 * h)
    final_s = np.full(ng, np.nan)
    tries = np.full(ng, -1)
    final_idx = np.zeros((ng, 2, size * size), dtype=np.int64)
    final_coeff = np.zeros((ng, 2, size * size))
    pending = np.arange(ng)
    dj, di = np.divmod(np.arange(size * size), size)
    for attempt in range(8):
        if not len(pending):
            break
        sp_ = s[pending]
        idx = np.zeros((len(pending), 2, size * size), dtype=np.int64)
        coeff = np.zeros((len(pending), 2, size * size))
        for k, ell in enumerate((sp_, 2 * sp_)):
            px = x[ii[pending]] + (d[pending] + ell) * nxg[pending]
            py = y[jj[pending]] + (d[pending] + ell) * nyg[pending]
            fi, fj = (px - x[0]) / (x[1] - x[0]), (py - y[0]) / (y[1] - y[0])
            if quadratic:
                i0 = np.floor(fi + .5).astype(np.int64) - 1
                j0 = np.floor(fj + .5).astype(np.int64) - 1
            else:
                i0 = np.floor(fi).astype(np.int64)
                j0 = np.floor(fj).astype(np.int64)
            if np.any((i0 < 0) | (i0 > len(x) - size) | (j0 < 0) | (j0 > len(y) - size)):
                raise ValueError('Ghost probe reaches an outer boundary')
            idx[:, k] = mapping[j0[:, None] + dj[None, :], i0[:, None] + di[None, :]]
            a, b = fi - i0, fj - j0
            if quadratic:
                wx = np.stack((.5 * (a - 1) * (a - 2), -a * (a - 2), .5 * a * (a - 1)), axis=1)
                wy = np.stack((.5 * (b - 1) * (b - 2), -b * (b - 2), .5 * b * (b - 1)), axis=1)
                coeff[:, k] = (wy[:, :, None] * wx[:, None, :]).reshape(len(pending), 9)
            else:
                coeff[:, k] = np.stack(((1 - a) * (1 - b), a * (1 - b), (1 - a) * b, a * b), axis=1)
        ok = np.all(idx >= 0, axis=(1, 2))
        done = pending[ok]
        final_s[done] = s[done]
        tries[done] = attempt
        final_idx[done] = idx[ok]
        final_coeff[done] = coeff[ok]
        s[pending[~ok]] *= 1.25
        pending = pending[~ok]
    if len(pending):
        raise ValueError('No exterior ghost stencil')
    sg, dg = final_s, d
    if neumann:
        w1 = (4 * sg * sg - dg * dg) / (3 * sg * sg)
        w2 = (dg * dg - sg * sg) / (3 * sg * sg)
    else:
        w1 = -dg * (dg + 2 * sg) / (sg * sg)
        w2 = dg * (dg + sg) / (2 * sg * sg)
    values = np.concatenate((w1[:, None] * final_coeff[:, 0], w2[:, None] * final_coeff[:, 1]), axis=1)
    node = jj * len(x) + ii
    rows = np.concatenate((ids, np.repeat(node, 2 * size * size)))
    cols = np.concatenate((np.arange(len(ids)), final_idx.reshape(ng, -1).ravel()))
    vals = np.concatenate((np.ones(len(ids)), values.ravel()))
    matrix = sp.csr_matrix((vals, (rows, cols)), shape=(active.size, len(ids)))
    info = dict(ghost_nodes=node, s=final_s, d=d, tries=tries)
    return matrix, ids, info


# ------------------------------------------------------------------------------------------
# Surface quadrature and wall-gradient stencils
# ------------------------------------------------------------------------------------------
def nearest_with_ties(centers, points, k):
    """For each point: indices of all centres within d_k (1+1e-12), ascending index.

    Replaces the reference's cKDTree ``query`` + ``query_ball_point`` pair
    (cut_cell_solver.py:53-55; optimized_ghost_cell.py:158-160).
    """
    out = []
    for p in points:
        diff = centers - p
        dist = np.sqrt(np.sum(diff * diff, axis=1))
        dk = np.partition(dist, k - 1)[k - 1]
        out.append(np.flatnonzero(dist <= dk * (1 + 1e-12)))
    return out


def exterior_pressure_extension(x, y, sdf, active):
    """Quadratic continuation of the pressure into the ghost cells (ghost_reconstruction.py:7-45).

    Each ghost cell (-2h <= d <= 1e-13) takes the least-squares quadratic through its 12 nearest
    active centres within 6h of the surface (ties kept), else 20, else 32, the first fit of rank
    6 with condition <= 1e8 that reproduces the quadratic basis to 1e-10 after the largest weight
    absorbs the partition-of-unity defect.  Returns (P, ids, info) like ``extension`` (identity
    rows on the active cells), with the fit's worst condition, reproduction error and stencil size.
    """
    xx, yy = np.meshgrid(x, y)
    dist = sdf.distance(xx, yy)
    dx, dy = x[1] - x[0], y[1] - y[0]
    h = max(dx, dy)
    ids = np.flatnonzero(active)
    label = np.full(active.size, -1, dtype=np.int64)
    label[ids] = np.arange(len(ids))
    near_set = np.flatnonzero(active.ravel() & (dist.ravel() < 6 * h))
    if len(near_set) < 12:
        raise ValueError('Pressure continuation needs twelve exterior centers')
    ghosts = np.flatnonzero((dist.ravel() <= 1e-13) & (dist.ravel() >= -2 * h))
    centers = np.column_stack((yy.ravel()[near_set], xx.ravel()[near_set]))
    target = np.array([1., 0., 0., 0., 0., 0.])
    rows, cols, vals = [ids], [np.arange(len(ids))], [np.ones(len(ids))]
    worst_cond, worst_err, widest = 0.0, 0.0, 0
    for g in ghosts:
        point = np.array([yy.ravel()[g], xx.ravel()[g]])
        for count in (12, 20, 32):
            near = nearest_with_ties(centers, point[None], min(count, len(centers)))[0]
            sy, sx = ((centers[near] - point) / [dy, dx]).T
            basis = np.array([np.ones(len(near)), sx, sy, sx * sx, sx * sy, sy * sy])
            w, _, rank, _ = np.linalg.lstsq(basis, target, rcond=1e-12)
            cond = float(np.linalg.cond(basis))
            w[np.argmax(abs(w))] += 1 - w.sum()
            err = float(np.max(abs(basis @ w - target)))
            if rank == 6 and np.isfinite(cond) and cond <= 1e8 and err <= 1e-10:
                break
        else:
            raise ValueError('Unresolved exterior pressure polynomial')
        rows.append(np.full(len(near), g)); cols.append(label[near_set[near]]); vals.append(w)
        worst_cond, worst_err, widest = max(worst_cond, cond), max(worst_err, err), max(widest, len(near))
    P = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(active.size, len(ids)))
    return P, ids, dict(ghost_nodes=ghosts, stencil_condition_max=worst_cond, polynomial_reproduction_linf=worst_err,
                        neighbor_count_max=widest)


def pressure_reconstruction(geom, points):
    """Pressure values at points (y, x) from the 12 nearest exterior cell centres (ties kept) by a
    least-squares quadratic (cut_cell_solver.py:25-67): CSR (len(points) x nx*ny)."""
    points = np.asarray(points, dtype=float)
    padding = 4 * max(geom.dx, geom.dy)
    lo = np.maximum(0, np.floor((points.min(axis=0) - padding) / [geom.dy, geom.dx]).astype(int))
    hi = np.minimum([geom.ny, geom.nx], np.ceil((points.max(axis=0) + padding) / [geom.dy, geom.dx]).astype(int) + 1)
    jj, ii = np.meshgrid(np.arange(lo[0], hi[0]), np.arange(lo[1], hi[1]), indexing='ij')
    y, x = (jj + .5) * geom.dy, (ii + .5) * geom.dx
    exterior = (geom.cell_type[jj, ii] != -1) & (np.hypot(x - geom.xc, y - geom.yc) >= geom.R * (1 - 1e-12))
    indices = (jj * geom.nx + ii)[exterior]
    if len(indices) < 12:
        raise ValueError('Pressure reconstruction needs at least 12 exterior centers')
    centers = np.column_stack((y[exterior], x[exterior]))
    rows, cols, vals = [], [], []
    for k, near in enumerate(nearest_with_ties(centers, points, 12)):
        dyy, dxx = ((centers[near] - points[k]) / [geom.dy, geom.dx]).T
        basis = np.array([np.ones(len(near)), dxx, dyy, dxx * dxx, dxx * dyy, dyy * dyy])
        w, _, rank, _ = np.linalg.lstsq(basis, [1., 0., 0., 0., 0., 0.], rcond=1e-12)
        if rank != 6:
            raise ValueError('Pressure stencil cannot resolve a quadratic polynomial')
        rows.extend([k] * len(near)); cols.extend(indices[near]); vals.extend(w)
    return sp.csr_matrix((vals, (rows, cols)), shape=(len(points), geom.nx * geom.ny))


def surface_pressure_quadrature(geom):
    """Points, pressure stencils, weighted normals (cut_cell_solver.py:70-77)."""
    count = max(360, 4 * int(np.ceil(2 * np.pi * geom.R / min(geom.dx, geom.dy))))
    theta = np.arange(count) * (2 * np.pi / count)
    normals = np.column_stack((np.cos(theta), np.sin(theta)))
    points = np.column_stack((geom.yc + geom.R * normals[:, 1], geom.xc + geom.R * normals[:, 0]))
    return points, pressure_reconstruction(geom, points), normals * (2 * np.pi * geom.R / count)


def wall_gradient(x, y, active, points, geom, h):
    """d/dn at the surface points from 14 active neighbours (optimized_ghost_cell.py:152-172)."""
    xx, yy = np.meshgrid(x, y)
    radius = np.hypot(xx - geom.xc, yy - geom.yc)
    ids = np.flatnonzero(active & (radius < geom.R + 6 * h))
    coords = np.column_stack((yy.ravel()[ids], xx.ravel()[ids]))
    rows, cols, vals = [], [], []
    for k, near in enumerate(nearest_with_ties(coords, points, 14)):
        ys, xs = coords[near].T
        theta = np.arctan2(points[k, 0] - geom.yc, points[k, 1] - geom.xc)
        angle = np.arctan2(ys - geom.yc, xs - geom.xc) - theta
        n = (np.hypot(xs - geom.xc, ys - geom.yc) - geom.R) / h
        t = geom.R * np.arctan2(np.sin(angle), np.cos(angle)) / h
        w, _, rank, _ = np.linalg.lstsq(np.array([n, n * n, n * t]), [1 / h, 0., 0.], rcond=1e-12)
        if rank != 3:
            raise ValueError('Unresolved wall-gradient stencil')
        rows.extend([k] * len(near)); cols.extend(ids[near]); vals.extend(w)
    return sp.csr_matrix((vals, (rows, cols)), shape=(len(points), active.size))


# ------------------------------------------------------------------------------------------
# Operators of the ghost integrator (optimized_ghost_cell.py:62-128)
# ------------------------------------------------------------------------------------------
class GhostOperators:
    """E, P, D, G, A, L and offsets exactly as assembled by the reference integrator."""

    def __init__(self, nx, ny, Lx, Ly, D, xc, yc, nu, rho, u_inlet, interpolation='bilinear',
                 pressure_extension='neumann'):
        dx, dy = Lx / nx, Ly / ny
        self.nx, self.ny, self.dx, self.dy, self.nu, self.rho = nx, ny, dx, dy, nu, rho
        self.Lx, self.Ly, self.D, self.xc, self.yc, self.R = Lx, Ly, D, xc, yc, D / 2
        gc = self.gc = GhostCellGeometry(nx, ny, dx, dy, CircleSDF(xc, yc, D / 2.0))
        active_u = gc.dist_u > 1e-13
        active_v = gc.dist_v > 1e-13
        active_u[:, 0] = False
        active_v[[0, -1]] = False
        active_p = gc.dist_p > 1e-13
        self.active_u, self.active_v, self.active_p = active_u, active_v, active_p
        self.Eu, self.ui, self.ghost_u = extension(gc.xu, gc.yu, gc.sdf, active_u, interpolation=interpolation)
        self.Ev, self.vi, self.ghost_v = extension(gc.xv, gc.yv, gc.sdf, active_v, interpolation=interpolation)
        if pressure_extension == 'quadratic':   # optimized_ghost_cell.py:75-79
            self.P, self.pi, self.ghost_p = exterior_pressure_extension(gc.xp, gc.yp, gc.sdf, active_p)
        elif pressure_extension == 'neumann':
            self.P, self.pi, self.ghost_p = extension(gc.xp, gc.yp, gc.sdf, active_p, neumann=True)
        else:
            raise ValueError('Unknown Ghost pressure extension')
        self.E = sp.block_diag((self.Eu, self.Ev), format='csr')
        self.ushape, self.vshape, self.pshape = (ny, nx + 1), (ny + 1, nx), (ny, nx)
        self.usize, self.vsize = ny * (nx + 1), (ny + 1) * nx
        self.nu_act = len(self.ui)
        self.qsize = len(self.ui) + len(self.vi)
        self.offset = np.zeros(self.usize + self.vsize)
        self.offset[:self.usize].reshape(self.ushape)[:, 0] = u_inlet
        j, i = np.unravel_index(self.pi, self.pshape)
        r = np.arange(len(self.pi))
        uf, vf = j * (nx + 1) + i, self.usize + j * nx + i
        self.Dfull = sp.csr_matrix(
            (np.concatenate((-np.ones(len(r)) / dx, np.ones(len(r)) / dx, -np.ones(len(r)) / dy, np.ones(len(r)) / dy)),
             (np.tile(r, 4), np.concatenate((uf, uf + 1, vf, vf + nx)))), shape=(len(r), self.E.shape[0]))
        self.Dop = (self.Dfull @ self.E).tocsr()
        self.div_offset = self.Dfull @ self.offset
        ju, iu = np.unravel_index(self.ui, self.ushape)
        jv, iv = np.unravel_index(self.vi, self.vshape)
        self.outlet_u = np.flatnonzero(iu == nx)
        ru = np.arange(len(self.ui)); interior = iu < nx
        rv = len(ru) + np.arange(len(self.vi))
        raw_g = sp.csr_matrix(
            (np.concatenate((np.full(interior.sum(), 1 / dx), np.full(interior.sum(), -1 / dx),
                             np.full((~interior).sum(), -2 / dx), np.full(len(rv), 1 / dy), np.full(len(rv), -1 / dy))),
             (np.concatenate((ru[interior], ru[interior], ru[~interior], rv, rv)),
              np.concatenate((ju[interior] * nx + iu[interior], ju[interior] * nx + iu[interior] - 1,
                              ju[~interior] * nx + nx - 1, jv * nx + iv, (jv - 1) * nx + iv)))),
            shape=(self.qsize, nx * ny))
        self.raw_g = raw_g
        self.G = (raw_g @ self.P / rho).tocsr()
        self.A = (self.Dop @ self.G).tocsc()
        rows, cols, vals = [], [], []

        def add(row, col, value):
            rows.append(row); cols.append(col); vals.append(np.broadcast_to(value, row.shape))
        add(ru, self.ui, -2 / dx**2 - np.where((ju == 0) | (ju == ny - 1), 3., 2.) / dy**2)
        add(ru, self.ui - 1, np.where(iu == nx, 2., 1.) / dx**2)
        add(ru[interior], self.ui[interior] + 1, 1 / dx**2)
        for shift, keep in ((nx + 1, ju < ny - 1), (-nx - 1, ju > 0)):
            add(ru[keep], self.ui[keep] + shift, 1 / dy**2)
        add(rv, self.usize + self.vi, -np.where(iv == 0, 3., np.where(iv == nx - 1, 1., 2.)) / dx**2 - 2 / dy**2)
        for shift, keep in ((1, iv < nx - 1), (-1, iv > 0), (nx, np.ones(len(rv), bool)), (-nx, np.ones(len(rv), bool))):
            add(rv[keep], self.usize + self.vi[keep] + shift, 1 / (dx**2 if abs(shift) == 1 else dy**2))
        self.raw_l = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                                   shape=(self.qsize, self.E.shape[0]))
        self.L = (nu * self.raw_l @ self.E).tocsr()
        self.diff_offset = nu * (self.raw_l @ self.offset)
        self.geom = SimpleNamespace(nx=nx, ny=ny, dx=dx, dy=dy, Lx=Lx, Ly=Ly, xc=xc, yc=yc, R=D / 2,
                                    cell_type=np.where(active_p, 1, -1))
        self.surface_points, self.pressure_surface, self.normal_weights = surface_pressure_quadrature(self.geom)
        self.normals = self.normal_weights / (np.pi * D / len(self.surface_points))
        h = max(dx, dy)
        self.du_dn = wall_gradient(gc.xu, gc.yu, active_u, self.surface_points, self.geom, h)
        self.dv_dn = wall_gradient(gc.xv, gc.yv, active_v, self.surface_points, self.geom, h)

    def helmholtz_actual(self, dt):
        """``I - dt/2 L`` on the u and v blocks and their rectangle indices (optimized_ghost_cell.py:247-252)."""
        n = self.nu_act
        ju, iu = np.unravel_index(self.ui, self.ushape)
        jv, iv = np.unravel_index(self.vi, self.vshape)
        out = []
        for kind, selected, block in (('u', ju * self.nx + iu - 1, self.L[:n, :n]),
                                      ('v', (jv - 1) * self.nx + iv, self.L[n:, n:])):
            actual = (sp.eye(block.shape[0], format='csc') - .5 * dt * block).tocsc()
            out.append((kind, selected, actual))
        return out

    # assembled operators and offsets: needed to build the device tables, not to step or index states
    OPERATOR_ATTRIBUTES = ('raw_l', 'L', 'A', 'G', 'raw_g', 'Dfull', 'Dop', 'E', 'Eu', 'Ev', 'P', 'offset',
                           'diff_offset', 'div_offset', 'du_dn', 'dv_dn', 'pressure_surface')

    def layout(self):
        """A copy holding the geometry and index maps only (the assembled operators are None)."""
        lay = object.__new__(GhostOperators)
        lay.__dict__.update(self.__dict__)
        for name in self.OPERATOR_ATTRIBUTES:
            lay.__dict__[name] = None
        return lay


# ------------------------------------------------------------------------------------------
# Rectangle operators and their y-eigenbasis (rectangular_inverse.py:8-46)
# ------------------------------------------------------------------------------------------
RECT_ENDS = {'pressure': (1., 3., 1., 1.), 'u': (2., 2., 3., 3.), 'v': (3., 1., 2., 2.)}


def rectangle_shape(kind, nx, ny):
    return (ny - 1, nx) if kind == 'v' else (ny, nx)


def rectangle_matrix(kind, nx, ny, dx, dy, nu=0., dt=0., rho=1.):
    """Sparse cylinder-free operator of rectangular_inverse.py:32-39, 43-46."""
    yy, xx = rectangle_shape(kind, nx, ny)
    left, right, lower, upper = RECT_ENDS[kind]
    diagx = np.full(xx, -2.); diagx[0] = -left; diagx[-1] = -right
    diagy = np.full(yy, -2.); diagy[0] = -lower; diagy[-1] = -upper
    lowerx = np.ones(xx - 1)
    if kind == 'u':
        lowerx[-1] = 2.
    lx = sp.diags((lowerx, diagx, np.ones(xx - 1)), (-1, 0, 1), format='csr') / dx**2
    ly = sp.diags((np.ones(yy - 1), diagy, np.ones(yy - 1)), (-1, 0, 1), format='csr') / dy**2
    lap = sp.kron(sp.eye(yy), lx) + sp.kron(ly, sp.eye(xx))
    if kind == 'pressure':
        return (lap / rho).tocsc()
    return (sp.eye(xx * yy) - .5 * dt * nu * lap).tocsc()


def rectangle_coefficients(kind, nu=0., dt=0., rho=1.):
    """(a, b) with operator = a I + b (lx + ly)."""
    if kind == 'pressure':
        return 0., 1. / rho
    return 1., -.5 * dt * nu


def y_eigenbasis(kind, ny, dy):
    """Orthonormal eigenvectors and eigenvalues of the 1-D y operator of ``kind``.

    Eigenvalues use the reference's closed form -4 sin^2(s_y)/dy^2
    (rectangular_inverse.py:11-28, 40); eigenvectors are the matching DCT-II,
    DST-II or DST-I columns, normalised numerically.
    """
    if kind == 'pressure':
        n = ny; sy = np.arange(ny) * np.pi / (2 * ny)
        j = np.arange(n) + .5
        V = np.cos(np.outer(j, 2 * sy))
    elif kind == 'u':
        n = ny; sy = (np.arange(ny) + 1) * np.pi / (2 * ny)
        j = np.arange(n) + .5
        V = np.sin(np.outer(j, 2 * sy))
    elif kind == 'v':
        n = ny - 1; sy = (np.arange(ny - 1) + 1) * np.pi / (2 * ny)
        j = np.arange(n) + 1.
        V = np.sin(np.outer(j, 2 * sy))
    else:
        raise ValueError('Unknown rectangular MAC field')
    V /= np.sqrt(np.sum(V * V, axis=0))[None, :]
    lam = -4 * np.sin(sy)**2 / dy**2
    return V, lam


def x_tridiagonal(kind, xx, dx):
    """(lower, diag, upper) of the 1-D x operator of ``kind`` (rectangular_inverse.py:32-37)."""
    left, right, _, _ = RECT_ENDS[kind]
    diag = np.full(xx, -2.); diag[0] = -left; diag[-1] = -right
    lower = np.ones(xx - 1)
    if kind == 'u':
        lower[-1] = 2.
    return lower / dx**2, diag / dx**2, np.ones(xx - 1) / dx**2


class RectangleFactor:
    """Separable inverse data: y-eigenbasis and per-mode tridiagonal x systems."""

    def __init__(self, kind, nx, ny, dx, dy, nu=0., dt=0., rho=1.):
        self.kind = kind
        self.shape = rectangle_shape(kind, nx, ny)
        self.a, self.b = rectangle_coefficients(kind, nu, dt, rho)
        self.V, self.lam = y_eigenbasis(kind, ny, dy)
        self.lower, self.diag, self.upper = x_tridiagonal(kind, self.shape[1], dx)
        self.dx = dx
        self.matrix = rectangle_matrix(kind, nx, ny, dx, dy, nu, dt, rho)

    def banded(self, k):
        xx = self.shape[1]
        ab = np.zeros((3, xx))
        ab[0, 1:] = self.b * self.upper
        ab[1] = self.a + self.b * self.lam[k] + self.b * self.diag
        ab[2, :-1] = self.b * self.lower
        return ab

    def solve(self, rhs):
        """Host inverse (eigenbasis along y, banded solves along x); setup use only."""
        g = self.V.T @ rhs
        out = np.empty_like(g)
        for k in range(self.shape[0]):
            out[k] = sla.solve_banded((1, 1), self.banded(k), g[k])
        return self.V @ out

    def columns_at(self, sources, targets):
        """Entries R^{-1}[targets, sources] for flat rectangle indices (setup only)."""
        yy, xx = self.shape
        js, is_ = np.divmod(np.asarray(sources), xx)
        jt, it = np.divmod(np.asarray(targets), xx)
        src_cols, src_inv = np.unique(is_, return_inverse=True)
        tgt_cols, tgt_inv = np.unique(it, return_inverse=True)
        rhs = np.zeros((xx, len(src_cols)))
        rhs[src_cols, np.arange(len(src_cols))] = 1.
        Z = np.empty((yy, len(tgt_cols), len(src_cols)))
        for k in range(yy):
            Z[k] = sla.solve_banded((1, 1), self.banded(k), rhs)[tgt_cols]
        # R^{-1}[t, s] = sum_k V[jt,k] V[js,k] Z[k, it, is]
        out = np.empty((len(targets), len(sources)))
        Vt, Vs = self.V[jt], self.V[js]
        for a in range(len(sources)):
            out[:, a] = np.einsum('tk,k,kt->t', Vt, Vs[a], Z[:, tgt_inv, src_inv[a]])
        return out


def boundary_rows(actual, indices, rect):
    """Rows where the embedded operator differs from the rectangle (pressure_schur.py:78-92)."""
    indices = np.asarray(indices)
    covered = np.ones(rect.shape[0], bool)
    covered[indices] = False
    coo = actual.tocoo()
    extended = sp.csr_matrix((coo.data, (indices[coo.row], indices[coo.col])), shape=rect.shape)
    virtual = rect.tocoo()
    keep = covered[virtual.row] & covered[virtual.col]
    extended += sp.csr_matrix((virtual.data[keep], (virtual.row[keep], virtual.col[keep])), shape=rect.shape)
    delta = (extended - rect).tocsr()
    tolerance = 8 * EPS * max(1., float(np.max(abs(rect.diagonal()))))
    delta.data[abs(delta.data) < tolerance] = 0.
    delta.eliminate_zeros()
    rows = np.flatnonzero(np.diff(delta.indptr))
    return rows, delta[rows].tocsr(), tolerance


class SchurFactor:
    """Capacitance data of the boundary-row correction (pressure_schur.py:75-111).

    ``S = I + C R^{-1} E_rows`` is formed on the host from the rectangle's
    eigenbasis and inverted once; the device applies ``S^{-1}`` as a dense
    matrix-vector product each solve.
    """

    def __init__(self, actual, indices, factor):
        self.indices = np.asarray(indices)
        self.rows, self.C, self.tolerance = boundary_rows(actual, self.indices, factor.matrix)
        if len(self.rows) > 2048:
            raise ValueError('Ghost inverse exceeds the bounded boundary-row rank')
        self.rank = len(self.rows)
        C = self.C.tocsc()
        self.points = np.flatnonzero(np.diff(C.indptr))          # columns of C that are used
        Cp = self.C[:, self.points].tocsr()
        W = factor.columns_at(self.rows, self.points)             # R^{-1}[points, rows]
        self.S = np.eye(self.rank) + Cp @ W
        self.Sinv = np.linalg.inv(self.S)
        self.Cp = Cp                                              # C restricted to the points
        self.SC = np.ascontiguousarray((Cp.T @ self.Sinv.T).T)   # S^{-1} C, dense r x npts
