"""CPU oracle for the ghost-cell AB2/CN cylinder step (FP64, NumPy/SciPy).

TEST INFRASTRUCTURE ONLY.  This module is a plain restatement of the reference
algorithm and is used as the checker: tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it.  The product package
never imports, links or executes anything under ``oracle/``.

Every function names the reference file:line it follows (paths are relative to
the reference repository root).  The arithmetic follows the reference's
expression order wherever a value feeds an integer decision (ghost-node
selection, probe cells, stencil growth, nearest-neighbour sets) so that the
index tables it produces are the reference's.

Pins (each test lives in tests/test_oracle_*.py):
  * CircleSDF                -- exact values (tests/test_ghost_cell_sharp.py:9-27)
  * extension                -- flat-wall quadratic (tests/test_gemini_followup.py:29-44),
                                exact linear / Neumann reproduction on a flat wall
  * rectangle                -- residual self-check (rectangular_inverse.py:54-57),
                                dense solve on small grids, independent stencil assembly
  * boundary_row_inverse     -- dense solve of ``actual`` on small grids, Woodbury residual
  * integrator operators     -- L against the flux-form diffusion
                                (tests/test_numba_ghost_cell.py:22-29), A phi = div/dt
  * ab2_cn_step              -- closed-form constrained ODE
                                (tests/test_cut_cell_time_restart.py:65-86), time order on
                                the 88x16 channel (tests/test_numba_ghost_cell.py:37-46)
  * surface_forces           -- p = x gives F_x,p = -pi R^2; torque converges to 2 pi nu R^2
                                (tests/test_numba_ghost_cell.py:49-69)
  * divergence reduction     -- the reference C++ ``calculate_mac_divergence_cpp`` built
                                from solver/src/bindings.cpp into oracle/_ref
  * checkpoint               -- restart parity (tests/test_numba_ghost_cell.py:87-106)
  * index tables and values  -- golden vectors written by the reference's own functions
                                (tests/golden/gen_reference_tables.py -> reference_tables.npz) on the
                                88x16, 40x30, 440x82, 1320x246 and 800x820 grids: active and ghost
                                node sets, probe cells, stencil growth (biquadratic), weights, the
                                kNN-with-ties surface stencils, Schur ranks, rectangle / Woodbury /
                                Helmholtz values (tests/test_reference_golden.py)
  * knn_with_ties            -- scipy cKDTree query + query_ball_point, the reference's calls
  * advection                -- closed forms for linear u, v: interior faces, outlet row, wall row
  * OpenMPKernels            -- the C++/OpenMP flux loops and CSR products of the CPU baseline
                                (oracle/cpu_kernels.cpp) equal flux_rhs and SciPy's products bit
                                for bit, and whole steps with them equal the NumPy steps
  * extra_refinements        -- (a study option, not the reference's rule) residual stays at its
                                round-off floor; its force moves set the long-run force bars

Options beyond the reference: ``openmp_threads`` (the CPU baseline's OpenMP loops) and
``extra_refinements`` (refinements past the reference's stopping rule, to measure how far that rule
determines a step) of ``CylinderFlow``.
"""
import json
import os
import time
from pathlib import Path
import tempfile
import warnings

import numpy as np
import scipy.sparse as sp
from scipy.fft import dct, idct, dst, idst
from scipy.linalg import lu_factor, lu_solve, LinAlgWarning
from scipy.sparse.linalg import LinearOperator, gmres

EPS = np.finfo(float).eps


# --------------------------------------------------------------------------------------
# Geometry (solver/ghost_cell.py:12-62)
# --------------------------------------------------------------------------------------
class CircleSDF:
    """Signed distance to a circle, negative inside (solver/ghost_cell.py:12-30)."""

    def __init__(self, xc, yc, radius):
        self.xc, self.yc, self.radius = xc, yc, radius

    def distance(self, x, y):
        r = np.sqrt((x - self.xc)**2 + (y - self.yc)**2)
        return r - self.radius

    def normal(self, x, y):
        r = np.sqrt((x - self.xc)**2 + (y - self.yc)**2)
        r_safe = np.maximum(r, 1e-14)
        return (x - self.xc) / r_safe, (y - self.yc) / r_safe


class GhostGeometry:
    """Staggered coordinates and SDF fields (solver/ghost_cell.py:41-62)."""

    def __init__(self, nx, ny, dx, dy, sdf):
        self.nx, self.ny, self.dx, self.dy, self.sdf = nx, ny, dx, dy, sdf
        self.xu = np.arange(nx + 1) * dx
        self.yu = (np.arange(ny) + 0.5) * dy
        self.xv = (np.arange(nx) + 0.5) * dx
        self.yv = np.arange(ny + 1) * dy
        self.xp = self.xv
        self.yp = self.yu
        XU, YU = np.meshgrid(self.xu, self.yu)
        XV, YV = np.meshgrid(self.xv, self.yv)
        XP, YP = np.meshgrid(self.xp, self.yp)
        self.dist_u = sdf.distance(XU, YU)
        self.dist_v = sdf.distance(XV, YV)
        self.dist_p = sdf.distance(XP, YP)


# --------------------------------------------------------------------------------------
# Ghost extension (solver/optimized_ghost_cell.py:14-58)
# --------------------------------------------------------------------------------------
def extension(x, y, sdf, active, *, neumann=False, interpolation='bilinear', return_probes=False):
    """Quadratic-normal ghost extension as a CSR matrix (optimized_ghost_cell.py:14-58).

    Returns ``(E, ids)``: ``E`` maps the active values (row-major ``flatnonzero``
    order) to the full array; identity rows on active nodes, two-probe
    extrapolation rows on ghost nodes (-2h <= d <= 1e-13), zero rows elsewhere.
    With ``return_probes`` a list of per-ghost records (node, s, tries) is added.
    """
    if interpolation not in ('bilinear', 'biquadratic'):
        raise ValueError('Unknown ghost probe interpolation')
    xx, yy = np.meshgrid(x, y)
    distance = sdf.distance(xx, yy)
    ids = np.flatnonzero(active)
    mapping = np.full(active.shape, -1, dtype=int)
    mapping.ravel()[ids] = np.arange(len(ids))
    rows, cols, values = [ids], [np.arange(len(ids))], [np.ones(len(ids))]
    h = max(x[1] - x[0], y[1] - y[0])
    jj, ii = np.nonzero((distance <= 1e-13) & (distance >= -2 * h))
    probes = []
    quadratic = interpolation == 'biquadratic'
    size = 3 if quadratic else 2
    for j, i in zip(jj, ii):
        d = max(0., -distance[j, i])
        nx_, ny_ = sdf.normal(x[i], y[j])
        s = max(d, 1.5 * h)
        for attempt in range(8):
            stencils = []
            for ell in (s, 2 * s):
                px, py = x[i] + (d + ell) * nx_, y[j] + (d + ell) * ny_
                fi, fj = (px - x[0]) / (x[1] - x[0]), (py - y[0]) / (y[1] - y[0])
                if quadratic:
                    i0, j0 = int(np.floor(fi + .5)) - 1, int(np.floor(fj + .5)) - 1
                else:
                    i0, j0 = int(np.floor(fi)), int(np.floor(fj))
                if not (0 <= i0 <= len(x) - size and 0 <= j0 <= len(y) - size):
                    raise ValueError('Ghost probe reaches an outer boundary')
                indices = mapping[j0:j0 + size, i0:i0 + size].ravel()
                a, b = fi - i0, fj - j0
                if quadratic:
                    wx = np.array([.5 * (a - 1) * (a - 2), -a * (a - 2), .5 * a * (a - 1)])
                    wy = np.array([.5 * (b - 1) * (b - 2), -b * (b - 2), .5 * b * (b - 1)])
                    coeff = np.outer(wy, wx).ravel()
                else:
                    coeff = np.array([(1 - a) * (1 - b), a * (1 - b), (1 - a) * b, a * b])
                stencils.append((indices, coeff, i0, j0))
            if all((st[0] >= 0).all() for st in stencils):
                break
            s *= 1.25
        else:
            raise ValueError('No exterior ghost stencil')
        if neumann:
            weights = ((4 * s * s - d * d) / (3 * s * s), (d * d - s * s) / (3 * s * s))
        else:
            weights = (-d * (d + 2 * s) / (s * s), d * (d + s) / (2 * s * s))
        for factor, (indices, coeff, _, _) in zip(weights, stencils):
            rows.append(np.full(len(indices), j * len(x) + i))
            cols.append(indices)
            values.append(factor * coeff)
        probes.append(dict(node=int(j * len(x) + i), s=float(s), tries=attempt,
                           cells=[(st[3], st[2]) for st in stencils]))
    matrix = sp.csr_matrix((np.concatenate(values), (np.concatenate(rows), np.concatenate(cols))),
                           shape=(active.size, len(ids)))
    if return_probes:
        return matrix, ids, probes
    return matrix, ids


# --------------------------------------------------------------------------------------
# k-nearest neighbours with ties (replaces cKDTree query + query_ball_point)
# --------------------------------------------------------------------------------------
def exterior_pressure_extension(x, y, sdf, active):
    """Quadratic exterior pressure continuation (solver/ghost_reconstruction.py:7-45), one ghost cell
    at a time: 12, 20 then 32 nearest active centres within 6h (ties kept), least squares on the
    basis [1, x, y, x^2, xy, y^2], largest weight absorbs the partition-of-unity defect; accepted at
    rank 6, condition <= 1e8 and reproduction error <= 1e-10.  Returns (P, ids, info)."""
    xx, yy = np.meshgrid(x, y)
    distance = sdf.distance(xx, yy)
    dx, dy = x[1] - x[0], y[1] - y[0]
    h = max(dx, dy)
    flat_active = active.ravel()
    ids = np.flatnonzero(flat_active)
    labels = -np.ones(active.size, dtype=np.int64)
    labels[ids] = np.arange(len(ids))
    pool = [k for k in range(active.size) if flat_active[k] and distance.ravel()[k] < 6 * h]
    if len(pool) < 12:
        raise ValueError('Pressure continuation needs twelve exterior centers')
    pool = np.array(pool)
    centers = np.column_stack((yy.ravel()[pool], xx.ravel()[pool]))
    data = {}
    for k in ids:
        data[(k, labels[k])] = 1.0
    cond_max = err_max = 0.0
    count_max = 0
    for node in range(active.size):
        d = distance.ravel()[node]
        if not (-2 * h <= d <= 1e-13):
            continue
        point = np.array([yy.ravel()[node], xx.ravel()[node]])
        for count in (12, 20, 32):
            near = knn_with_ties(centers, point, min(count, len(centers)))
            rel_y = (centers[near, 0] - point[0]) / dy
            rel_x = (centers[near, 1] - point[1]) / dx
            basis = np.vstack((np.ones(len(near)), rel_x, rel_y, rel_x ** 2, rel_x * rel_y, rel_y ** 2))
            want = np.zeros(6)
            want[0] = 1.0
            weights = np.linalg.lstsq(basis, want, rcond=1e-12)
            w, rank = weights[0], weights[2]
            cond = float(np.linalg.cond(basis))
            big = int(np.argmax(np.abs(w)))
            w[big] += 1.0 - np.sum(w)
            err = float(np.max(np.abs(basis @ w - want)))
            if rank == 6 and np.isfinite(cond) and cond <= 1e8 and err <= 1e-10:
                break
        else:
            raise ValueError('Unresolved exterior pressure polynomial')
        for q, c in enumerate(near):
            data[(node, labels[pool[c]])] = data.get((node, labels[pool[c]]), 0.0) + w[q]
        cond_max, err_max, count_max = max(cond_max, cond), max(err_max, err), max(count_max, len(near))
    keys = list(data)
    P = sp.csr_matrix(([data[k] for k in keys], ([k[0] for k in keys], [k[1] for k in keys])),
                      shape=(active.size, len(ids)))
    return P, ids, dict(stencil_condition_max=cond_max, polynomial_reproduction_linf=err_max,
                        neighbor_count_max=count_max)


def knn_with_ties(centers, point, k):
    """All centres within the k-th smallest distance x (1+1e-12), ascending index order.

    Brute-force restatement of ``tree.query(point, k)`` followed by
    ``tree.query_ball_point(point, d_k*(1+1e-12))`` (cut_cell_solver.py:53-55,
    optimized_ghost_cell.py:158-160).
    """
    diff = centers - point
    dist = np.sqrt(np.sum(diff * diff, axis=1))
    dk = np.partition(dist, k - 1)[k - 1]
    return np.flatnonzero(dist <= dk * (1 + 1e-12))


def build_pressure_reconstruction(geom, points):
    """Quadratic LSQ point-pressure stencils (solver/cut_cell_solver.py:25-67)."""
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
    rows, columns, weights = [], [], []
    for k in range(len(points)):
        nearby = knn_with_ties(centers, points[k], 12)
        dy_, dx_ = ((centers[nearby] - points[k]) / [geom.dy, geom.dx]).T
        basis = np.array([np.ones(len(nearby)), dx_, dy_, dx_ * dx_, dx_ * dy_, dy_ * dy_])
        w, _, rank, _ = np.linalg.lstsq(basis, [1., 0., 0., 0., 0., 0.], rcond=1e-12)
        if rank != 6:
            raise ValueError('Pressure stencil cannot resolve a quadratic polynomial')
        rows.extend([k] * len(nearby))
        columns.extend(indices[nearby])
        weights.extend(w)
    return sp.csr_matrix((weights, (rows, columns)), shape=(len(points), geom.nx * geom.ny))


def build_surface_pressure_quadrature(geom):
    """Surface points, pressure stencils and weighted normals (cut_cell_solver.py:70-77)."""
    count = max(360, 4 * int(np.ceil(2 * np.pi * geom.R / min(geom.dx, geom.dy))))
    theta = np.arange(count) * (2 * np.pi / count)
    normals = np.column_stack((np.cos(theta), np.sin(theta)))
    points = np.column_stack((geom.yc + geom.R * normals[:, 1], geom.xc + geom.R * normals[:, 0]))
    return points, build_pressure_reconstruction(geom, points), normals * (2 * np.pi * geom.R / count)


def wall_gradient(x, y, active, points, geom, h):
    """Wall-normal derivative stencils (solver/optimized_ghost_cell.py:152-172)."""
    xx, yy = np.meshgrid(x, y)
    radius = np.hypot(xx - geom.xc, yy - geom.yc)
    ids = np.flatnonzero(active & (radius < geom.R + 6 * h))
    coords = np.column_stack((yy.ravel()[ids], xx.ravel()[ids]))
    rows, cols, values = [], [], []
    for k in range(len(points)):
        near = knn_with_ties(coords, points[k], 14)
        ys, xs = coords[near].T
        theta = np.arctan2(points[k, 0] - geom.yc, points[k, 1] - geom.xc)
        angle = np.arctan2(ys - geom.yc, xs - geom.xc) - theta
        n = (np.hypot(xs - geom.xc, ys - geom.yc) - geom.R) / h
        t = geom.R * np.arctan2(np.sin(angle), np.cos(angle)) / h
        weights, _, rank, _ = np.linalg.lstsq(np.array([n, n * n, n * t]), [1 / h, 0., 0.], rcond=1e-12)
        if rank != 3:
            raise ValueError('Unresolved wall-gradient stencil')
        rows.extend([k] * len(near))
        cols.extend(ids[near])
        values.extend(weights)
    return sp.csr_matrix((values, (rows, cols)), shape=(len(points), active.size))


# --------------------------------------------------------------------------------------
# Separable rectangle inverse (solver/rectangular_inverse.py:8-58)
# --------------------------------------------------------------------------------------
def rectangle_parameters(kind, nx, ny):
    """Boundary diagonals and transform types of each rectangle (rectangular_inverse.py:9-31)."""
    if kind == 'pressure':
        return dict(xx=nx, yy=ny, ends=(1., 3., 1., 1.), sx=(np.arange(nx) + .5) * np.pi / (2 * nx),
                    sy=np.arange(ny) * np.pi / (2 * ny), tx=4, ty=2, cos=True)
    if kind == 'u':
        return dict(xx=nx, yy=ny, ends=(2., 2., 3., 3.), sx=(np.arange(nx) + .5) * np.pi / (2 * nx),
                    sy=(np.arange(ny) + 1) * np.pi / (2 * ny), tx=3, ty=2, cos=False)
    if kind == 'v':
        return dict(xx=nx, yy=ny - 1, ends=(3., 1., 2., 2.), sx=(np.arange(nx) + .5) * np.pi / (2 * nx),
                    sy=(np.arange(ny - 1) + 1) * np.pi / (2 * ny), tx=4, ty=1, cos=False)
    raise ValueError('Unknown rectangular MAC field')


def rectangle(nx, ny, dx, dy, kind, nu=0., dt=0., rho=1., workers=None):
    """(matrix, solve, shape) of the cylinder-free operator (rectangular_inverse.py:8-58)."""
    prm = rectangle_parameters(kind, nx, ny)
    xx, yy = prm['xx'], prm['yy']
    left, right, lower, upper = prm['ends']
    diagx = np.full(xx, -2.); diagx[0] = -left; diagx[-1] = -right
    diagy = np.full(yy, -2.); diagy[0] = -lower; diagy[-1] = -upper
    lowerx = np.ones(xx - 1)
    if kind == 'u':
        lowerx[-1] = 2.
    lx = sp.diags((lowerx, diagx, np.ones(xx - 1)), (-1, 0, 1), format='csr') / dx**2
    ly = sp.diags((np.ones(yy - 1), diagy, np.ones(yy - 1)), (-1, 0, 1), format='csr') / dy**2
    lap = sp.kron(sp.eye(yy), lx) + sp.kron(ly, sp.eye(xx))
    rates = 4 * np.sin(prm['sy'][:, None])**2 / dy**2 + 4 * np.sin(prm['sx'][None, :])**2 / dx**2
    if kind == 'pressure':
        matrix = (lap / rho).tocsc()
        divisor = -rates / rho
    else:
        matrix = (sp.eye(xx * yy) - .5 * dt * nu * lap).tocsc()
        divisor = 1 + .5 * dt * nu * rates
    forward, backward = (dct, idct) if prm['cos'] else (dst, idst)
    tx, ty = prm['tx'], prm['ty']
    if workers is None:
        workers = int(os.environ.get('FFT_WORKERS', 4))

    def solve(rhs):
        transformed = forward(forward(rhs, type=tx, axis=1, workers=workers), type=ty, axis=0,
                              overwrite_x=True, workers=workers)
        transformed /= divisor
        return backward(backward(transformed, type=ty, axis=0, overwrite_x=True, workers=workers),
                        type=tx, axis=1, overwrite_x=True, workers=workers)

    probe = np.random.default_rng(17).normal(size=(yy, xx))
    error = np.max(abs(matrix @ solve(probe).ravel() - probe.ravel()))
    if error > 1e-10:
        raise RuntimeError(f'Rectangular transform mismatch: {kind}, {error:g}')
    return matrix, solve, (yy, xx)


def _scipy_matvec(matrix, x):
    return matrix @ x


class OpenMPKernels:
    """The reference's Numba prange loops (numba_kernels.py:17-142 flux-form advection/diffusion,
    :347-354 CSR product) as the OpenMP loops of oracle/cpu_kernels.cpp (``make -C oracle cpu``),
    for the benchmark's CPU baseline (``CylinderFlow(openmp_threads=n)``).  Products and fluxes equal
    the SciPy/NumPy ones bit for bit: each CSR row is summed from 0.0 in stored order, as SciPy's
    csr_matvec does (a CSC matrix is converted to CSR, whose sorted rows sum in csc_matvec's order),
    and the flux expressions are evaluated in flux_rhs's order without FP contraction."""

    def __init__(self, threads):
        import ctypes
        path = Path(__file__).resolve().with_name('libghost_cpu.so')
        if not path.exists():
            raise RuntimeError(f'{path} is missing: build it with make -C oracle cpu')
        lib = ctypes.CDLL(str(path))
        ptr = np.ctypeslib.ndpointer
        f8, i4 = ptr(np.float64, flags='C'), ptr(np.int32, flags='C')
        lib.gc_csr_matvec.argtypes = [ctypes.c_int64, i4, i4, f8, f8, f8]
        lib.gc_flux_rhs.argtypes = [ctypes.c_int, ctypes.c_int, f8, f8] + [ctypes.c_double] * 5 + [f8, f8, f8]
        lib.gc_set_threads.argtypes = [ctypes.c_int]
        lib.gc_get_threads.restype = ctypes.c_int
        lib.gc_set_threads(int(threads))
        self.lib = lib
        self.threads = lib.gc_get_threads()
        self._csr = {}

    def matvec(self, matrix, x):
        entry = self._csr.get(id(matrix))
        if entry is None or entry[0] is not matrix:
            c = matrix.tocsr()
            entry = (matrix, np.ascontiguousarray(c.indptr, np.int32), np.ascontiguousarray(c.indices, np.int32),
                     np.ascontiguousarray(c.data, np.float64), c.shape)
            self._csr[id(matrix)] = entry
        _, indptr, indices, data, shape = entry
        x = np.ascontiguousarray(x, np.float64)
        if x.shape != (shape[1],):
            raise ValueError('matvec shape mismatch')
        y = np.empty(shape[0])
        self.lib.gc_csr_matvec(shape[0], indptr, indices, data, x, y)
        return y

    def flux_rhs(self, u, v, dx, dy, nu, include_advection, rho=1.):
        ny, nxp1 = u.shape
        nx = nxp1 - 1
        if v.shape != (ny + 1, nx):
            raise ValueError('flux_rhs shape mismatch')
        ru, rv = np.empty_like(u), np.empty_like(v)
        work = np.empty(2 * (ny + 1) * (nx + 1))
        self.lib.gc_flux_rhs(ny, nx, np.ascontiguousarray(u, np.float64), np.ascontiguousarray(v, np.float64),
                             dx, dy, nu, 1. if include_advection else 0., rho, ru, rv, work)
        return ru, rv


# --------------------------------------------------------------------------------------
# Boundary-row (Woodbury) correction and the monitored pressure solve
# (solver/pressure_schur.py:16-125)
# --------------------------------------------------------------------------------------
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
    return rows, delta[rows], tolerance


def boundary_row_inverse(actual, indices, rect, inverse, shape, extra=0, matvec=None):
    """Woodbury correction of a separable inverse with <= 2 refinements (pressure_schur.py:75-125).

    ``extra`` (a study option, not the reference's rule) adds that many refinements after the
    reference's stopping rule.  Each one adds the inverse applied to a residual already at the
    round-off floor, which moves the solution along the operator's weakly determined directions
    while the residual stays at the floor: the size of that move is how far the rule determines the
    solution, and so how closely two correct implementations can agree.  ``matvec(M, x)`` computes
    the step-time sparse products (default ``M @ x``)."""
    mv = matvec or _scipy_matvec
    indices = np.asarray(indices)
    rows, correction, tolerance = boundary_rows(actual, indices, rect)
    if len(rows) > 2048:
        raise ValueError('Ghost inverse exceeds the bounded boundary-row rank')
    schur = np.eye(len(rows))
    impulse = np.zeros(shape)
    for column, row in enumerate(rows):
        impulse.ravel()[row] = 1.
        schur[:, column] += correction @ inverse(impulse).ravel()
        impulse.ravel()[row] = 0.
    with warnings.catch_warnings():
        warnings.simplefilter('error', LinAlgWarning)
        factor = lu_factor(schur)
    rhs_buffer = np.zeros(shape)

    def apply(values):
        rhs = rhs_buffer
        rhs.fill(0.)
        rhs.ravel()[indices] = values
        first = inverse(rhs)
        rhs.fill(0.)
        rhs.ravel()[rows] = lu_solve(factor, mv(correction, first.ravel()))
        return (first - inverse(rhs)).ravel()[indices]

    details = dict(rank=len(rows), dense_bytes=int(schur.nbytes), roundoff_drop_tolerance=tolerance,
                   rows=rows, schur=schur)

    def solve(values):
        result = apply(values)
        target = max(1e-10, 32 * EPS * max(1., float(np.max(abs(values)))))
        count = 0
        for _ in range(2):
            residual = values - mv(actual, result)
            if np.max(abs(residual)) <= target:
                break
            result += apply(residual)
            count += 1
        for _ in range(extra):
            result += apply(values - mv(actual, result))
        details['last_refinement_iterations'] = count
        return result

    return solve, details


def monitored_pressure_solver(actual, indices, rect, inverse, shape, iterative=False, extra=0, matvec=None):
    """Monitored pressure solve (pressure_schur.py:16-72).

    Direct: the boundary-row corrected rectangle inverse.  Iterative: SciPy's GMRES on ``actual``,
    left-preconditioned by the rectangle inverse of the zero-filled rectangle right-hand side
    restricted to ``indices``, with the reference's arguments (restart 50, 100 cycles,
    rtol = atol = 1e-12, Arnoldi steps counted by the pr_norm callback), then at most two
    refinement solves on the true residual while it exceeds the target.  Both end with the same
    true-residual gate.  ``gmres`` is looked up in this module at call time.
    """
    mv = matvec or _scipy_matvec
    indices = np.asarray(indices)
    absolute = abs(actual).tocsr()
    roundoff_weights = (2 * np.diff(absolute.indptr) + 2) * EPS
    if iterative:
        def precondition(values):
            full = np.zeros(shape)
            full.ravel()[indices] = values
            return inverse(full).ravel()[indices]
        preconditioner = LinearOperator(actual.shape, matvec=precondition, dtype=float)
        info = dict(backend='gmres_rectangle', preconditioner='separable_rectangle', krylov_rtol=1e-12,
                    krylov_atol=1e-12)
    else:
        direct, info = boundary_row_inverse(actual, indices, rect, inverse, shape, extra, matvec)
        info['backend'] = 'fft_boundary_schur'

    def solve(values):
        started = time.perf_counter()
        if not np.isfinite(values).all():
            raise ValueError('Pressure RHS must be finite')
        steps = [0]

        def counted(_):
            steps[0] += 1

        def krylov(rhs):
            return gmres(actual, rhs, M=preconditioner, rtol=info['krylov_rtol'], atol=info['krylov_atol'],
                         restart=50, maxiter=100, callback=counted, callback_type='pr_norm')

        if iterative:
            result, status = krylov(values)
        else:
            result, status = direct(values), 0
        residual = values - mv(actual, result)
        target = max(1e-10, 64 * EPS * max(1., float(np.max(abs(values)))))
        refinements = info.get('last_refinement_iterations', 0)
        if iterative:
            refinements = 0
            while refinements < 2 and not status and np.max(abs(residual)) > target:
                correction, status = krylov(residual)
                result = result + correction
                residual = values - mv(actual, result)
                refinements += 1
        linf = float(np.max(abs(residual)))
        relative = float(np.linalg.norm(residual) / max(np.linalg.norm(values), np.finfo(float).tiny))
        action_scale = mv(absolute, abs(result)) + abs(values)
        roundoff_bound = float(np.max(roundoff_weights * action_scale))
        target = max(target, roundoff_bound)
        info['last_solve'] = dict(iterations=steps[0], residual_linf=linf, residual_relative_l2=relative,
                                  refinement_iterations=refinements, residual_target_linf=target,
                                  roundoff_bound_linf=roundoff_bound,
                                  backward_error_linf=linf / max(float(np.max(action_scale)), np.finfo(float).tiny),
                                  status=int(status), seconds=time.perf_counter() - started)
        if status or not np.isfinite(result).all() or linf > target:
            raise RuntimeError(f'Pressure solve failed: {info["last_solve"]}')
        return result

    return solve, info


# --------------------------------------------------------------------------------------
# AB2 / Crank-Nicolson step with discrete incremental pressure (solver/time_integration.py:7-145)
# --------------------------------------------------------------------------------------
def ab2_cn_step(q, pressure, dt, *, advection, diffusion, gradient, project, nu, rho=1.,
                previous_advection=None, previous_dt=None, helmholtz_solve=None, H_matvec=None,
                offset=None):
    """One AB2/CN step; returns (q, half_pressure, output_pressure, now, diagnostics).

    Follows time_integration.py:7-145 (the vector form of lines 89-145).
    """
    if (not np.isfinite(dt) or dt <= 0 or not np.isfinite(q).all() or not np.isfinite(pressure).all()):
        raise ValueError('Finite fields and a positive timestep are required')
    now = advection(q)
    if previous_advection is None:
        # projected midpoint startup (time_integration.py:33-47)
        midpoint = ((now + diffusion(q)) - gradient(pressure)) * (0.5 * dt) + q
        half, _ = project(midpoint, .5 * dt)
        explicit = advection(half)
    else:
        if previous_dt is None or not np.isfinite(previous_dt) or previous_dt <= 0:
            raise ValueError('AB2 history requires a finite positive previous timestep')
        alpha = dt / (2 * previous_dt)
        explicit = (1.0 + alpha) * now - alpha * previous_advection
    if offset is None:
        offset = diffusion(np.zeros_like(q))
    diff_q = diffusion(q)
    grad_p = gradient(pressure)
    if H_matvec is None:
        def H_matvec(x):
            return x - .5 * dt * (diffusion(x) - offset)
    rhs = q + dt * (explicit + 0.5 * diff_q - grad_p) + (0.5 * dt) * offset
    if helmholtz_solve is None:
        from scipy.sparse.linalg import LinearOperator, gmres
        H = LinearOperator((q.size, q.size), dtype=float, matvec=H_matvec)
        star, info = gmres(H, rhs, x0=q, rtol=1e-12, atol=1e-13, restart=30, maxiter=100)
    else:
        star, info = helmholtz_solve(rhs), 0
    residual = float(np.max(abs(H_matvec(star) - rhs)))
    if info or not np.isfinite(star).all() or residual > 1e-10 * max(1., float(np.max(abs(rhs)))):
        raise RuntimeError(f'CN momentum solve failed: info={info}, residual={residual:g}')
    result, _ = project(star, dt)
    projected_action = H_matvec(result)
    momentum_defect = rhs - projected_action
    _, increment = project(result + momentum_defect, dt)
    half_pressure = pressure + increment
    factor = 1. if previous_advection is None else dt / (dt + previous_dt)
    output_pressure = half_pressure + factor * increment
    coupled_residual = projected_action - rhs + dt * gradient(increment)
    return result, half_pressure, output_pressure, now, dict(
        momentum_iterations=0, momentum_residual_linf=residual,
        coupled_momentum_residual_linf=float(np.max(abs(coupled_residual))))


# --------------------------------------------------------------------------------------
# Flux-form advection / diffusion (solver/numba_kernels.py:17-142), vectorised
# --------------------------------------------------------------------------------------
def flux_rhs(u, v, dx, dy, nu, include_advection, rho=1.):
    """rhs_u (ny,nx+1), rhs_v (ny+1,nx) of compute_flux_* + accumulate_rhs_* (numba_kernels.py:17-142).

    Centred transport, unit apertures, no-slip walls.
    """
    ny, nxp1 = u.shape
    nx = nxp1 - 1
    adv = 1. if include_advection else 0.
    # F_xx (numba_kernels.py:23-35)
    uf = 0.5 * (u[:, :-1] + u[:, 1:])
    Fxx = rho * nu * (u[:, 1:] - u[:, :-1]) / dx * dy - adv * (rho * uf * uf * dy)
    # F_xy (numba_kernels.py:37-62)
    Fxy = np.zeros((ny + 1, nx + 1))
    vface = np.empty((ny - 1, nx + 1))
    vface[:, 0] = v[1:ny, 0]
    vface[:, nx] = v[1:ny, nx - 1]
    vface[:, 1:nx] = 0.5 * (v[1:ny, :-1] + v[1:ny, 1:])
    uadv = 0.5 * (u[:-1, :] + u[1:, :])
    Fxy[1:ny] = rho * nu * (u[1:, :] - u[:-1, :]) / dy * dx - adv * (rho * vface * uadv * dx)
    Fxy[0] = 2.0 * rho * nu * u[0] / dy * dx
    Fxy[ny] = -2.0 * rho * nu * u[ny - 1] / dy * dx
    ru = np.zeros_like(u)
    ru[:, 1:nx] = (Fxx[:, 1:] - Fxx[:, :-1]) + (Fxy[1:, 1:nx] - Fxy[:-1, 1:nx])
    ext_x = -rho * u[:, nx] * u[:, nx] * dy * adv
    ru[:, nx] = ext_x - Fxx[:, nx - 1] + 0.5 * (Fxy[1:, nx] - Fxy[:-1, nx])
    # F_yy (numba_kernels.py:82-94)
    vf = 0.5 * (v[:-1, :] + v[1:, :])
    Fyy = rho * nu * (v[1:, :] - v[:-1, :]) / dy * dx - adv * (rho * vf * vf * dx)
    # F_yx (numba_kernels.py:96-132)
    Fyx = np.zeros((ny + 1, nx + 1))
    uface = np.empty((ny + 1, nx - 1))
    uface[0] = u[0, 1:nx]
    uface[ny] = u[ny - 1, 1:nx]
    uface[1:ny] = 0.5 * (u[:-1, 1:nx] + u[1:, 1:nx])
    vadv = 0.5 * (v[:, :-1] + v[:, 1:])
    Fyx[:, 1:nx] = rho * nu * (v[:, 1:] - v[:, :-1]) / dx * dy - adv * (rho * uface * vadv * dy)
    Fyx[:, 0] = 2.0 * rho * nu * v[:, 0] / dx * dy
    uout = np.empty(ny + 1)
    uout[0] = u[0, nx]
    uout[ny] = u[ny - 1, nx]
    uout[1:ny] = 0.5 * (u[:-1, nx] + u[1:, nx])
    Fyx[:, nx] = -rho * uout * v[:, nx - 1] * dy * adv
    rv = np.zeros_like(v)
    rv[1:ny] = (Fyx[1:ny, 1:] - Fyx[1:ny, :-1]) + (Fyy[1:] - Fyy[:-1])
    return ru, rv


# --------------------------------------------------------------------------------------
# Integrator (solver/optimized_ghost_cell.py:61-286)
# --------------------------------------------------------------------------------------
class GhostIntegrator:
    """Operators, inverses and the AB2/CN driver of optimized_ghost_cell.py:61-286."""

    def __init__(self, solver, fft_workers=None, extra_refinements=0, kernels=None):
        self.solver = s = solver
        self.kernels = kernels
        self._mv = kernels.matvec if kernels is not None else _scipy_matvec
        gc = solver.gc
        nx, ny, dx, dy = s.nx, s.ny, s.dx, s.dy
        self.fft_workers = fft_workers
        self.extra_refinements = extra_refinements
        active_u = gc.dist_u > 1e-13
        active_v = gc.dist_v > 1e-13
        active_u[:, 0] = False
        active_v[[0, -1]] = False
        active_p = gc.dist_p > 1e-13
        self.active_u, self.active_v, self.active_p = active_u, active_v, active_p
        interp = s.ghost_velocity_interpolation
        eu, self.ui = extension(gc.xu, gc.yu, gc.sdf, active_u, interpolation=interp)
        ev, self.vi = extension(gc.xv, gc.yv, gc.sdf, active_v, interpolation=interp)
        if s.ghost_pressure_extension == 'quadratic':   # optimized_ghost_cell.py:75-79
            self.P, self.pi, self.reconstruction_fit = exterior_pressure_extension(gc.xp, gc.yp, gc.sdf, active_p)
        else:
            self.P, self.pi = extension(gc.xp, gc.yp, gc.sdf, active_p, neumann=True)
        self.Eu, self.Ev = eu, ev
        self.E = sp.block_diag((eu, ev), format='csr')
        self.usize = s.u.size
        self.qsize = len(self.ui) + len(self.vi)
        self.offset = np.zeros(self.E.shape[0])
        self.offset[:self.usize].reshape(s.u.shape)[:, 0] = s.u_inlet
        # D = Dfull E (optimized_ghost_cell.py:85-93)
        j, i = np.unravel_index(self.pi, s.p.shape)
        r = np.arange(len(self.pi))
        uf = j * (nx + 1) + i
        vf = self.usize + j * nx + i
        self.Dfull = sp.csr_matrix((np.concatenate((-np.ones(len(r)) / dx, np.ones(len(r)) / dx,
                                                    -np.ones(len(r)) / dy, np.ones(len(r)) / dy)),
                                    (np.tile(r, 4), np.concatenate((uf, uf + 1, vf, vf + nx)))),
                                   shape=(len(r), self.E.shape[0]))
        self.D = (self.Dfull @ self.E).tocsr()
        self.div_offset = self.Dfull @ self.offset
        # G = raw_g P / rho (optimized_ghost_cell.py:94-107)
        rows, cols, values = [], [], []
        ju, iu = np.unravel_index(self.ui, s.u.shape)
        self.outlet_u = np.flatnonzero(iu == nx)
        ru = np.arange(len(self.ui)); interior = iu < nx
        rows.extend((ru[interior], ru[interior], ru[~interior]))
        cols.extend((ju[interior] * nx + iu[interior], ju[interior] * nx + iu[interior] - 1, ju[~interior] * nx + nx - 1))
        values.extend((np.full(interior.sum(), 1 / dx), np.full(interior.sum(), -1 / dx), np.full((~interior).sum(), -2 / dx)))
        jv, iv = np.unravel_index(self.vi, s.v.shape)
        rv = len(ru) + np.arange(len(self.vi))
        rows.extend((rv, rv)); cols.extend((jv * nx + iv, (jv - 1) * nx + iv))
        values.extend((np.full(len(rv), 1 / dy), np.full(len(rv), -1 / dy)))
        raw_g = sp.csr_matrix((np.concatenate(values), (np.concatenate(rows), np.concatenate(cols))),
                              shape=(self.qsize, s.p.size))
        self.G = (raw_g @ self.P / s.rho).tocsr()
        self.A = (self.D @ self.G).tocsc()
        self._pressure_rectangle = rectangle(nx, ny, dx, dy, 'pressure', rho=s.rho, workers=fft_workers)
        self.set_pressure_solver(s.pressure_solver)
        self._pressure_calls = []
        # L = nu raw_l E (optimized_ghost_cell.py:113-128)
        rows, cols, values = [], [], []

        def add(row, col, value):
            rows.append(row); cols.append(col); values.append(np.broadcast_to(value, row.shape))
        add(ru, self.ui, -2 / dx**2 - np.where((ju == 0) | (ju == ny - 1), 3., 2.) / dy**2)
        add(ru, self.ui - 1, np.where(iu == nx, 2., 1.) / dx**2)
        add(ru[interior], self.ui[interior] + 1, 1 / dx**2)
        for shift, keep in ((nx + 1, ju < ny - 1), (-nx - 1, ju > 0)):
            add(ru[keep], self.ui[keep] + shift, 1 / dy**2)
        add(rv, self.usize + self.vi, -np.where(iv == 0, 3., np.where(iv == nx - 1, 1., 2.)) / dx**2 - 2 / dy**2)
        for shift, keep in ((1, iv < nx - 1), (-1, iv > 0), (nx, np.ones(len(rv), bool)), (-nx, np.ones(len(rv), bool))):
            add(rv[keep], self.usize + self.vi[keep] + shift, 1 / (dx**2 if abs(shift) == 1 else dy**2))
        raw_l = sp.csr_matrix((np.concatenate(values), (np.concatenate(rows), np.concatenate(cols))),
                              shape=(self.qsize, self.E.shape[0]))
        self.L = (s.nu * raw_l @ self.E).tocsr()
        self._raw_l = raw_l
        self.diff_offset = s.nu * (raw_l @ self.offset)
        self.previous_advection = self.previous_dt = self.half_pressure = None
        self._dt = self._helmholtz = None
        self.helmholtz_inverse_info = []
        self.u = np.zeros_like(s.u); self.v = np.zeros_like(s.v)
        # surface quadrature and wall-gradient stencils (optimized_ghost_cell.py:142-148)
        from types import SimpleNamespace
        self.geom = SimpleNamespace(nx=nx, ny=ny, dx=dx, dy=dy, Lx=s.Lx, Ly=s.Ly, xc=s.xc, yc=s.yc,
                                    R=s.D / 2, cell_type=np.where(active_p, 1, -1))
        points, self.pressure_surface, self.normal_weights = build_surface_pressure_quadrature(self.geom)
        self.surface_points = points
        self.normals = self.normal_weights / (np.pi * s.D / len(points))
        h = max(s.dx, s.dy)
        self.du_dn = wall_gradient(gc.xu, gc.yu, active_u, points, self.geom, h)
        self.dv_dn = wall_gradient(gc.xv, gc.yv, active_v, points, self.geom, h)
        self.last_diagnostics = {}

    def set_pressure_solver(self, pressure_solver):
        """(Re)build the monitored pressure solve for 'direct' or 'iterative' (optimized_ghost_cell.py:109-111)."""
        rect, inverse, shape = self._pressure_rectangle
        self.pressure_solve, self.pressure_inverse_info = monitored_pressure_solver(
            self.A, self.pi, rect, inverse, shape, iterative=pressure_solver == 'iterative',
            extra=self.extra_refinements, matvec=self._mv)

    # --- packing (optimized_ghost_cell.py:174-184)
    def pack(self, u, v):
        return np.concatenate((u.ravel()[self.ui], v.ravel()[self.vi]))

    def unpack(self, q):
        field = self._mv(self.E, q) + self.offset
        self.u[:] = field[:self.usize].reshape(self.u.shape)
        self.v[:] = field[self.usize:].reshape(self.v.shape)
        return self.u, self.v

    # --- operators (optimized_ghost_cell.py:186-234)
    def advection(self, q):
        u, v = self.unpack(q)
        s = self.solver
        if self.kernels is not None:
            ru, rv = self.kernels.flux_rhs(u, v, s.dx, s.dy, 0., True)
        else:
            ru, rv = flux_rhs(u, v, s.dx, s.dy, 0., True)
        result = self.pack(ru, rv) / (s.dx * s.dy)
        result[self.outlet_u] *= 2
        return result

    def diffusion(self, q):
        return self._mv(self.L, q) + self.diff_offset

    def gradient(self, p):
        return self._mv(self.G, p.ravel() if p.ndim > 1 else p)

    def divergence(self, q):
        return self._mv(self.D, q) + self.div_offset

    def project(self, q, dt):
        rhs = self.divergence(q) / dt
        phi = self.pressure_solve(rhs)
        self._pressure_calls.append(dict(self.pressure_inverse_info['last_solve']))
        out = q - dt * self._mv(self.G, phi)
        error = float(np.max(np.abs(self.divergence(out))))
        if error > 1e-10 or not np.isfinite(out).all():
            raise RuntimeError(f'Ghost projection failed: divergence={error:g}')
        return out, phi

    def H_matvec(self, q, dt):
        return q - 0.5 * dt * self._mv(self.L, q)

    def build_helmholtz(self, dt):
        """Per-dt u/v Helmholtz inverses (optimized_ghost_cell.py:243-255)."""
        s = self.solver
        nu = len(self.ui)
        solves = []
        self.helmholtz_inverse_info = []
        ju, iu = np.unravel_index(self.ui, self.u.shape)
        jv, iv = np.unravel_index(self.vi, self.v.shape)
        for kind, selected, block in (('u', ju * s.nx + iu - 1, self.L[:nu, :nu]),
                                      ('v', (jv - 1) * s.nx + iv, self.L[nu:, nu:])):
            rect, inverse, shape = rectangle(s.nx, s.ny, s.dx, s.dy, kind, s.nu, dt, s.rho, workers=self.fft_workers)
            actual = (sp.eye(block.shape[0], format='csc') - .5 * dt * block).tocsc()
            solve, info = boundary_row_inverse(actual, selected, rect, inverse, shape, self.extra_refinements,
                                               self._mv)
            solves.append(solve)
            self.helmholtz_inverse_info.append(info)
        self._helmholtz = lambda rhs: np.concatenate((solves[0](rhs[:nu]), solves[1](rhs[nu:])))
        self._dt = dt

    def advance(self, u, v, p, dt):
        """optimized_ghost_cell.py:236-278."""
        self._pressure_calls = []
        inlet = self.offset[:self.usize].reshape(self.u.shape)[:, 0]
        if not np.array_equal(inlet, self.solver.u_inlet):
            inlet[:] = self.solver.u_inlet
            self.div_offset = self.Dfull @ self.offset
            self.diff_offset = self.solver.nu * (self._raw_l @ self.offset)
        if self._dt != dt:
            self.build_helmholtz(dt)
        pressure = p.ravel()[self.pi] if self.half_pressure is None else self.half_pressure
        q = self.pack(u, v)
        q, half, output, history, diagnostics = ab2_cn_step(
            q, pressure, dt, advection=self.advection, diffusion=self.diffusion, gradient=self.gradient,
            project=self.project, nu=self.solver.nu, rho=self.solver.rho,
            previous_advection=self.previous_advection, previous_dt=self.previous_dt,
            helmholtz_solve=self._helmholtz, H_matvec=lambda x: self.H_matvec(x, dt),
            offset=self.diff_offset)
        self.half_pressure, self.previous_advection, self.previous_dt = half, history, dt
        self.last_q = q
        calls = self._pressure_calls
        diagnostics.update(
            pressure_backend=self.pressure_inverse_info['backend'],
            pressure_solve_count=len(calls),
            pressure_iterations=sum(c['iterations'] for c in calls),
            pressure_refinement_iterations=sum(c['refinement_iterations'] for c in calls),
            pressure_residual_linf=max(c['residual_linf'] for c in calls),
            pressure_residual_relative_l2=max(c['residual_relative_l2'] for c in calls),
            pressure_seconds=sum(c['seconds'] for c in calls), pressure_solves=list(calls))
        self.last_diagnostics = diagnostics
        uu, vv = self.unpack(q)
        return uu, vv, self._mv(self.P, output).reshape(p.shape)

    def surface_forces(self, u, v, p, mu):
        """(Fx, Fy, Fx_p, Fy_p, Fx_v, Fy_v) (optimized_ghost_cell.py:280-286)."""
        pressure = self._mv(self.pressure_surface, p.ravel())
        dn = np.column_stack((self._mv(self.du_dn, u.ravel()), self._mv(self.dv_dn, v.ravel())))
        traction = mu * (dn + self.normals * np.sum(dn * self.normals, axis=1)[:, None])
        fp = -np.sum(pressure[:, None] * self.normal_weights, axis=0)
        fv = np.sum(traction, axis=0) * (np.pi * self.solver.D / len(pressure))
        return (*(fp + fv), *fp, *fv)


# --------------------------------------------------------------------------------------
# Initial-state helpers shared by the benchmark workloads (build definitions)
# --------------------------------------------------------------------------------------
def gaussian_v_perturbation(xv, yv, xc, yc, R, D, seed, amplitude=1e-3, width=0.5):
    """Seeded transverse perturbation: v += a xi exp(-r^2/(2 w^2)) for 0 < r-R <= D.

    xi ~ N(0,1) drawn for the whole (ny+1, nx) v array in row-major order from
    ``numpy.random.default_rng(seed)``.
    """
    X, Y = np.meshgrid(xv, yv)
    r = np.hypot(X - xc, Y - yc)
    xi = np.random.default_rng(seed).standard_normal(X.shape)
    band = (r - R > 0) & (r - R <= D)
    return np.where(band, amplitude * xi * np.exp(-r * r / (2 * width * width)), 0.)


# --------------------------------------------------------------------------------------
# Solver facade, ghost-cell / AB2-CN branch (solver/main.py:40-242, 245-501, 503-585)
# --------------------------------------------------------------------------------------
class CylinderFlow:
    """Ghost AB2/CN branch of ``MAC_CylinderFlow`` (solver/main.py)."""

    def __init__(self, nx=400, ny=100, Lx=20.0, Ly=5.0, D=1.0, xc=5.0, yc=2.5, nu=0.01, rho=1.0,
                 U_inf=1.0, inflow='uniform', U_m=None, slip_walls=False, verbose=False,
                 boundary='ghost_cell', transport='centered', pressure_solver='direct',
                 marker_spacing_factor=1.5, time_scheme='ab2_cn', ghost_pressure_extension='neumann',
                 ghost_velocity_interpolation='bilinear', initial='profile', v_perturbation=None,
                 fft_workers=None, extra_refinements=0, openmp_threads=None):
        if boundary != 'ghost_cell' or time_scheme != 'ab2_cn' or transport != 'centered':
            raise ValueError('The oracle covers boundary=ghost_cell, time_scheme=ab2_cn, centered transport')
        if pressure_solver not in ('direct', 'iterative') or slip_walls:
            raise ValueError('Ghost AB2/CN requires direct/iterative pressure and no-slip walls')
        if ghost_pressure_extension not in ('neumann', 'quadratic'):
            raise ValueError('Unknown Ghost pressure extension')
        if any(isinstance(n, bool) or not isinstance(n, (int, np.integer)) or n < 4 for n in (nx, ny)):
            raise ValueError('nx and ny must be integers >= 4')
        if not all(np.isfinite(x) and x > 0 for x in (Lx, Ly, D, nu, rho, U_inf)):
            raise ValueError('Domain lengths, D, nu, rho and U_inf must be finite and positive')
        if inflow not in ('uniform', 'parabolic'):
            raise ValueError('inflow must be uniform or parabolic')
        if not (D / 2 < xc < Lx - D / 2 and D / 2 < yc < Ly - D / 2):
            raise ValueError('Cylinder must lie strictly inside the channel')
        self.ghost_pressure_extension = ghost_pressure_extension
        self.ghost_velocity_interpolation = ghost_velocity_interpolation
        self.pressure_solver = pressure_solver
        self.configuration = dict(nx=nx, ny=ny, Lx=Lx, Ly=Ly, D=D, xc=xc, yc=yc, nu=nu, rho=rho,
                                  U_inf=U_inf, inflow=inflow, U_m=U_m, slip_walls=slip_walls,
                                  boundary=boundary, transport=transport, pressure_solver=pressure_solver,
                                  marker_spacing_factor=marker_spacing_factor, time_scheme=time_scheme,
                                  verbose=False, ghost_pressure_extension=ghost_pressure_extension,
                                  ghost_velocity_interpolation=ghost_velocity_interpolation)
        self.run_options = {}
        self.last_step_diagnostics = {}
        self.time = 0.0
        self.nx, self.ny, self.Lx, self.Ly = nx, ny, Lx, Ly
        self.dx, self.dy = Lx / nx, Ly / ny
        self.nu, self.rho, self.U_inf, self.inflow = nu, rho, U_inf, inflow
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
        # inlet profile (main.py:155-164)
        y = self.y_centers
        if inflow == 'parabolic':
            self.u_inlet = 4.0 * self.U_m * y * (Ly - y) / (Ly**2)
        else:
            self.u_inlet = np.full(ny, U_inf)
        self.u_inlet_base = self.u_inlet.copy()
        # initial velocity (main.py:166-174); 'zero' keeps only the inlet column
        if initial == 'profile':
            self.u[:, :] = self.u_inlet[:, None]
            jm, im = np.nonzero(self.mask)
            self.u[jm, im] = 0.; self.u[jm, im + 1] = 0.
            self.v[jm, im] = 0.; self.v[jm + 1, im] = 0.
        elif initial == 'zero':
            self.u[:, 0] = self.u_inlet
        else:
            raise ValueError("initial must be 'profile' or 'zero'")
        self.gc = GhostGeometry(nx, ny, self.dx, self.dy, CircleSDF(xc, yc, D / 2.0))
        if v_perturbation is not None:
            self.v += gaussian_v_perturbation(self.gc.xv, self.gc.yv, xc, yc, D / 2, D, **v_perturbation)
        self.gg = GhostIntegrator(self, fft_workers=fft_workers, extra_refinements=extra_refinements,
                                  kernels=None if openmp_threads is None else OpenMPKernels(openmp_threads))
        self.A = self.gg.A
        initial_q, _ = self.gg.project(self.gg.pack(self.u, self.v), 1.)
        self.u[:], self.v[:] = self.gg.unpack(initial_q)

    def step(self, dt, *, perturb_transient=False, perturb_duration=None, perturb_amp=0.15):
        """One iteration of the step loop (main.py:271-434); returns (t, cd, cl)."""
        if perturb_duration is None:
            perturb_duration = 2.5 * (self.D / self.U_ref)
        current_time = self.time + dt
        if perturb_transient and self.inflow == 'uniform' and current_time <= perturb_duration:
            decay = 1.0 - (current_time / perturb_duration)
            pert = perturb_amp * np.sin(2.0 * np.pi * self.y_centers / self.Ly) * decay
            self.u_inlet[:] = self.u_inlet_base * (1.0 + pert)
        else:
            self.u_inlet[:] = self.u_inlet_base
        self.u[:], self.v[:], self.p[:] = self.gg.advance(self.u, self.v, self.p, dt)
        fx, fy, fx_p, fy_p, fx_v, fy_v = self.gg.surface_forces(self.u, self.v, self.p, self.nu * self.rho)
        denom = 0.5 * self.rho * (self.U_ref**2) * self.D
        cd, cl = fx / denom, fy / denom
        if not np.isfinite(cd) or not np.isfinite(cl):
            raise FloatingPointError(f'Non-finite force at t={current_time:g}; results not saved')
        self.time = current_time
        div_max = float(np.max(abs(self.gg.divergence(self.gg.pack(self.u, self.v)))))
        self.last_step_diagnostics = dict(self.gg.last_diagnostics, time=self.time, dt=dt,
                                          max_divergence=div_max,
                                          cfl=float(dt * (np.max(abs(self.u)) / self.dx + np.max(abs(self.v)) / self.dy)),
                                          cd=float(cd), cl=float(cl), cd_pressure=float(fx_p / denom),
                                          cd_viscous=float(fx_v / denom), cl_pressure=float(fy_p / denom),
                                          cl_viscous=float(fy_v / denom))
        if div_max >= 1e-10 or self.last_step_diagnostics['cfl'] >= .5:
            raise RuntimeError('Ghost step failed divergence/CFL gate')
        return current_time, cd, cl

    def solve(self, n_steps=25000, dt=0.002, perturb_transient=True, perturb_duration=None,
              perturb_amp=0.15, save_artifacts=False, verbose=None, step_callback=None):
        """Step loop of main.py:245-501 (no plots or npz artefacts)."""
        if isinstance(n_steps, bool) or not isinstance(n_steps, (int, np.integer)) or n_steps <= 0:
            raise ValueError('n_steps must be a positive integer')
        if not np.isfinite(dt) or dt <= 0:
            raise ValueError('dt must be finite and positive')
        if perturb_duration is None:
            perturb_duration = 2.5 * (self.D / self.U_ref)
        self.run_options = dict(perturb_transient=perturb_transient, perturb_duration=perturb_duration,
                                perturb_amp=perturb_amp)
        t_hist, cd_hist, cl_hist = [], [], []
        for s in range(n_steps):
            if s % 100 == 0:
                rate = np.max(np.abs(self.u)) / self.dx + np.max(np.abs(self.v)) / self.dy
                if not np.isfinite(rate) or dt * rate > .5:
                    raise RuntimeError(f'Explicit timestep unsafe at t={self.time:g}: dt*rate={dt * rate:g}; reduce dt')
            t, cd, cl = self.step(dt, perturb_transient=perturb_transient, perturb_duration=perturb_duration,
                                  perturb_amp=perturb_amp)
            if step_callback is not None:
                step_callback(self, self.last_step_diagnostics.copy())
            t_hist.append(t); cd_hist.append(cd); cl_hist.append(cl)
        return np.array(t_hist), np.array(cd_hist), np.array(cl_hist)

    # --- checkpoint (main.py:503-585), without the source-hash manifest
    def save_checkpoint(self, path, *, overwrite=False):
        path = Path(path)
        if path.exists() and not overwrite:
            raise FileExistsError(path)
        arrays = {k: getattr(self, k) for k in ('u', 'v', 'p', 'phi', 'u_inlet')}
        for k in ('previous_advection', 'half_pressure'):
            if getattr(self.gg, k) is not None:
                arrays[k] = getattr(self.gg, k)
        meta = dict(schema=1, configuration=self.configuration, time=self.time, run_options=self.run_options,
                    prev_dt=self.gg.previous_dt, cn_dt=self.gg.previous_dt)
        with tempfile.NamedTemporaryFile(dir=path.parent, suffix='.npz', delete=False) as stream:
            np.savez(stream, checkpoint_metadata=json.dumps(meta), **arrays)
            name = stream.name
        os.replace(name, path)

    @classmethod
    def from_checkpoint(cls, path):
        with np.load(path, allow_pickle=False) as data:
            meta = json.loads(str(data['checkpoint_metadata']))
            if meta.get('schema') != 1 or meta['configuration'].get('boundary') != 'ghost_cell':
                raise ValueError('Unsupported checkpoint')
            solver = cls(**meta['configuration'])
            for name in ('u', 'v', 'p', 'phi', 'u_inlet'):
                value = data[name]
                target = getattr(solver, name)
                if value.shape != target.shape or not np.isfinite(value).all():
                    raise ValueError(f'Invalid checkpoint field: {name}')
                target[:] = value
            solver.time = float(meta['time'])
            solver.run_options = meta['run_options']
            for name, shape in (('previous_advection', (solver.gg.qsize,)), ('half_pressure', (len(solver.gg.pi),))):
                value = data[name].copy() if name in data else None
                if solver.time > 0 and (value is None or value.shape != shape or not np.isfinite(value).all()):
                    raise ValueError(f'Incomplete Ghost AB2/CN history: {name}')
                setattr(solver.gg, name, value)
            solver.gg.previous_dt = meta['prev_dt']
            return solver


# --------------------------------------------------------------------------------------
# Divergence statistics (solver/src/bindings.cpp:343-390)
# --------------------------------------------------------------------------------------
def divergence_stats(u, v, mask, dx, dy):
    """(max|div|, sum div^2, count, core sum div^2, core count) over cells with mask False."""
    div = (u[:, 1:] - u[:, :-1]) / dx + (v[1:] - v[:-1]) / dy
    fluid = ~mask
    core = fluid.copy()
    core[:, 1:] &= ~mask[:, :-1]
    core[:, :-1] &= ~mask[:, 1:]
    core[1:] &= ~mask[:-1]
    core[:-1] &= ~mask[1:]
    d2 = div * div
    return (float(np.max(np.abs(div[fluid]), initial=0.)), float(d2[fluid].sum()), int(fluid.sum()),
            float(d2[core].sum()), int(core.sum()))
