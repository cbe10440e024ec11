"""Write the reference-generated golden vectors of the ghost-cell tables (run in the build container).

The functions whose outputs no closed form fixes -- which nodes are ghosts, which probe cells and
stencil growth each ghost node takes, the kNN-with-ties sets of the surface stencils, the Schur
row set and rank -- are evaluated here by the reference's own code, and their outputs are stored
as data in ``tests/golden/reference_tables.npz``.  ``tests/test_reference_golden.py`` checks the
oracle and the product's host tables against that file; nothing else reads the reference.

Usage:  python tests/golden/gen_reference_tables.py [/root/reference]

What is imported from the reference (paths relative to its root):
  solver/ghost_cell.py             CircleSDF, GhostCellBoundary   (numpy/scipy only)
  solver/rectangular_inverse.py    rectangle                      (numpy/scipy only)
  solver/pressure_schur.py         boundary_row_inverse, monitored_pressure_solver (numpy/scipy only)
  solver/ghost_reconstruction.py   exterior_pressure_extension    (numpy/scipy only)
  solver/optimized_ghost_cell.py   extension (14-58) and GhostCellIntegrator (61-172), taken from the
                                   file's source: the module imports numba at the top, which this
                                   image lacks, and neither definition calls into numba during
                                   construction.
  solver/cut_cell_solver.py        build_pressure_reconstruction (25-67) and
                                   build_surface_pressure_quadrature (70-77), from the source for
                                   the same reason.
"""
import ast
import sys
import time
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import scipy.sparse as sp
from scipy.spatial import cKDTree

OUT = Path(__file__).resolve().parent / 'reference_tables.npz'

# name: (nx, ny, Lx, Ly, D, xc, yc, nu, inflow, dt, variants)
CASES = {
    'c88x16': (88, 16, 22., 4.1, 1., 2., 2., 0.01, 'uniform', 1e-3, ('bilinear/neumann',)),
    'c40x30': (40, 30, 4., 3., 1., 1.5, 1.5, 0.01, 'uniform', 1e-3,
               ('bilinear/neumann', 'biquadratic/neumann', 'bilinear/quadratic')),
    'c440x82': (440, 82, 22., 4.1, 1., 2., 2., 0.05, 'parabolic', 2e-3,
                ('bilinear/neumann', 'biquadratic/neumann', 'bilinear/quadratic')),
    'c1320x246': (1320, 246, 22., 4.1, 1., 2., 2., 0.01, 'parabolic', 5e-4, ('bilinear/neumann',)),
    'c800x820': (800, 820, 40., 41., 1., 10., 20.5, 0.01, 'uniform', 2e-3, ('bilinear/neumann',)),
}
# full solution vectors are stored up to this many unknowns, random projections above it
FULL_LIMIT = 5000


def load_definitions(path, names):
    """Execute the named top-level definitions of a reference file in a namespace of its
    numpy/scipy dependencies (the module's own import list is not executed)."""
    tree = ast.parse(Path(path).read_text())
    keep = [node for node in tree.body if isinstance(node, (ast.FunctionDef, ast.ClassDef)) and node.name in names]
    missing = set(names) - {node.name for node in keep}
    if missing:
        raise RuntimeError(f'{path}: definitions not found: {sorted(missing)}')
    return ast.Module(body=keep, type_ignores=[])


def reference_namespace(root):
    solver = Path(root) / 'solver'
    sys.path.insert(0, str(solver))
    import ghost_cell
    import ghost_reconstruction
    import pressure_schur
    import rectangular_inverse
    ns = dict(np=np, sp=sp, cKDTree=cKDTree, SimpleNamespace=SimpleNamespace,
              rectangle=rectangular_inverse.rectangle,
              boundary_row_inverse=pressure_schur.boundary_row_inverse,
              monitored_pressure_solver=pressure_schur.monitored_pressure_solver,
              exterior_pressure_extension=ghost_reconstruction.exterior_pressure_extension)
    exec(compile(load_definitions(solver / 'cut_cell_solver.py',
                                  ['build_pressure_reconstruction', 'build_surface_pressure_quadrature']),
                 str(solver / 'cut_cell_solver.py'), 'exec'), ns)
    exec(compile(load_definitions(solver / 'optimized_ghost_cell.py', ['extension', 'GhostCellIntegrator']),
                 str(solver / 'optimized_ghost_cell.py'), 'exec'), ns)
    return ns, ghost_cell, rectangular_inverse


def reference_solver(ghost_cell, nx, ny, Lx, Ly, D, xc, yc, nu, inflow, velocity, pressure):
    """The attributes GhostCellIntegrator reads from MAC_CylinderFlow (solver/main.py:84-230)."""
    dx, dy = Lx / nx, Ly / ny
    y = (np.arange(ny) + .5) * dy
    u_inlet = 4. * 1.5 * y * (Ly - y) / Ly**2 if inflow == 'parabolic' else np.ones(ny)
    gc = ghost_cell.GhostCellBoundary(nx, ny, dx, dy, ghost_cell.CircleSDF(xc, yc, D / 2.))
    return SimpleNamespace(gc=gc, nx=nx, ny=ny, dx=dx, dy=dy, Lx=Lx, Ly=Ly, D=D, xc=xc, yc=yc, nu=nu, rho=1.,
                           u=np.zeros((ny, nx + 1)), v=np.zeros((ny + 1, nx)), p=np.zeros((ny, nx)),
                           u_inlet=u_inlet, pressure_solver='direct', ghost_velocity_interpolation=velocity,
                           ghost_pressure_extension=pressure)


def ghost_rows(matrix, ids):
    """The non-identity rows of an extension matrix: row index, CSR pointer, columns, weights."""
    m = matrix.tocsr().copy()
    m.sum_duplicates()
    m.sort_indices()
    active = np.zeros(m.shape[0], bool)
    active[ids] = True
    rows = np.flatnonzero((np.diff(m.indptr) > 0) & ~active)
    sub = m[rows]
    return rows, sub.indptr, sub.indices, sub.data


def csr_parts(matrix):
    m = matrix.tocsr().copy()
    m.sum_duplicates()
    m.sort_indices()
    return m.indptr, m.indices, m.data


def projections(x, seed):
    w = np.random.default_rng(seed).normal(size=(4, x.size))
    return w @ x.ravel()


def main(root='/root/reference'):
    ns, ghost_cell, rectangular_inverse = reference_namespace(root)
    out = {}
    for name, (nx, ny, Lx, Ly, D, xc, yc, nu, inflow, dt, variants) in CASES.items():
        for variant in variants:
            t0 = time.perf_counter()
            velocity, pressure = variant.split('/')
            tag = name if variant == 'bilinear/neumann' else f'{name}_{velocity}_{pressure}'
            s = reference_solver(ghost_cell, nx, ny, Lx, Ly, D, xc, yc, nu, inflow, velocity, pressure)
            g = ns['GhostCellIntegrator'](s)
            gc = s.gc
            active_u = gc.dist_u > 1e-13
            active_u[:, 0] = False
            active_v = gc.dist_v > 1e-13
            active_v[[0, -1]] = False
            eu, ui = ns['extension'](gc.xu, gc.yu, gc.sdf, active_u, interpolation=velocity)
            ev, vi = ns['extension'](gc.xv, gc.yv, gc.sdf, active_v, interpolation=velocity)
            assert np.array_equal(ui, g.ui) and np.array_equal(vi, g.vi)
            out[f'{tag}/params'] = np.array([nx, ny, Lx, Ly, D, xc, yc, nu, inflow == 'parabolic', dt], dtype=float)
            out[f'{tag}/variant'] = np.array(variant)
            for field, matrix, ids, size in (('u', eu, ui, s.u.size), ('v', ev, vi, s.v.size),
                                             ('p', g.P, g.pi, s.p.size)):
                out[f'{tag}/{field}_inactive'] = np.setdiff1d(np.arange(size), ids).astype(np.int32)
                rows, ptr, col, val = ghost_rows(matrix, ids)
                out[f'{tag}/{field}_ghost_rows'] = rows.astype(np.int32)
                out[f'{tag}/{field}_ghost_ptr'] = ptr.astype(np.int32)
                out[f'{tag}/{field}_ghost_col'] = col.astype(np.int32)
                out[f'{tag}/{field}_ghost_w'] = val
            info = g.pressure_inverse_info
            out[f'{tag}/pressure_rank'] = np.array([info['rank']])
            if variant != 'bilinear/neumann':
                print(f'{tag}: {time.perf_counter() - t0:.1f} s', flush=True)
                continue
            # surface quadrature and wall gradients (kNN-with-ties sets + min-norm LSQ weights)
            out[f'{tag}/surface_points'] = g.surface_points
            out[f'{tag}/normal_weights'] = g.normal_weights
            for key in ('pressure_surface', 'du_dn', 'dv_dn'):
                ptr, col, val = csr_parts(getattr(g, key))
                out[f'{tag}/{key}_ptr'] = ptr.astype(np.int32)
                out[f'{tag}/{key}_col'] = col.astype(np.int32)
                out[f'{tag}/{key}_w'] = val
            # the rows of the per-dt Helmholtz boundary-row inverses (optimized_ghost_cell.py:243-255)
            nq = len(g.ui)
            ju, iu = np.unravel_index(g.ui, s.u.shape)
            jv, iv = np.unravel_index(g.vi, s.v.shape)
            for kind, selected, block in (('u', ju * nx + iu - 1, g.L[:nq, :nq]), ('v', (jv - 1) * nx + iv, g.L[nq:, nq:])):
                rect, inverse, shape = rectangular_inverse.rectangle(nx, ny, s.dx, s.dy, kind, nu, dt, 1.)
                actual = (sp.eye(block.shape[0], format='csc') - .5 * dt * block).tocsc()
                solve, details = ns['boundary_row_inverse'](actual, selected, rect, inverse, shape)
                out[f'{tag}/helmholtz_{kind}_rank'] = np.array([details['rank']])
                b = np.random.default_rng(31).normal(size=block.shape[0])
                x = solve(b)
                out[f'{tag}/helmholtz_{kind}_x' if x.size <= FULL_LIMIT else f'{tag}/helmholtz_{kind}_proj'] = (
                    x if x.size <= FULL_LIMIT else projections(x, 32))
                out[f'{tag}/dt'] = np.array([dt])
            # rectangle inverses and the monitored Woodbury pressure solve on seeded right-hand sides
            for k, kind in enumerate(('pressure', 'u', 'v')):
                _, inverse, shape = rectangular_inverse.rectangle(nx, ny, s.dx, s.dy, kind, nu, dt, 1.)
                b = np.random.default_rng(40 + k).normal(size=shape)
                x = inverse(b.copy())
                if x.size <= FULL_LIMIT:
                    out[f'{tag}/rect_{kind}_x'] = x
                else:
                    out[f'{tag}/rect_{kind}_proj'] = projections(x, 50 + k)
            b = np.random.default_rng(60).normal(size=len(g.pi))
            x = g.pressure_solve(b)
            if x.size <= FULL_LIMIT:
                out[f'{tag}/woodbury_x'] = x
            else:
                out[f'{tag}/woodbury_proj'] = projections(x, 61)
            out[f'{tag}/woodbury_refinements'] = np.array([info['last_solve']['refinement_iterations']])
            print(f'{tag}: {time.perf_counter() - t0:.1f} s, pressure rank {info["rank"]}', flush=True)
    np.savez_compressed(OUT, **out)
    print(f'wrote {OUT} ({OUT.stat().st_size / 1e3:.0f} kB, {len(out)} arrays)')


if __name__ == '__main__':
    main(*sys.argv[1:])
