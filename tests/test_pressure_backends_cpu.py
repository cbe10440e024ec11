"""CPU pins of the monitored pressure solve's two backends in the oracle (pressure_schur.py:16-72)
and of the host assembly of the per-solve diagnostics records.

The GMRES backend of the oracle is SciPy's gmres called as the reference calls it; these tests pin
its monitoring around the Krylov call (refinement rule, status and residual gate) with the
reference's own failure-reporting cases (tests/test_gemini_followup.py:118-139), and both backends
against a sparse LU solve and against each other over 12 steps (test_gemini_followup.py:47-64).
"""
from unittest.mock import patch

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import ghost_oracle as O

REF = dict(nx=40, ny=30, Lx=4., Ly=3., xc=1.5, yc=1.5)   # test_gemini_followup.py:24-26


def test_gmres_failure_reporting():
    matrix = 2 * sp.eye(2, format='csr')
    solve, info = O.monitored_pressure_solver(matrix, np.arange(2), matrix, lambda a: a / 2, (1, 2), iterative=True)
    assert np.max(abs(solve(np.array([2., 4.])) - [1., 2.])) < 1e-12
    last = info['last_solve']
    assert last['iterations'] > 0 and last['status'] == 0
    assert last['residual_target_linf'] >= 1e-10 and last['backward_error_linf'] < 1e-12
    assert np.array_equal(solve(np.zeros(2)), np.zeros(2))
    # a first answer off by 1e-7 is refined once
    with patch('oracle.ghost_oracle.gmres', side_effect=[(np.array([1. + 1e-7, 2.]), 0),
                                                         (np.array([-1e-7, 0.]), 0)]) as calls:
        assert np.max(abs(solve(np.array([2., 4.])) - [1., 2.])) < 1e-12
        assert calls.call_count == 2 and info['last_solve']['refinement_iterations'] == 1
    # a false convergence and a reported failure are both rejected
    for status in (0, 1):
        with patch('oracle.ghost_oracle.gmres', return_value=(np.zeros(2), status)):
            with pytest.raises(RuntimeError):
                solve(np.ones(2))
    with pytest.raises(ValueError):
        solve(np.array([np.nan, 1.]))


def test_backends_match_sparse_lu():
    rhs = None
    for backend in ('direct', 'iterative'):
        g = O.CylinderFlow(pressure_solver=backend, **REF).gg
        if rhs is None:
            rhs = np.random.default_rng(21).normal(size=len(g.pi))
            exact = spla.splu(g.A.tocsc()).solve(rhs)
        x = g.pressure_solve(rhs)
        last = g.pressure_inverse_info['last_solve']
        assert np.max(abs(x - exact)) <= 1e-10 * np.max(abs(exact)), backend
        assert last['residual_linf'] <= last['residual_target_linf']
        assert (last['iterations'] > 0) == (backend == 'iterative')
        assert g.pressure_inverse_info['backend'] == {'direct': 'fft_boundary_schur',
                                                      'iterative': 'gmres_rectangle'}[backend]


def test_backends_agree_over_steps():
    outputs, last = [], {}
    for backend in ('direct', 'iterative'):
        s = O.CylinderFlow(pressure_solver=backend, **REF)
        saved = []
        s.solve(n_steps=12, dt=.001, perturb_transient=False, step_callback=lambda _, d: saved.append(d))
        assert all(len(d['pressure_solves']) == d['pressure_solve_count'] for d in saved)
        assert saved[0]['pressure_solve_count'] == 3 and saved[-1]['pressure_solve_count'] == 2
        d = s.last_step_diagnostics
        assert d['pressure_residual_linf'] < 1e-9 and d['pressure_iterations'] >= 0
        outputs.append((s.u.copy(), s.v.copy(), s.p.copy()))
        last[backend] = d
    errors = [float(np.max(abs(a - b))) for a, b in zip(*outputs)]
    assert max(errors[:2]) < 1e-9 and errors[2] < 1e-8, errors
    assert last['iterative']['pressure_iterations'] > 0 and last['direct']['pressure_iterations'] == 0


def test_diagnostics_records(pkg):
    import importlib
    flow = importlib.import_module(pkg.__name__ + '.flow')
    nat = importlib.import_module(pkg.__name__ + '._native')
    row = np.zeros(nat.DIAG_LEN)
    row[nat.DIAG_NAMES.index('pressure_solve_count')] = 3
    for q in range(3):
        rec = row[nat.DIAG_SOLVE0 + nat.SOLVE_LEN * q:]
        rec[:9] = [10 + q, 1e-13 * (q + 1), 1e-14, q % 2, 1e-10, 1e-15, 1e-16, 0, .5]
    d = flow.diagnostics_from_row(row, 'gmres_rectangle')
    assert d['pressure_backend'] == 'gmres_rectangle' and d['pressure_solve_count'] == 3
    assert [c['iterations'] for c in d['pressure_solves']] == [10, 11, 12]
    assert d['pressure_iterations'] == 33 and d['pressure_refinement_iterations'] == 1
    assert d['pressure_residual_linf'] == 1e-13 * 3 and d['pressure_seconds'] == 1.5
    assert all(isinstance(c['status'], int) for c in d['pressure_solves'])
    row[nat.DIAG_NAMES.index('pressure_solve_count')] = 2
    assert len(flow.diagnostics_from_row(row)['pressure_solves']) == 2
