"""Wake observables of a ghost-cell run: stagnation pressures, the mean centre-line wake and the
recirculation length.

The quantities and the on-disk record are the reference's (postprocessing/cylinder_observables.py:1-121):
``<name>.bin`` holds little-endian float64 rows ``[t, p_front, p_back, u(x_0) .. u(x_m)]`` and
``<name>.json`` describes them (``schema``, ``columns``, ``dtype``, ``stride``, ``x``, ``rear``,
``diameter``, ``quantities``, ``sampling``).  p_front / p_back are the 12-centre quadratic
reconstructions of p at the cylinder's upstream and downstream points; u(x) is the streamwise
velocity on the cylinder's axis, interpolated linearly in y between the two u rows around it, at the
u faces from the rear face to the outlet.

Here the stagnation operator can be handed to the device, which then evaluates both pressures on
every step's output pressure (``GhostDevice.set_probes``), and rows are buffered in blocks and
written with one call per block.  Window means integrate the piecewise-linear record exactly, in
vectorised blocks.
"""
import json
from pathlib import Path

import numpy as np

from .tables import pressure_reconstruction

SAMPLING = 'Linear MAC interpolation in y; temporal trapezoids clipped to the requested window'
DEFINITION = 'First reattachment of the connected mean reverse-flow region; distance from cylinder rear / D'
CAVEAT = ('Temporal profile sampling error requires a stride-sensitivity check; no instantaneous checkpoint is '
          'substituted for the mean.')
BLOCK_ROWS = 256


def recirculation_length(x, velocity, rear, diameter):
    """Length of the mean reverse-flow bubble behind the cylinder, in diameters from the rear face.

    The bubble must touch the rear (reverse flow at one of the first two samples); its end is the
    first upward zero crossing of u after that, located by linear interpolation."""
    x = np.asarray(x, dtype=float)
    u = np.asarray(velocity, dtype=float)
    valid = (x.ndim == 1 and u.shape == x.shape and x.size >= 2 and np.isfinite(rear) and np.isfinite(diameter)
             and diameter > 0 and np.isfinite(x).all() and np.isfinite(u).all() and bool(np.all(x[1:] > x[:-1]))
             and x[0] >= rear - 1e-10)
    if not valid:
        raise ValueError('Invalid wake centerline')
    reverse = u < 0
    if not reverse.any():
        return dict(status='no_resolved_reverse_flow', length_D=None)
    head = int(np.argmax(reverse))
    if head > 1:
        return dict(status='detached_negative_region', length_D=None)
    ups = np.flatnonzero(reverse[head:-1] & ~reverse[head + 1:]) + head
    if ups.size == 0:
        return dict(status='reattachment_outside_sampled_domain', length_D=None)
    k = int(ups[0])
    x0 = x[k] - u[k] * (x[k + 1] - x[k]) / (u[k + 1] - u[k])
    return dict(status='resolved', length_D=float((x0 - rear) / diameter), x_reattachment=float(x0),
                definition=DEFINITION)


class WakeLine:
    """Where a solver's wake profile is read: u faces ``i0 .. nx`` on the cylinder's axis, between u
    rows ``j`` and ``j + 1`` with weight ``w`` on the upper one."""

    def __init__(self, solver):
        self.i0 = int(np.ceil((solver.xc + solver.D / 2) / solver.dx - 1e-12))
        self.x = np.arange(self.i0, solver.nx + 1) * solver.dx
        position = solver.yc / solver.dy - 0.5
        self.j = int(np.floor(position))
        self.w = position - self.j
        if not 0 <= self.j < solver.ny - 1:
            raise ValueError('Wake centerline interpolation reaches an outer wall')

    def profile(self, u):
        lower, upper = u[self.j, self.i0:], u[self.j + 1, self.i0:]
        return (1 - self.w) * lower + self.w * upper


def stagnation_operator(solver):
    """2 x (ny nx) CSR: the quadratic reconstructions of p at (yc, xc - D/2) and (yc, xc + D/2)."""
    points = np.array([[solver.yc, solver.xc - solver.D / 2], [solver.yc, solver.xc + solver.D / 2]])
    return pressure_reconstruction(solver.gg.ops.geom, points)


class CylinderSamples:
    """Disk record of a ``MAC_CylinderFlow``'s wake samples (the reference's ``CylinderSamples``).

    ``record()`` counts its calls and writes a row on every ``stride``-th one (or when forced), with
    the solver's current time, stagnation pressures (given, or reconstructed from ``solver.p``) and
    wake profile.  Rows are kept in a block buffer and reach the file when the block fills, on
    ``flush()`` and on ``close()``."""

    def __init__(self, solver, path, stride=10):
        if getattr(solver, 'gg', None) is None or isinstance(stride, bool) or not isinstance(stride, int) or stride < 1:
            raise ValueError('Cylinder sampling requires Ghost AB2/CN and a positive integer stride')
        self.solver, self.path, self.stride = solver, Path(path), stride
        self.line = WakeLine(solver)
        self.pressure_operator = stagnation_operator(solver)
        self.x = self.line.x
        self.columns = 3 + self.x.size
        self.calls, self.last = 0, -np.inf
        self._block = np.empty((BLOCK_ROWS, self.columns), dtype='<f8')
        self._fill = 0
        sidecar = dict(schema=1, columns=self.columns, dtype='<f8', stride=stride, x=self.x.tolist(),
                       rear=solver.xc + solver.D / 2, diameter=solver.D,
                       quantities=['time', 'p_front', 'p_back', 'u_centerline'], sampling=SAMPLING)
        self.path.parent.mkdir(parents=True, exist_ok=True)
        with self.path.with_suffix('.json').open('x', encoding='utf-8') as f:
            json.dump(sidecar, f, indent=2)
        self.stream = self.path.open('xb')

    def pressures(self):
        """(p_front, p_back) reconstructed from the solver's current pressure field."""
        return np.asarray(self.pressure_operator @ self.solver.p.ravel())

    def record(self, *, force=False, pressures=None):
        self.calls += 1
        if self.calls % self.stride and not force:
            return
        s = self.solver
        if s.time == self.last:
            return
        if s.time < self.last:
            raise ValueError('Sample times must increase')
        row = self._block[self._fill]
        row[0] = s.time
        row[1:3] = self.pressures() if pressures is None else pressures
        row[3:] = self.line.profile(s.u)
        if not np.isfinite(row).all():
            raise ValueError('Nonfinite cylinder sample')
        self._fill += 1
        self.last = s.time
        if self._fill == BLOCK_ROWS:
            self.flush()

    def flush(self):
        if self._fill:
            self.stream.write(self._block[:self._fill].tobytes())
            self._fill = 0
        self.stream.flush()

    def close(self):
        if not self.stream.closed:
            self.flush()
            self.stream.close()


def _sample_file(path):
    """The sidecar and a read-only (rows, columns) view of a sample record, validated."""
    path = Path(path)
    meta = json.loads(path.with_suffix('.json').read_text(encoding='utf-8'))
    ncol = meta['columns']
    nbytes = path.stat().st_size
    if meta['schema'] != 1 or ncol != 3 + len(meta['x']) or meta['dtype'] != '<f8' or nbytes % (8 * ncol):
        return meta, None
    count = nbytes // (8 * ncol)
    return meta, (np.memmap(path, dtype='<f8', mode='r', shape=(count, ncol)) if count else np.empty((0, ncol)))


def window_integral(rows, left, right, block=4096):
    """Exact integral over [left, right] of the piecewise-linear interpolant of the rows' columns
    1.. in time (column 0), the covered duration and the largest sample spacing used."""
    total = np.zeros(rows.shape[1] - 1)
    covered, max_dt = 0.0, 0.0
    n = rows.shape[0]
    for b in range(0, n - 1, block):
        seg = np.asarray(rows[b:min(n, b + block + 1)])
        if not np.isfinite(seg).all():
            raise ValueError('Nonfinite, duplicate or unordered sample')
        t = seg[:, 0]
        step = np.diff(t)
        if np.any(step <= 0):
            raise ValueError('Nonfinite, duplicate or unordered sample')
        a = np.maximum(left, t[:-1])
        e = np.minimum(right, t[1:])
        use = np.flatnonzero(e > a)
        if use.size == 0:
            continue
        lo, hi = seg[use, 1:], seg[use + 1, 1:]
        slope = (hi - lo) / step[use, None]
        va = lo + (a[use] - t[use])[:, None] * slope
        vb = lo + (e[use] - t[use])[:, None] * slope
        total += np.sum(0.5 * (e[use] - a[use])[:, None] * (va + vb), axis=0)
        covered += float(np.sum(e[use] - a[use]))
        max_dt = max(max_dt, float(np.max(step[use])))
    return total, covered, max_dt


def mean_observables(path, left, right):
    """Window means of a sample record over [left, right] and the recirculation length of the mean
    wake profile (the reference's ``mean_observables``: same keys and checks)."""
    meta, rows = _sample_file(path)
    if rows is None or not (np.isfinite(left) and np.isfinite(right)) or right <= left:
        raise ValueError('Invalid sample file or integration window')
    if rows.shape[0] < 2:
        raise ValueError('At least two samples are required')
    if rows[0, 0] > left or rows[-1, 0] < right:
        raise ValueError('Samples do not cover the full requested window')
    if not np.isfinite(rows[0]).all():
        raise ValueError('Nonfinite sample')
    total, covered, max_dt = window_integral(rows, left, right)
    if not np.isclose(covered, right - left, rtol=1e-12, atol=1e-12):
        raise ValueError('Incomplete integration window')
    mean = total / covered
    return dict(window=[float(left), float(right)], p_front_mean=float(mean[0]), p_back_mean=float(mean[1]),
                delta_p_mean=float(mean[0] - mean[1]), x=meta['x'], u_centerline_mean=mean[2:].tolist(),
                sample_max_dt=max_dt,
                recirculation=recirculation_length(meta['x'], mean[2:], meta['rear'], meta['diameter']),
                caveat=CAVEAT)
