"""Wake samples of a ghost-cell run on disk and their window means
(postprocessing/cylinder_observables.py:1-121).

``CylinderSamples`` appends little-endian float64 rows ``[t, p_front, p_back, u_centerline...]``
(front / back stagnation pressures from the 12-centre quadratic reconstruction, the streamwise
velocity on the cylinder's centre line from its rear face to the outlet, linearly interpolated in
y) every ``stride`` calls, with a JSON sidecar describing the columns; ``mean_observables``
integrates the rows exactly (piecewise linear in time) over a window and measures the mean
recirculation length.
"""
import json
from pathlib import Path

import numpy as np

from .tables import pressure_reconstruction


def recirculation_length(x, velocity, rear, diameter):
    """First reattachment of the mean reverse-flow region attached to the rear face, in diameters
    from the rear (cylinder_observables.py:8-26)."""
    x, velocity = np.asarray(x, dtype=float), np.asarray(velocity, dtype=float)
    if (x.ndim != 1 or velocity.shape != x.shape or len(x) < 2 or not np.isfinite(rear)
            or not np.isfinite(diameter) or diameter <= 0 or not np.all(np.isfinite(x))
            or not np.all(np.isfinite(velocity)) or np.any(np.diff(x) <= 0) or x[0] < rear - 1e-10):
        raise ValueError('Invalid wake centerline')
    negative = np.flatnonzero(velocity < 0)
    if len(negative) == 0:
        return dict(status='no_resolved_reverse_flow', length_D=None)
    first = int(negative[0])
    if first > 1:
        return dict(status='detached_negative_region', length_D=None)
    up = np.flatnonzero((velocity[first:-1] < 0) & (velocity[first + 1:] >= 0))
    if len(up) == 0:
        return dict(status='reattachment_outside_sampled_domain', length_D=None)
    i = first + int(up[0])
    xr = x[i] - velocity[i] * (x[i + 1] - x[i]) / (velocity[i + 1] - velocity[i])
    return dict(status='resolved', length_D=float((xr - rear) / diameter), x_reattachment=float(xr),
                definition='First reattachment of the connected mean reverse-flow region; distance from cylinder rear / D')


class CylinderSamples:
    """Disk-backed wake samples of a ``MAC_CylinderFlow`` (cylinder_observables.py:29-87)."""

    def __init__(self, solver, path, stride=10):
        if getattr(solver, 'gg', None) is None or isinstance(stride, bool) or not isinstance(stride, int) or stride < 1:
            raise ValueError('Cylinder sampling requires Ghost AB2/CN and a positive integer stride')
        self.solver, self.path, self.stride = solver, Path(path), stride
        self.calls, self.last = 0, -np.inf
        stagnation = np.array([[solver.yc, solver.xc - solver.D / 2], [solver.yc, solver.xc + solver.D / 2]])
        self.pressure_operator = pressure_reconstruction(solver.gg.ops.geom, stagnation)
        self.first = int(np.ceil((solver.xc + solver.D / 2) / solver.dx - 1e-12))
        self.x = np.arange(self.first, solver.nx + 1) * solver.dx
        fy = solver.yc / solver.dy - .5
        self.j = int(np.floor(fy))
        self.weight = fy - self.j
        if not 0 <= self.j < solver.ny - 1:
            raise ValueError('Wake centerline interpolation reaches an outer wall')
        meta = dict(schema=1, columns=3 + len(self.x), dtype='<f8', stride=stride, x=self.x.tolist(),
                    rear=solver.xc + solver.D / 2, diameter=solver.D,
                    quantities=['time', 'p_front', 'p_back', 'u_centerline'],
                    sampling='Linear MAC interpolation in y; temporal trapezoids clipped to the requested window')
        self.path.parent.mkdir(parents=True, exist_ok=True)
        with self.path.with_suffix('.json').open('x', encoding='utf-8') as f:
            json.dump(meta, f, indent=2)
        self.stream = self.path.open('xb')

    def pressures(self):
        return np.asarray(self.pressure_operator @ self.solver.p.ravel())

    def record(self, *, force=False, pressures=None):
        self.calls += 1
        if not force and self.calls % self.stride:
            return
        s = self.solver
        if s.time == self.last:
            return
        if s.time < self.last:
            raise ValueError('Sample times must increase')
        p = self.pressures() if pressures is None else np.asarray(pressures)
        line = (1 - self.weight) * s.u[self.j, self.first:] + self.weight * s.u[self.j + 1, self.first:]
        row = np.concatenate(([s.time], p, line)).astype('<f8')
        if not np.all(np.isfinite(row)):
            raise ValueError('Nonfinite cylinder sample')
        self.stream.write(row.tobytes())
        self.last = s.time

    def flush(self):
        self.stream.flush()

    def close(self):
        self.stream.close()


def mean_observables(path, left, right):
    """Window means of the recorded samples (exact integral of the piecewise-linear rows over
    [left, right]) and the mean recirculation length (cylinder_observables.py:90-121)."""
    path = Path(path)
    meta = json.loads(path.with_suffix('.json').read_text(encoding='utf-8'))
    ncol = meta['columns']
    size = path.stat().st_size
    if (meta['schema'] != 1 or ncol != 3 + len(meta['x']) or meta['dtype'] != '<f8' or size % (8 * ncol)
            or not (np.isfinite(left) and np.isfinite(right)) or right <= left):
        raise ValueError('Invalid sample file or integration window')
    count = size // (8 * ncol)
    if count < 2:
        raise ValueError('At least two samples are required')
    rows = np.memmap(path, dtype='<f8', mode='r', shape=(count, ncol))
    if rows[0, 0] > left or rows[-1, 0] < right:
        raise ValueError('Samples do not cover the full requested window')
    total = np.zeros(ncol - 1)
    covered, max_dt = 0.0, 0.0
    prev = np.array(rows[0])
    if not np.all(np.isfinite(prev)):
        raise ValueError('Nonfinite sample')
    for cur in rows[1:]:
        cur = np.array(cur)
        if not np.all(np.isfinite(cur)) or cur[0] <= prev[0]:
            raise ValueError('Nonfinite, duplicate or unordered sample')
        a, b = max(left, prev[0]), min(right, cur[0])
        if b > a:
            dt = cur[0] - prev[0]
            slope = (cur[1:] - prev[1:]) / dt
            total += .5 * (b - a) * (2 * prev[1:] + (a - prev[0] + b - prev[0]) * slope)
            covered += b - a
            max_dt = max(max_dt, float(dt))
        prev = cur
    if not np.isclose(covered, right - left, rtol=1e-12, atol=1e-12):
        raise ValueError('Incomplete integration window')
    mean = total / covered
    return dict(window=[float(left), float(right)], p_front_mean=float(mean[0]), p_back_mean=float(mean[1]),
                delta_p_mean=float(mean[0] - mean[1]), x=meta['x'], u_centerline_mean=mean[2:].tolist(),
                sample_max_dt=max_dt,
                recirculation=recirculation_length(meta['x'], mean[2:], meta['rear'], meta['diameter']),
                caveat='Temporal profile sampling error requires a stride-sensitivity check; no instantaneous '
                       'checkpoint is substituted for the mean.')
