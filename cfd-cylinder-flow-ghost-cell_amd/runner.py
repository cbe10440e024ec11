"""Stage-JSON runner of one ghost-cell case on the GPU.

    python cfd_amd_loader.py run-case --config-path case.json --output out/ [--restart checkpoint.npz]

The case file and the output directory follow the reference's runner
(validation/run_ghost_validation.py:19-102).  A case holds ``dt``, ``end_time``, ``sample_start``,
``kind`` ('steady', 'periodic' or 'diagnostic'), ``solver`` (``MAC_CylinderFlow`` keywords; ghost_cell /
ab2_cn only) and optionally ``profile_stride`` (default 10).  The output directory must not exist;
it receives

* ``manifest.json``: the case, the start time, the restart file and its SHA-256, the solver metadata
  and the SHA-256 of this runner, the observables and the analysis modules;
* ``history.jsonl``: one row per step, the step's diagnostics dict plus ``step``, ``delta_p`` (front
  minus back stagnation pressure) and ``wall_seconds``;
* ``centerline.bin`` + ``centerline.json``: the wake samples (``observables.CylinderSamples``);
* ``checkpoint.npz``: every 45 s of wall time, at the end, and after a failure;
* ``summary.json``: progress while running; at the end the force statistics over the sample window
  (for 'periodic' cases the last complete shedding cycles), or the error and the failing step's
  diagnostics;
* ``mean_observables.json``: window means of the wake samples.

The steps run on the device in segments that end where a wake sample is due (every ``profile_stride``
steps, as the reference's sampler counts them).  History rows come from the device's per-step
diagnostics rows, with the stagnation pressures evaluated on the device every step; the host reads
the state back once per segment.  ``wall_seconds`` of the rows inside a segment is interpolated
linearly over the segment's wall time.  After a failed step the state (and the checkpoint) is the one
at the end of the segment that holds it.
"""
import argparse
import hashlib
import json
import os
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _native as nat
from .analysis import periodic_metrics, window_statistics
from .flow import MAC_CylinderFlow, diagnostics_from_row
from .observables import CylinderSamples, mean_observables

HERE = Path(__file__).resolve().parent
KINDS = ('steady', 'periodic', 'diagnostic')
MAXIMA = ('max_divergence', 'cfl', 'pressure_residual_linf')
CHECKPOINT_SECONDS = 45.0
PROBE0, PROBE1 = nat.DIAG_NAMES.index('probe0'), nat.DIAG_NAMES.index('probe1')


def digest(path):
    return hashlib.sha256(Path(path).read_bytes()).hexdigest()


@dataclass(frozen=True)
class Case:
    """A validated stage-JSON case."""
    spec: dict
    dt: float
    end_time: float
    sample_start: float
    kind: str
    solver: dict
    profile_stride: int

    @classmethod
    def load(cls, path):
        spec = json.loads(Path(path).read_text(encoding='utf-8'))
        times = np.array([spec['dt'], spec['end_time'], spec['sample_start']], dtype=float)
        solver = spec['solver']
        if (not np.isfinite(times).all() or times[0] <= 0 or not 0 <= times[2] < times[1] or spec['kind'] not in KINDS
                or solver.get('boundary') != 'ghost_cell' or solver.get('time_scheme') != 'ab2_cn'):
            raise ValueError('Invalid explicit Ghost validation case')
        return cls(spec, float(spec['dt']), float(spec['end_time']), float(spec['sample_start']), spec['kind'],
                   solver, spec.get('profile_stride', 10))

    def steps_from(self, t):
        """Steps from time t to the end time; the sample window must start at or after t."""
        n = round((self.end_time - t) / self.dt)
        if n <= 0 or abs(t + n * self.dt - self.end_time) > 1e-9 or self.sample_start < t:
            raise ValueError('Window/end must be covered by this run at its declared timestep')
        return n

    def segments(self, n):
        """Segment lengths whose ends fall on the sampler's due steps: the sampler's first call is the
        forced initial sample, so its stride-th call is step stride - 1, then every stride steps."""
        k, out, done = self.profile_stride, [], 0
        first = k - 1 if k > 1 else k
        while done < n:
            out.append(min(first if not out else k, n - done))
            done += out[-1]
        return out


class Telemetry:
    """history.jsonl, the running maxima of the gated quantities, and summary.json."""

    def __init__(self, directory):
        self.dir = directory
        self.stream = (directory / 'history.jsonl').open('x', encoding='utf-8')
        self.maxima = dict.fromkeys(MAXIMA, 0.0)
        self.steps = 0
        self.started = time.perf_counter()

    def seconds(self):
        return time.perf_counter() - self.started

    def add(self, rows, backend, wall0, wall1):
        for i, row in enumerate(rows):
            d = diagnostics_from_row(row, backend)
            self.steps += 1
            d.update(step=self.steps, delta_p=float(row[PROBE0] - row[PROBE1]),
                     wall_seconds=wall0 + (wall1 - wall0) * (i + 1) / len(rows))
            self.stream.write(json.dumps(d, allow_nan=False) + '\n')
            self.fold(d)

    def fold(self, d):
        for key in MAXIMA:
            if key in d:
                self.maxima[key] = max(self.maxima[key], float(d[key]))

    def summary(self, **fields):
        tmp = self.dir / 'summary.tmp'
        tmp.write_text(json.dumps(dict(fields, **self.maxima), indent=2, allow_nan=False), encoding='utf-8')
        os.replace(tmp, self.dir / 'summary.json')

    def close(self):
        if not self.stream.closed:
            self.stream.close()

    def history(self):
        rows = [json.loads(line) for line in (self.dir / 'history.jsonl').read_text(encoding='utf-8').splitlines()]
        return {key: np.array([r[key] for r in rows]) for key in ('time', 'cd', 'cl', 'delta_p')}


def open_flow(case, restart):
    flow = MAC_CylinderFlow.from_checkpoint(restart) if restart else MAC_CylinderFlow(**case.solver)
    if flow.configuration != MAC_CylinderFlow.configuration_of(**case.solver):
        raise ValueError('Restart/configuration mismatch')
    return flow


def finite_or_none(values):
    return {k: (None if isinstance(v, float) and not np.isfinite(v) else v) for k, v in values.items()}


def run(config_path, output, restart=None):
    case = Case.load(config_path)
    out = Path(output)
    out.mkdir(parents=True, exist_ok=False)
    flow = open_flow(case, restart)
    n = case.steps_from(flow.time)
    t_begin = flow.time
    samples = CylinderSamples(flow, out / 'centerline.bin', stride=case.profile_stride)
    device = flow.gg.device
    device.set_probes(samples.pressure_operator)
    backend = device.pressure_info['backend']
    samples.record(force=True)
    manifest = dict(case=case.spec, initial_time=t_begin, restart=str(restart) if restart else None,
                    restart_sha256=digest(restart) if restart else None,
                    metadata=flow.metadata(case.dt, n, perturb_transient=False), runner_sha256=digest(__file__),
                    observables_sha256=digest(HERE / 'observables.py'), analysis_sha256=digest(HERE / 'analysis.py'))
    (out / 'manifest.json').write_text(json.dumps(manifest, indent=2), encoding='utf-8')
    tel = Telemetry(out)
    checkpoint = out / 'checkpoint.npz'
    saved = tel.started
    try:
        wall = 0.0
        for rows in flow.advance(n, case.dt, segment=case.segments(n), perturb_transient=False):
            wall0, wall = wall, tel.seconds()
            tel.add(rows, backend, wall0, wall)
            samples.record(force=True, pressures=rows[-1, [PROBE0, PROBE1]])
            if time.perf_counter() - saved > CHECKPOINT_SECONDS:
                samples.flush()
                tel.stream.flush()
                flow.save_checkpoint(checkpoint, overwrite=True)
                saved = time.perf_counter()
                tel.summary(completed=False, steps=tel.steps, time=flow.time, seconds=tel.seconds())
        samples.record(force=True)
    except BaseException as error:
        rows = getattr(error, 'diagnostics', None)
        failed = getattr(error, 'failed_step', None)
        if rows is not None and failed is not None:
            tel.add(rows[:failed], backend, tel.seconds(), tel.seconds())
            tel.fold(flow.last_step_diagnostics)
        samples.close()
        tel.close()
        flow.save_checkpoint(checkpoint, overwrite=True)
        tel.summary(completed=False, steps=tel.steps, time=flow.time, error=f'{type(error).__name__}: {error}',
                    last_step_diagnostics=dict(flow.last_step_diagnostics), seconds=tel.seconds())
        raise
    samples.close()
    tel.close()
    flow.save_checkpoint(checkpoint, overwrite=True)
    h = tel.history()
    window = [max(case.sample_start, float(h['time'][0])), float(h['time'][-1])]
    periodic = None
    if case.kind == 'periodic':
        periodic = periodic_metrics(h['time'], h['cd'], h['cl'], case.sample_start, D=flow.D, U_ref=flow.U_ref)
        if periodic['cycles'] < 5:
            tel.summary(completed=True, accepted=False, steps=tel.steps,
                        reason='Insufficient complete cycles; no periodic mean-wake claim')
            return
        window = [periodic['sample_start'], periodic['sample_end']]
    stats = window_statistics(h['time'], np.column_stack((h['cd'], h['cl'], h['delta_p'])), *window)
    mean = mean_observables(out / 'centerline.bin', *window)
    (out / 'mean_observables.json').write_text(json.dumps(mean, indent=2, allow_nan=False), encoding='utf-8')
    tel.summary(completed=True, steps=tel.steps, time=flow.time, seconds=tel.seconds(), statistics=stats,
                periodic=finite_or_none(periodic) if periodic else None, mean_flow_available=True,
                sample_window=window)


def main(argv=None):
    ap = argparse.ArgumentParser(description='One ghost-cell AB2/CN case from a stage JSON file, on the GPU')
    ap.add_argument('--config-path', required=True)
    ap.add_argument('--output', required=True)
    ap.add_argument('--restart')
    run(**vars(ap.parse_args(argv)))
