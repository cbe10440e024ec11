"""Stage-JSON runner of one ghost-cell case with telemetry (validation/run_ghost_validation.py:19-111).

    python cfd_amd_loader.py run-case --config-path case.json --output out/ [--restart checkpoint.npz]

A case file holds ``dt``, ``end_time``, ``sample_start``, ``kind`` ('steady', 'periodic' or
'diagnostic'), ``solver`` (the ``MAC_CylinderFlow`` keywords; ghost_cell / ab2_cn only) and
optionally ``profile_stride``.  The output directory (created, must not exist) receives
``manifest.json``, ``history.jsonl`` (one row per step: the step's diagnostics plus ``step``,
``delta_p`` = front minus back stagnation pressure and ``wall_seconds``), ``centerline.bin`` +
``.json`` (wake samples), ``checkpoint.npz`` (every 45 s and at the end), ``summary.json``
(status; on failure the error and the failed step's diagnostics) and ``mean_observables.json``.
"""
import argparse
import hashlib
import json
import os
import time
from pathlib import Path

import numpy as np

from .analysis import periodic_metrics, window_statistics
from .flow import MAC_CylinderFlow
from .observables import CylinderSamples, mean_observables

HERE = Path(__file__).resolve().parent


def _sha256(path):
    return hashlib.sha256(Path(path).read_bytes()).hexdigest()


def _finite_or_none(d):
    return {k: (None if isinstance(v, float) and not np.isfinite(v) else v) for k, v in d.items()}


def run(config_path, output, restart=None):
    case = json.loads(Path(config_path).read_text(encoding='utf-8'))
    dt, end, left = case['dt'], case['end_time'], case['sample_start']
    solver_opts = case['solver']
    if (not np.all(np.isfinite([dt, end, left])) or dt <= 0 or not 0 <= left < end
            or case['kind'] not in ('steady', 'periodic', 'diagnostic')
            or solver_opts.get('boundary') != 'ghost_cell' or solver_opts.get('time_scheme') != 'ab2_cn'):
        raise ValueError('Invalid explicit Ghost validation case')
    output = Path(output)
    output.mkdir(parents=True, exist_ok=False)
    s = MAC_CylinderFlow.from_checkpoint(restart) if restart else MAC_CylinderFlow(**solver_opts)
    if s.configuration != MAC_CylinderFlow.configuration_of(**solver_opts):
        raise ValueError('Restart/configuration mismatch')
    steps = round((end - s.time) / dt)
    if steps <= 0 or abs(s.time + steps * dt - end) > 1e-9 or left < s.time:
        raise ValueError('Window/end must be covered by this run at its declared timestep')
    t_begin = s.time
    samples = CylinderSamples(s, output / 'centerline.bin', stride=case.get('profile_stride', 10))
    samples.record(force=True)
    clock0 = time.perf_counter()
    last_checkpoint = clock0
    count = 0
    worst = dict(max_divergence=0.0, cfl=0.0, pressure_residual_linf=0.0)
    manifest = dict(case=case, initial_time=t_begin, restart=str(restart) if restart else None,
                    restart_sha256=_sha256(restart) if restart else None,
                    metadata=s.metadata(dt, steps, perturb_transient=False), runner_sha256=_sha256(__file__),
                    observables_sha256=_sha256(HERE / 'observables.py'), analysis_sha256=_sha256(HERE / 'analysis.py'),
                    ram_stop_guard=False, ready_for_paper1=False)
    (output / 'manifest.json').write_text(json.dumps(manifest, indent=2), encoding='utf-8')

    def status(value):
        tmp = output / 'summary.tmp'
        tmp.write_text(json.dumps(value, indent=2, allow_nan=False), encoding='utf-8')
        os.replace(tmp, output / 'summary.json')

    with (output / 'history.jsonl').open('x', encoding='utf-8') as history:
        def observe(state, diagnostics):
            nonlocal count, last_checkpoint
            count += 1
            p = samples.pressures()
            row = dict(diagnostics, step=count, delta_p=float(p[0] - p[1]), wall_seconds=time.perf_counter() - clock0)
            history.write(json.dumps(row, allow_nan=False) + '\n')
            for key in worst:
                worst[key] = max(worst[key], float(row[key]))
            samples.record(pressures=p)
            if time.perf_counter() - last_checkpoint > 45:
                samples.flush()
                history.flush()
                state.save_checkpoint(output / 'checkpoint.npz', overwrite=True)
                last_checkpoint = time.perf_counter()
                status(dict(completed=False, steps=count, time=state.time, seconds=time.perf_counter() - clock0, **worst))

        try:
            s.solve(n_steps=steps, dt=dt, perturb_transient=False, save_artifacts=False, verbose=False,
                    step_callback=observe)
            samples.record(force=True)
            samples.close()
            history.flush()
            s.save_checkpoint(output / 'checkpoint.npz', overwrite=True)
        except BaseException as error:
            samples.close()
            history.flush()
            s.save_checkpoint(output / 'checkpoint.npz', overwrite=True)
            failed = dict(s.last_step_diagnostics)
            if failed.get('time') == s.time:
                for key in worst:
                    if key in failed:
                        worst[key] = max(worst[key], float(failed[key]))
            status(dict(completed=False, steps=count, time=s.time, error=f'{type(error).__name__}: {error}',
                        last_step_diagnostics=failed, seconds=time.perf_counter() - clock0, **worst))
            raise
    rows = [json.loads(line) for line in (output / 'history.jsonl').read_text(encoding='utf-8').splitlines()]
    t = np.array([r['time'] for r in rows])
    cd = np.array([r['cd'] for r in rows])
    cl = np.array([r['cl'] for r in rows])
    dp = np.array([r['delta_p'] for r in rows])
    window = [max(left, float(t[0])), float(t[-1])]
    periodic = None
    if case['kind'] == 'periodic':
        periodic = periodic_metrics(t, cd, cl, left, D=s.D, U_ref=s.U_ref)
        if periodic['cycles'] < 5:
            status(dict(completed=True, accepted=False, ready_for_paper1=False, steps=count,
                        reason='Insufficient complete cycles; no periodic mean-wake claim', **worst))
            return
        window = [periodic['sample_start'], periodic['sample_end']]
    stats = window_statistics(t, np.column_stack((cd, cl, dp)), *window)
    mean = mean_observables(output / 'centerline.bin', *window)
    (output / 'mean_observables.json').write_text(json.dumps(mean, indent=2, allow_nan=False), encoding='utf-8')
    status(dict(completed=True, ready_for_paper1=False, steps=count, time=s.time, seconds=time.perf_counter() - clock0,
                statistics=stats, periodic=_finite_or_none(periodic) if periodic else None, mean_flow_available=True,
                benchmark_accepted=False,
                acceptance_note='Reference-specific force, pressure, Lr, stationarity and domain checks are separate.',
                sample_window=window, **worst))


def main(argv=None):
    ap = argparse.ArgumentParser(description='One ghost-cell AB2/CN case from a stage JSON file, on the GPU')
    ap.add_argument('--config-path', required=True)
    ap.add_argument('--output', required=True)
    ap.add_argument('--restart')
    run(**vars(ap.parse_args(argv)))
