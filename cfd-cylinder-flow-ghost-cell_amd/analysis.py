"""Force-history statistics of a run (postprocessing/analysis.py:8-103).

``window_statistics`` integrates piecewise-linear histories exactly over an arbitrary window;
``periodic_metrics`` finds complete lift cycles (same-direction zero crossings of the centred lift)
and reports means, extrema, RMS, the period and Strouhal number, and cycle-to-cycle drifts.  No
frequency is reported without resolved cycles.  Host-side NumPy: these run once per history, not
per step.
"""
import numpy as np


def window_statistics(t, values, left, right):
    """Time-weighted statistics of the columns of ``values`` over [left, right]
    (analysis.py:8-48): mean and centred RMS of the linear interpolant (exact trapezoid integrals
    of the data and of its square), minimum, maximum, their mid value, half amplitude and amplitude
    over the window's samples and end points."""
    t = np.asarray(t, dtype=float)
    values = np.asarray(values, dtype=float)
    ok = (t.ndim == 1 and len(t) >= 2 and values.ndim == 2 and values.shape[0] == len(t)
          and np.all(np.isfinite(t)) and np.all(np.isfinite(values)) and np.all(np.diff(t) > 0)
          and np.isfinite(left) and np.isfinite(right) and t[0] <= left < right <= t[-1])
    if not ok:
        raise ValueError('Require finite increasing times, vector data and a covered window')
    inner = np.flatnonzero((t > left) & (t < right))
    ends = np.array([[np.interp(x, t, col) for col in values.T] for x in (left, right)])
    nodes = np.concatenate(([left], t[inner], [right]))
    data = np.concatenate((ends[:1], values[inner], ends[1:]))
    h = np.diff(nodes)[:, None]
    span = right - left
    a, b = data[:-1], data[1:]
    mean = np.sum(h * (a + b), axis=0) / (2 * span)
    ca, cb = a - mean, b - mean
    var = np.sum(h * (ca * ca + ca * cb + cb * cb), axis=0) / (3 * span)
    lo, hi = data.min(axis=0), data.max(axis=0)
    return dict(mean=mean.tolist(), rms=np.sqrt(np.maximum(var, 0.0)).tolist(), minimum=lo.tolist(),
                maximum=hi.tolist(), mid_extrema=((lo + hi) / 2).tolist(), half_amplitude=((hi - lo) / 2).tolist(),
                amplitude=(hi - lo).tolist(), window=[float(left), float(right)])


def _upward_zeros(t, signal):
    """Times where the signal crosses zero upwards, by linear interpolation."""
    k = np.flatnonzero((signal[:-1] < 0) & (signal[1:] >= 0))
    return t[k] - signal[k] * (t[k + 1] - t[k]) / (signal[k + 1] - signal[k])


def periodic_metrics(t, cd, cl, t_start, D=.4, U_ref=1.):
    """Complete-cycle statistics of C_D / C_L histories from t_start on (analysis.py:51-103).
    A period is resolved with at least two upward crossings' worth of periods, a lift range above
    1e-10 and every period at least ten samples long; statistics then cover the first to the last
    crossing.  With four or more cycles, drifts between the first and last halves of the cycles
    (lift half amplitude, period, mean drag) are reported and ``stationary`` is their maximum
    below 2%."""
    t, cd, cl = (np.asarray(x, dtype=float) for x in (t, cd, cl))
    if (t.ndim != 1 or cd.shape != t.shape or cl.shape != t.shape or len(t) < 3
            or not (np.all(np.isfinite(t)) and np.all(np.isfinite(cd)) and np.all(np.isfinite(cl)))
            or np.any(np.diff(t) <= 0)):
        raise ValueError('Histories must be finite, equal-length vectors with strictly increasing time')
    if not (np.isfinite(t_start) and np.isfinite(D) and np.isfinite(U_ref)) or D <= 0 or U_ref <= 0:
        raise ValueError('Require finite t_start and positive D, U_ref')
    keep = t >= t_start
    t, cd, cl = t[keep], cd[keep], cl[keep]
    if len(t) < 3:
        raise ValueError('Analysis window contains fewer than three samples')
    centred = cl - np.trapezoid(cl, t) / (t[-1] - t[0])
    zeros = _upward_zeros(t, centred)
    periods = np.diff(zeros)
    resolved = len(periods) >= 2 and np.ptp(cl) > 1e-10 and periods.min() >= 10 * np.diff(t).max()
    cycles = len(periods) if resolved else 0
    a, b = (zeros[0], zeros[-1]) if resolved else (t[0], t[-1])
    st = window_statistics(t, np.column_stack((cd, cl)), a, b)
    out = {}
    for q, name in enumerate(('cd', 'cl')):
        out.update({f'{name}_mean': st['mean'][q], f'{name}_max': st['maximum'][q], f'{name}_min': st['minimum'][q],
                    f'{name}_mid_extrema': st['mid_extrema'][q], f'{name}_half_amplitude': st['half_amplitude'][q],
                    f'{name}_amplitude': st['amplitude'][q], f'{name}_rms': st['rms'][q]})
    out['cl_amp'] = st['half_amplitude'][1]
    period = float(periods.mean()) if resolved else float('nan')
    out.update(period=period, strouhal=float(D / U_ref / period) if resolved else float('nan'), cycles=cycles,
               sample_start=float(a), sample_end=float(b), stationary=False, amplitude_drift=float('nan'),
               period_drift=float('nan'), drag_drift=float('nan'))
    if cycles >= 4:
        amps, drags = [], []
        for lo, hi in zip(zeros[:-1], zeros[1:]):
            tt = np.concatenate(([lo], t[(t > lo) & (t < hi)], [hi]))
            amps.append(np.ptp(np.interp(tt, t, cl)) / 2)
            drags.append(np.trapezoid(np.interp(tt, t, cd), tt) / (hi - lo))
        half = cycles // 2

        def drift(x):
            x = np.asarray(x)
            late, early = np.mean(x[-half:]), np.mean(x[:half])
            return float(abs(late - early) / max(abs(late), 1e-12))

        out.update(amplitude_drift=drift(amps), period_drift=drift(periods), drag_drift=drift(drags))
        out['stationary'] = max(out['amplitude_drift'], out['period_drift'], out['drag_drift']) < .02
    return out
