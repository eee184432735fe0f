"""Trajectory accuracy of one cycle run against another, with the metrics of
the reference's scoring tool (test_hmatrix/rsf_accuracy.py:65-99): the
log10 max|v| monitor trace of the run interpolated onto the reference run's
monitor times over their common span (maximum and RMS absolute deviation in
log10 units), the first-onset time difference and the mean-recurrence
difference, all signed run minus reference.  Extended with the per-event
onset differences and the peak slip rates of matched events.

Inputs are the rows of rsfg_monitor (10 columns: step, t [s], t [yr], dt,
log10 max|v|, ...) and rsfg_events (8 columns: t [s], +1 onset / -1 arrest,
log10 max|v|, mean slip, mean mu, peak |v| since onset, ...).
"""
import numpy as np

SEC_PER_YEAR = 31557600.0


def _trace(mon):
    mon = np.asarray(mon, np.float64).reshape(-1, 10)
    t, lv = mon[:, 2], mon[:, 4]
    keep = np.concatenate(([True], np.diff(t) > 0))   # strictly increasing times
    return t[keep], lv[keep]


def trace_deviation(ref_mon, run_mon):
    """(max, rms) |log10 max|v|_run - log10 max|v|_ref| on the reference's
    monitor times within the common time span (rsf_accuracy.py trace_dev)."""
    tr, lr = _trace(ref_mon)
    tx, lx = _trace(run_mon)
    if len(tr) < 2 or len(tx) < 2:
        return float("nan"), float("nan")
    lo, hi = max(tr[0], tx[0]), min(tr[-1], tx[-1])
    m = (tr >= lo) & (tr <= hi)
    if m.sum() < 2:
        return float("nan"), float("nan")
    dv = np.abs(np.interp(tr[m], tx, lx) - lr[m])
    return float(dv.max()), float(np.sqrt(np.mean(dv ** 2)))


def trace_deviation_uniform(ref_mon, run_mon, npts=2000):
    """(max, rms) of the same difference on npts equally spaced times of the
    common span.  The monitor writes its rows on state change, so most of
    them fall inside events, where log10 max|v| moves by decades within
    seconds and a small onset shift dominates the row-sampled metric; on a
    uniform time grid the interseismic trajectory carries the weight."""
    tr, lr = _trace(ref_mon)
    tx, lx = _trace(run_mon)
    if len(tr) < 2 or len(tx) < 2:
        return float("nan"), float("nan")
    lo, hi = max(tr[0], tx[0]), min(tr[-1], tx[-1])
    if not hi > lo:
        return float("nan"), float("nan")
    tt = np.linspace(lo, hi, npts)
    dv = np.abs(np.interp(tt, tx, lx) - np.interp(tt, tr, lr))
    return float(dv.max()), float(np.sqrt(np.mean(dv ** 2)))


def onsets(ev):
    ev = np.asarray(ev, np.float64).reshape(-1, 8)
    return np.sort(ev[ev[:, 1] > 0, 0]) / SEC_PER_YEAR


def event_peaks(ev):
    """(onset [yr], peak |v| [m/s]) of every event that has an arrest row"""
    ev = np.asarray(ev, np.float64).reshape(-1, 8)
    out, t_on = [], None
    for row in ev:
        if row[1] > 0:
            t_on = row[0] / SEC_PER_YEAR
        elif t_on is not None:
            out.append((t_on, row[5]))
            t_on = None
    return np.array(out).reshape(-1, 2)


def score(ref_mon, ref_ev, run_mon, run_ev):
    """dict of maxdev, rmsdev (log10 units, on the reference's monitor
    rows), maxdev_uniform, rmsdev_uniform (on a uniform time grid), dt1
    (first-onset difference [yr]), drec (mean-recurrence difference [yr]), nev_ref, nev_run,
    rec_ref [yr], donset (onset differences of the events both runs have
    [yr]) and dpeak (relative peak slip-rate differences of those events)."""
    maxdev, rmsdev = trace_deviation(ref_mon, run_mon)
    umax, urms = trace_deviation_uniform(ref_mon, run_mon)
    er, ex = onsets(ref_ev), onsets(run_ev)
    dt1 = float(ex[0] - er[0]) if len(er) and len(ex) else float("nan")
    rec_ref = float(np.mean(np.diff(er))) if len(er) >= 2 else float("nan")
    rec_run = float(np.mean(np.diff(ex))) if len(ex) >= 2 else float("nan")
    k = min(len(er), len(ex))
    pr, px = event_peaks(ref_ev), event_peaks(run_ev)
    kp = min(len(pr), len(px))
    return dict(maxdev=maxdev, rmsdev=rmsdev, maxdev_uniform=umax, rmsdev_uniform=urms,
                dt1=dt1, drec=rec_run - rec_ref,
                nev_ref=int(len(er)), nev_run=int(len(ex)), rec_ref=rec_ref,
                donset=[float(d) for d in ex[:k] - er[:k]],
                dpeak=[float(d) for d in (px[:kp, 1] - pr[:kp, 1]) / pr[:kp, 1]])
