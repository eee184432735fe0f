"""Writers of the reference's text outputs from the device logs, in the
reference's column layouts, so that scripts reading rsf_solve's files read
these unchanged:

  rsf_monitor.dat  rsf_TS_Monitor rows (src/rsf_output.c:832-842), from
                   rsfg_monitor rows (10 columns);
  rsf_events.dat   rsf_post_event rows (src/rsf_output.c:1028-1032), from
                   rsfg_events rows;
  rsf_catalog.dat  rsf_finalize_event rows (src/rsf_output.c:1150-1178),
                   from rsfg_catalog rows (11 columns), numbered from 1.

Every numeric field uses the reference's printf conversion; Python's %
operator formats %i, %e and %f as C does.  The comment lines name the
columns.  file names: filenames.h:17-24 of the reference.
"""
import numpy as np

SEC_PER_YEAR = 31557600.0        # src/includes/constants.h:31

MONITOR_FILE = "rsf_monitor.dat"
EVENTS_FILE = "rsf_events.dat"
CATALOG_FILE = "rsf_catalog.dat"

MONITOR_FMT = "%9i %20.8e %17.10f %15.8e %12.7f %15.8e %10.6f %15.8e %15.8e %12.7f\n"
EVENT_FMT = "%20.8e %17.10f %2i %12.7f %15.8e %10.6f\n"
CATALOG_FMT = ("%5i %17.10f %17.10f %13.6e %8i %14.6e %13.6e %13.6e %12.6e %12.6e "
               "%13.6e %13.6e %8.4f\n")


def monitor_lines(rows):
    """rsfg_monitor rows -> rsf_monitor.dat data lines"""
    rows = np.asarray(rows, np.float64).reshape(-1, 10)
    return [MONITOR_FMT % (int(r[0]), r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9])
            for r in rows]


def event_lines(rows):
    """rsfg_events rows -> rsf_events.dat data lines (time in seconds and
    years, +1 onset / -1 arrest, log10 max|v|, mean slip, mean mu)"""
    rows = np.asarray(rows, np.float64).reshape(-1, 8)
    return [EVENT_FMT % (r[0], r[0] / SEC_PER_YEAR, int(round(r[1])), r[2], r[3], r[4])
            for r in rows]


def catalog_lines(rows, first=1):
    """rsfg_catalog rows -> rsf_catalog.dat data lines: event number,
    onset and arrest [yr], duration [s], ruptured cells, area [m^2], mean
    and max |slip change| [m], mean and max stress drop [MPa], peak slip
    rate [m/s], M0 [N m], Mw"""
    rows = np.asarray(rows, np.float64).reshape(-1, 11)
    return [CATALOG_FMT % (first + k, r[0] / SEC_PER_YEAR, r[1] / SEC_PER_YEAR, r[1] - r[0],
                           int(round(r[2])), r[3], r[4], r[5], r[6] / 1e6, r[7] / 1e6,
                           r[8], r[9], r[10])
            for k, r in enumerate(rows)]


def write_monitor(fname, rows):
    with open(fname, "w") as f:
        f.write("# columns: step, t [s], t [yr], dt [s], log10 max|v| [m/s], mean slip [m], "
                "mean mu, max sigma [Pa], min sigma [Pa], log10 mean|v| [m/s]\n")
        f.writelines(monitor_lines(rows))


def write_events(fname, rows, vel_event=1e-3):
    with open(fname, "w") as f:
        f.write("# columns: t [s], t [yr], +1 onset / -1 arrest, log10 max|v| [m/s], "
                "mean slip [m], mean mu; event threshold |v| = %.3e m/s\n" % vel_event)
        f.writelines(event_lines(rows))


def write_catalog(fname, rows, vel_event=1e-3, rupture_vth=None):
    rv = vel_event if rupture_vth is None else rupture_vth
    with open(fname, "w") as f:
        f.write("# one row per closed event that ruptured a cell; thresholds: events "
                "%.3e m/s, rupture %.3e m/s\n" % (vel_event, rv))
        f.write("# stress drop |tau_onset| - |tau_arrest|, slip as |slip change|, means over "
                "the ruptured cells\n")
        f.write("# columns: event, onset [yr], arrest [yr], duration [s], ruptured cells, "
                "area [m^2], mean slip [m], max slip [m], mean drop [MPa], max drop [MPa], "
                "peak slip rate [m/s], M0 [N m], Mw\n")
        f.writelines(catalog_lines(rows))


def write_outputs(fs, directory=".", vel_event=1e-3, rupture_vth=None):
    """the three files of a FaultSystem's last solve, in directory"""
    import os
    write_monitor(os.path.join(directory, MONITOR_FILE), fs.monitor())
    write_events(os.path.join(directory, EVENTS_FILE), fs.events(), vel_event)
    write_catalog(os.path.join(directory, CATALOG_FILE), fs.catalog(), vel_event, rupture_vth)
