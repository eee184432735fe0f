"""The text outputs (rsf_monitor.dat, rsf_events.dat, rsf_catalog.dat)
and the oracle's event catalog (rsf_finalize_event, src/rsf_output.c:
1078-1180), on the CPU.

The writers are pinned against files the reference wrote: the head of a
dense run's rsf_monitor.dat (test_hmatrix/out_2f_2.0km_sep4.0km_260yr_np4,
nine columns: the reference has since appended log10 mean|v| as a tenth)
and the BP5 1 km rk3bs rsf_events.dat.  The oracle's catalog is pinned by
the reference's published slider summary (slider/slider_physics.dat) and
by closed forms on the one-patch slider (area 4 l w, M0 = G A |slip|, the
slip and drop of the event rows), and on the 1000-patch c1 fault by the
event rows' ruptured counts and drops and the uniform cell area.
"""
import os

import numpy as np

import oracle
from conftest import GOLDEN
from helpers import SEC_PER_YEAR, initial_state

G_BP5 = 3.204e10


def test_monitor_writer_reproduces_the_reference_file():
    from interact_amd import output
    fname = os.path.join(GOLDEN, "ref_results", "monitor_2f_2km_dense_head.dat")
    lines = [l for l in open(fname) if not l.startswith("#")]
    rows = np.loadtxt(fname)
    assert rows.shape == (40, 9)
    ours = output.monitor_lines(np.hstack([rows, np.full((len(rows), 1), -12.0)]))
    for ref, mine in zip(lines, ours):
        # nine columns byte for byte; the tenth is appended after them
        assert mine[:len(ref) - 1] == ref[:-1], (ref, mine)
        assert mine[len(ref) - 1:] == " %12.7f\n" % -12.0


def test_event_writer_reproduces_the_reference_file():
    from interact_amd import output
    fname = os.path.join(GOLDEN, "ref_results", "bp5_1km_dense_3bs_rtol1e-4_events.dat")
    lines = [l for l in open(fname) if not l.startswith("#")]
    ref = np.loadtxt(fname)
    rows = np.zeros((len(ref), 8))
    rows[:, 0], rows[:, 1], rows[:, 2], rows[:, 3], rows[:, 4] = (ref[:, 0], ref[:, 2],
                                                                  ref[:, 3], ref[:, 4], ref[:, 5])
    ours = output.event_lines(rows)
    assert len(ours) == len(lines) > 10
    for a, b in zip(lines, ours):
        assert len(a) == len(b)
        ta, tb = a.split(), b.split()
        # the years column is t [s] / year, which the reference formed
        # from the unrounded time: equal to its printed precision
        assert [ta[0]] + ta[2:] == [tb[0]] + tb[2:], (a, b)
        assert abs(float(ta[1]) - float(tb[1])) <= 1e-8 * max(1.0, abs(float(ta[1])))


def test_catalog_writer_layout_and_round_trip(tmp_path):
    """rsf_catalog.dat rows: 13 fields in the reference's widths, the
    event number from 1, times in years, drops in MPa"""
    from interact_amd import output
    rows = np.array([[1.0e9, 1.0e9 + 15.0, 1, 1e8, 2.5, 2.5, 6.6e6, 6.6e6, 1.5, 8.0e18, 6.569],
                     [3.1e9, 3.1e9 + 9.5, 12, 4.8e7, 0.25, 1.75, 2.2e6, 5.3e6, 0.9, 3.8e17, 5.68]])
    f = tmp_path / "rsf_catalog.dat"
    output.write_catalog(str(f), rows)
    back = np.loadtxt(str(f))
    assert back.shape == (2, 13)
    assert list(back[:, 0]) == [1, 2]
    assert np.allclose(back[:, 1], rows[:, 0] / SEC_PER_YEAR, rtol=1e-9)
    assert np.allclose(back[:, 3], rows[:, 1] - rows[:, 0])
    assert list(back[:, 4]) == [1, 12]
    assert np.allclose(back[:, 8], rows[:, 6] / 1e6, rtol=1e-6)
    assert np.allclose(back[:, 11], rows[:, 9], rtol=1e-6)
    data = [l for l in open(str(f)) if not l.startswith("#")]
    # %5i %17.10f %17.10f %13.6e %8i %14.6e %13.6e %13.6e %12.6e %12.6e %13.6e %13.6e %8.4f
    assert all(len(l) == 5 + 17 + 17 + 13 + 8 + 14 + 13 + 13 + 12 + 12 + 13 + 13 + 8 + 12 + 1
               for l in data)


def _slider_problem():
    g = oracle.read_patches(os.path.join(GOLDEN, "ref_results", "geom_slider10.in"))
    K = oracle.traction_tensor(g, 0, 1)[:, :, 0, 0] * G_BP5 / 1e4
    vpl, sig, eta = 1e-9, 50e6, G_BP5 / (2 * 3464.0)
    P = oracle.RSFProblem(K, [0.01], [0.015], [0.008], K @ [-vpl], eta=eta, nthreads=1)
    X0 = initial_state(0.6 * sig + eta * vpl, sig, vpl, np.array([0.01]), 1e-6)
    return g, P, X0


def test_oracle_catalog_on_the_slider():
    """rk5dp rtol 1e-9 over 450 yr: the catalog's seven events reproduce the
    reference's slider summary, and each row agrees with closed forms"""
    g, P, X0 = _slider_problem()
    area = 4.0 * g["l"][0] * g["w"][0]
    assert area == 1e8
    r = P.solve(X0, 450 * SEC_PER_YEAR, rtol=1e-9, rdx_monitor=1e-5, rupture_vth=0.1,
                cat_area=[area], cat_mu=G_BP5)
    cat, ev = r["catalog"], r["events"]
    ref = {l.split()[0]: [float(v) for v in l.split()[1:]]
           for l in open(os.path.join(GOLDEN, "ref_results", "slider_physics.dat"))
           if not l.startswith("#")}["aging"]
    n_ev, recur, drop, peakv, dur = ref
    assert len(cat) == n_ev
    assert abs(np.diff(cat[:, 0] / SEC_PER_YEAR)[1:].mean() - recur) < 0.005
    assert abs(cat[1:, 6].mean() / 1e6 - drop) < 0.005
    assert abs(cat[1:, 8].mean() - peakv) < 0.0005
    assert abs((cat[1:, 1] - cat[1:, 0]).mean() - dur) < 0.05
    # each row against its onset and arrest event rows (one patch: the
    # event rows' mean slip and mean drop are that patch's)
    for row in cat:
        on = ev[(ev[:, 1] > 0) & (ev[:, 0] == row[0])]
        ar = ev[(ev[:, 1] < 0) & (ev[:, 0] == row[1])]
        assert len(on) == 1 and len(ar) == 1
        slip = abs(ar[0, 3] - on[0, 3])
        assert row[2] == 1 and row[3] == area
        assert abs(row[4] - slip) <= 1e-12 * slip and row[5] == row[4]
        assert abs(row[6] - ar[0, 6]) <= 1e-12 * abs(ar[0, 6]) and row[7] == row[6]
        assert row[8] == ar[0, 5]
        # M0 = G A D (Aki), Mw inverted back to M0 with the reference's
        # offset 9.05 (src/geometry/llgeo.c:757-760)
        assert abs(row[9] - G_BP5 * area * slip) <= 1e-12 * row[9]
        assert abs(10.0 ** (1.5 * row[10] + 9.05) - row[9]) <= 1e-12 * row[9]
        assert 5.0 < row[10] < 7.5


def test_oracle_catalog_on_the_c1_fault():
    """c1 (1000 two-component patches of 2 km) through the first event:
    the catalog row's ruptured count and mean drop are the arrest event
    row's, every cell's area is 4 l w = 4e6 m^2, so area = 4e6 count and
    M0 = G area mean slip; means do not exceed maxima"""
    from interact_amd import faults
    from helpers import vector_operator
    g = faults.planar_grid(2.0)
    f = faults.cycle_fields(g)
    n = len(g["l"])
    K = vector_operator(oracle.traction_tensor(g, 0, 8), f["shear_modulus"] / 1e4)
    xpl = np.zeros(2 * n)
    xpl[0::2] = -f["vpl"]
    P = oracle.RSFProblem(K, f["a"], f["b"], f["dc"], K @ xpl, ncomp=2, eta=f["eta"],
                          v0=f["v0"], f0=f["f0"], nthreads=8)
    X0 = initial_state(f["tau"], f["sigma"], f["vel"], f["a"], f["v0"], ncomp=2)
    area = 4.0 * g["l"] * g["w"]
    assert np.all(area == 4e6)
    r = P.solve(X0, 40 * SEC_PER_YEAR, rtol=1e-8, cat_area=area, cat_mu=f["shear_modulus"])
    cat, ev = r["catalog"], r["events"]
    assert len(cat) == 1
    row = cat[0]
    ar = ev[(ev[:, 1] < 0) & (ev[:, 0] == row[1])][0]
    assert row[2] == ar[7] > 10
    assert abs(row[6] - ar[6]) <= 1e-12 * abs(ar[6])
    assert row[3] == 4e6 * row[2]
    assert abs(row[9] - f["shear_modulus"] * row[3] * row[4]) <= 1e-12 * row[9]
    assert row[4] <= row[5] and row[6] <= row[7]
    assert ev[(ev[:, 1] > 0) & (ev[:, 0] == row[0])].shape[0] == 1
