"""The deck-driven drivers (rptunnel.driver) on the device.

* PIMD-TI from a deck in angstrom with a path.xyz (pimd_par.f90:60-424) gives the same
  integrands, bit for bit, as the batched TI of rptunnel.ti.run_ti on the deck's path.
  The water-cluster potential depends only on interatomic geometry, so the deck's
  alignment does not change the physics.
* Restart files (verletmodule.f90:162-185, pimd_par.f90:327-370): a run of K1 steps
  written to restart files and resumed for K2 steps ends in the state of one
  uninterrupted K1 + K2 step run, bit for bit (positions, momenta, random streams), and
  its integrand is the reference's (running sum)/(NMC + restartnmc - imin).
* The instanton driver (rpi_ser.f90) from a 1-D deck: its action, lndetJ and lndetJ0
  agree with the oracle at the device's optimum, and its splitting with E1 - E0 of the
  double well from a fine-grid diagonalisation.
"""
import math

import numpy as np
import pytest
import scipy.linalg

import oracle as O
from helpers import oracle_pes
from rptunnel import Batch, Pes, driver, restart as R, specs, ti
from rptunnel.inputs import write_xyz

pytestmark = pytest.mark.gpu

ANG = 0.529177


def _water_deck(d, sp, extra):
    frames = sp.extra["path"].reshape(-1, sp.natom, 3)
    labels = ["O", "H", "H"] * (sp.natom // 3)
    with open(d / "masses.dat", "w") as f:
        for lab, m in zip(labels, sp.mass):
            f.write(f"{lab} {float(m)!r}\n")
    for name, w in (("well1.dat", sp.well1), ("well2.dat", sp.well2)):
        with open(d / name, "w") as f:
            for row in w.reshape(-1, 3):
                f.write(" ".join(repr(float(v * ANG)) for v in row) + "\n")
    write_xyz(d / "path.xyz", frames, labels)
    text = (f"&MCDATA n=16, beta=2000.0, npath={len(frames)}, natom={sp.natom}, xunit=2, "
            f"tau=20.0, dt=1.0, gamma=1.0d-3 {extra} /")
    (d / "input").write_text(text)
    return text


@pytest.mark.parametrize("thermostat", [2, 1])
def test_pimd_deck_equals_run_ti(tmp_path, thermostat):
    sp = specs.water_cluster(nmol=2)
    pes = Pes(sp)
    _water_deck(tmp_path, sp, f"nintegral=3, nrep=2, thermostat={thermostat}, imin=20, "
                              f"nmc=150, noutput=40")
    deck = driver.pimd_deck(tmp_path, pes)
    nml = deck.nml
    assert deck.path.npath == 9
    res = driver.run_pimd(deck, pes, seed=5)

    def factory(pols):
        return Batch(pes, deck.mass, pols, nml["tau"], nml["dt"], gamma=nml["gamma"],
                     thermostat=thermostat, noutput=nml["noutput"], seed=5)

    want = ti.run_ti(factory, sp, nml["n"], nml["beta"], 3, 2, 150, 20, path=deck.path.flat())
    assert np.array_equal(res["integrands"], want["integrands"])
    assert res["deltaA"] == want["deltaA"] and np.all(np.isfinite(res["integrands"]))


def test_restart_resume_is_the_uninterrupted_run(tmp_path):
    sp = specs.water_cluster(nmol=2)
    pes = Pes(sp)
    K1, K2 = 130, 70
    _water_deck(tmp_path, sp, f"nintegral=2, nrep=1, thermostat=2, restart=1, nmc={K1}, "
                              "noutput=50")
    deck = driver.pimd_deck(tmp_path, pes)
    n, natom, betan = deck.nml["n"], deck.natom, deck.nml["betan"]
    driver.run_pimd(deck, pes, workdir=tmp_path, seed=3)
    st1 = [R.read_restart(tmp_path / R.restart_filename(0, k), n, natom) for k in (1, 2)]
    assert all(s.nmc == K1 for s in st1)
    deck.nml.update(restart=2)
    res = driver.run_pimd(deck, pes, workdir=tmp_path, seed=3, nmc=K2)
    st2 = [R.read_restart(tmp_path / R.restart_filename(0, k), n, natom) for k in (1, 2)]
    assert all(s.nmc == K1 + K2 for s in st2)
    # the same tasks run uninterrupted for K1 + K2 steps
    tt = driver.ti_tasks(deck)
    pols = [dict(n=n, betan=betan, xi=t["xi"], a=tt["start"], b=t["b"], dbdl=t["dbdl"],
                 stream=t["k"]) for t in tt["tasks"]]
    nml = deck.nml
    b = Batch(pes, deck.mass, pols, nml["tau"], nml["dt"], gamma=nml["gamma"], thermostat=2,
              noutput=nml["noutput"], seed=3)
    b.set_path(deck.path.lampath, deck.path.flat())
    b.init_path()
    b.run(K1 + K2)
    s, cnt, _ = b.ti_readout()
    cur = b.cursors()
    for q in range(2):
        x, p = b.get_state(q)
        assert np.array_equal(st2[q].x, x) and np.array_equal(st2[q].p, p)
        assert st2[q].cursor == cur[q]
        # integrand: (file sum + resumed sum) / (NMC + restartnmc - imin) / betan^2
        assert res["integrands"][q] == pytest.approx(s[q] / (K1 + K2) / betan ** 2, rel=1e-12)
        assert st2[q].dhdr == pytest.approx(s[q], rel=1e-12)


@pytest.mark.parametrize("thermostat", [1, 2])
def test_restart_state_round_trip_through_the_device(tmp_path, thermostat):
    """restart = 2 with NMC = 0: the files written back hold the state that was read."""
    sp = specs.water_cluster(nmol=2)
    pes = Pes(sp)
    _water_deck(tmp_path, sp, f"nintegral=2, thermostat={thermostat}, restart=1, nmc=25, "
                              "noutput=10")
    deck = driver.pimd_deck(tmp_path, pes)
    driver.run_pimd(deck, pes, workdir=tmp_path, seed=1)
    n, natom = deck.nml["n"], deck.natom
    before = [R.read_restart(tmp_path / R.restart_filename(0, k), n, natom) for k in (1, 2)]
    deck.nml.update(restart=2)
    driver.run_pimd(deck, pes, workdir=tmp_path, seed=1, nmc=0)
    after = [R.read_restart(tmp_path / R.restart_filename(0, k), n, natom) for k in (1, 2)]
    for u, v in zip(before, after):
        assert np.array_equal(u.x, v.x) and np.array_equal(u.p, v.p)
        assert u.dhdr == v.dhdr and u.nmc == v.nmc == 25
        assert (u.cursor, u.kick_cursor) == (v.cursor, v.kick_cursor)


def test_rpi_deck_double_well_splitting(tmp_path):
    V0, a, m = 0.02, 1.0, 1836.0
    sp = specs.quartic(V0=V0, a=a, mass=m)
    pes = Pes(sp)
    (tmp_path / "masses.dat").write_text(f"H {m!r}\n")
    (tmp_path / "well1.dat").write_text(f"{-a!r}\n")
    (tmp_path / "well2.dat").write_text(f"{a!r}\n")
    n, beta = 256, 12000.0
    deck = driver.rpi_deck(tmp_path, pes, text=f"&RPIDATA n={n}, beta={beta!r}, ndim=1, "
                                                "readpath=.false. /")
    out = driver.run_rpi(deck, pes, opts=dict(pgtol=1e-8, factr=1e3))
    assert out["status"] == 0, out["status"]
    betan = beta / n
    op = oracle_pes(sp)
    ch = O.Chain(n, 1, 1, [m], betan, True)
    wa, wb = np.array([-a]), np.array([a])
    xo = out["instanton"]
    assert out["skink"] == pytest.approx(betan * ch.UM(op, xo, wa, wb), rel=1e-12)
    ev = O.detj_eigs(ch.UMhessian(op, xo, False))
    ev0 = O.detj_eigs(ch.UMhessian(op, np.full((1, n), -a), True))
    assert out["lndetj"] == pytest.approx(O.lndetj(ev), rel=1e-9)
    assert out["lndetj0"] == pytest.approx(O.lndetj0(ev0), rel=1e-9)
    # E1 - E0 on a fine grid (3-point Laplacian)
    x = np.linspace(-2.6, 2.6, 6001)
    h = x[1] - x[0]
    d = 1.0 / (m * h * h) + V0 * ((x / a) ** 2 - 1) ** 2
    e = np.full(x.size - 1, -0.5 / (m * h * h))
    E = scipy.linalg.eigh_tridiagonal(d, e, select="i", select_range=(0, 1), eigvals_only=True)
    assert abs(out["delta"] / (E[1] - E[0]) - 1) < 0.1, (out["delta"], E[1] - E[0])


def test_rpi_angular_rows_equal_single_orientation_runs(tmp_path):
    """Every orientation of the batched angular quadrature equals the non-angular driver
    run with well 2 replaced by that orientation's rotated well (one chain per Ring)."""
    k = np.array([2e-3, 3e-3, 5e-3])
    sp = specs.harmonic(k, np.zeros(3), [1836.0], ndim=3)
    pes = Pes(sp)
    (tmp_path / "masses.dat").write_text("H 1836.0\n")
    (tmp_path / "well1.dat").write_text("0.0 0.0 0.0\n")
    (tmp_path / "well2.dat").write_text("0.4 0.1 -0.2\n")
    text = "&RPIDATA n=12, beta=600.0, natom=1, readpath=.false., angular=.true., npoints=2 /"
    deck = driver.rpi_deck(tmp_path, pes, text=text)
    opts = dict(pgtol=1e-10, factr=10.0)
    ang = driver.run_rpi(deck, pes, opts=opts)
    assert len(ang["rows"]) == 8
    for row in ang["rows"]:
        assert row["status"] == 0
        single = driver.rpi_deck(tmp_path, pes, text=text.replace(".true.", ".false."))
        single.well2 = driver.rotated_well(deck.well2, row["angles"])
        out = driver.run_rpi(single, pes, opts=opts)
        assert out["lndetj0"] == pytest.approx(ang["lndetj0"], rel=1e-13)
        assert row["skink"] == pytest.approx(out["skink"], rel=1e-10)
        assert row["lndetj"] == pytest.approx(out["lndetj"], rel=1e-10, abs=1e-10)
        want = driver.orientation_ibeta(out["theta"], 12, deck.nml["ndim"])
        assert row["Ibeta"] == pytest.approx(want, rel=1e-9)
