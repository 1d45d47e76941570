"""Input decks and restart files (rptunnel.inputs, rptunnel.restart, the set-up half of
rptunnel.driver), on the CPU.

The readers are pinned by what the mathematics fixes rather than by restating them:
alignment by the geometry it must produce (atom1 at the origin, atom2 on +x, atom3 in the
xz half-plane z >= 0, distances and handedness kept), splines by scipy's natural cubic
spline, findmiddle by a path whose energy maximum is known, the centring map by its fixed
points, and restart files by exact round trips and by the record layout the reference's
reader (pimd_par.f90:355-370) consumes."""
import math
import os

import numpy as np
import pytest
import scipy.interpolate

from rptunnel import driver, inputs as I, restart as R, ti


# ---------------------------------------------------------------- namelists

def test_mcdata_defaults_are_the_reference_values():
    nml = I.mcdata("&MCDATA /")
    # pimd_par.f90:60-88
    want = dict(n=100, beta=100.0, nmc=5000000, noutput=100000, dt=1e-3, nintegral=5, nrep=1,
                thermostat=1, ndim=3, natom=1, xunit=1, imin=0, tau=1.0, gamma=1.0,
                instapath=False, centre=False, dhdrlimit=-1.0, readpath=True,
                alignwell=False, fixedends=True, cayley=False, readhess=False, atom1=1,
                atom2=2, atom3=3, basename="", restart=0)
    for k, v in want.items():
        assert nml[k] == v and type(nml[k]) is type(v), k
    assert nml["betan"] == 100.0 / 101.0


def test_namelist_syntax():
    text = """ some preamble the reader skips
    &mcdata
      N = 32, Beta=1.5d3   ! comment with a / slash inside a comment
      instapath=.TRUE., centre=T, fixedends=.f.
      basename='a/b ''q''', dt=2.5E-2
      nmc=200000 noutput=50
    /
    trailing text"""
    nml = I.mcdata(text)
    assert nml["n"] == 32 and nml["beta"] == 1500.0 and nml["betan"] == 1500.0 / 33
    assert nml["instapath"] is True and nml["centre"] is True and nml["fixedends"] is False
    assert nml["basename"] == "a/b 'q'" and nml["dt"] == 0.025
    assert nml["nmc"] == 200000 and nml["noutput"] == 50
    # old $group ... $end form
    assert I.mcdata("$MCDATA n=4 $END")["n"] == 4


@pytest.mark.parametrize("text, msg", [
    ("&MCDATA bogus=1 /", "unknown item"),
    ("&MCDATA n=5e6 /", "bad integer"),
    ("&MCDATA instapath=2 /", "bad logical"),
    ("&MCDATA n=3", "not terminated"),
    ("&RPIDATA n=3 /", "not found"),
    ("&MCDATA n= /", "missing value"),
    ("&MCDATA n='3' /", "string given"),
])
def test_namelist_errors(text, msg):
    with pytest.raises(ValueError, match=msg):
        I.mcdata(text)


def test_rpidata_defaults_and_rules():
    nml = I.rpidata("&RPIDATA n=64, beta=3200.0, angular=.true., fixedends=.false. /")
    assert nml["betan"] == 50.0  # rpi_ser.f90:45: beta/n
    assert nml["fixedends"] is True  # rpi_ser.f90:53
    assert nml["npoints"] == 10 and nml["cutofftheta"] == 6.5 and nml["cutoffphi"] == 3.2


# ---------------------------------------------------------------- list-directed files

def test_list_reader_records():
    r = I.ListReader(["1 2", "3, 4 5", "2*7 8", "  9  "])
    assert r.read(3) == ["1", "2", "3"]      # spans records, drops "4 5"
    assert r.read(2) == ["7", "7"]           # repeat count, drops "8"
    r.skip_record()
    with pytest.raises(EOFError):
        r.read(1)


def test_masses_wells_units(tmp_path):
    (tmp_path / "masses.dat").write_text("He 7294.3\nO, 29156.9\n")
    labels, m = I.read_masses(tmp_path / "masses.dat", 2)
    assert labels == ["H", "O"]  # character(len=1) labels keep the first letter
    assert np.array_equal(m, [7294.3, 29156.9])
    (tmp_path / "w1").write_text("1.0 2.0 3.0 extra\n4.0d0 5.0 6.0\n")
    (tmp_path / "w2").write_text("0.5\n0.25 0.125\n1 2 3\n")
    w1, w2 = I.read_wells(tmp_path / "w1", tmp_path / "w2", 3, 2, xunit=2)
    assert np.array_equal(w1, np.array([[1, 2, 3], [4, 5, 6]]) / 0.529177)
    assert np.array_equal(w2, np.array([[0.5, 0.25, 0.125], [1, 2, 3]]) / 0.529177)
    w1b, _ = I.read_wells(tmp_path / "w1", tmp_path / "w2", 3, 2, xunit=1)
    assert np.array_equal(w1b, [[1, 2, 3], [4, 5, 6]])


# ---------------------------------------------------------------- alignment

def _signed_volume(x, a, b, c, d):
    return float(np.linalg.det(np.stack([x[b] - x[a], x[c] - x[a], x[d] - x[a]])))


def _dists(x):
    return np.linalg.norm(x[:, None] - x[None], axis=-1)


@pytest.mark.parametrize("seed, atoms", [(0, (1, 2, 3)), (1, (3, 1, 5)), (2, (5, 4, 2))])
def test_alignment_geometry(seed, atoms):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((6, 3)) * 2.0 + rng.standard_normal(3) * 5.0
    a1, a2, a3 = atoms
    t1, t2, t3, origin = I.get_align(x, a1, a2, a3)
    assert np.array_equal(origin, x[a1 - 1])
    y = I.align_atoms(x, t1, t2, t3, a1)
    assert np.allclose(y[a1 - 1], 0.0, atol=1e-14)
    assert y[a2 - 1][0] > 0 and abs(y[a2 - 1][1]) < 1e-13 and abs(y[a2 - 1][2]) < 1e-13
    assert abs(y[a3 - 1][1]) < 1e-13 and y[a3 - 1][2] >= 0.0
    assert np.allclose(_dists(y), _dists(x), rtol=1e-13, atol=1e-13)
    others = [i for i in range(6) if i not in (a1 - 1, a2 - 1, a3 - 1)]
    v0 = _signed_volume(x, a1 - 1, a2 - 1, a3 - 1, others[0])
    v1 = _signed_volume(y, a1 - 1, a2 - 1, a3 - 1, others[0])
    assert v1 == pytest.approx(v0, rel=1e-12)  # a proper rotation, no reflection


def test_rotate_atoms_convention():
    th = 0.7
    # the z rotation takes (cos th, sin th, 0) onto +x (what get_align relies on)
    v = I.rotate_atoms(np.array([[math.cos(th), math.sin(th), 0.0]]), 3, th)
    assert np.allclose(v, [[1.0, 0.0, 0.0]], atol=1e-15)
    # about x: (0, sin, cos) -> (0, 0, 1) with theta = -atan2(y, z)... applied as +angle
    w = I.rotate_atoms(np.array([[0.0, 0.0, 1.0]]), 1, th)
    assert np.allclose(w, [[0.0, math.sin(th), math.cos(th)]], atol=1e-15)
    assert np.array_equal(I.rotate_atoms(v, 2, 1e-11), v)  # below the 1e-10 threshold


def test_align_with_zero_angles_does_not_translate():
    """instantonmod.f90:283-305: the translation happens only inside the rotation branch."""
    x = np.array([[1.0, 2.0, 3.0], [2.5, 2.0, 3.0], [1.0, 2.0, 4.0]])
    t = I.get_align(x)
    assert all(abs(v) <= 1e-10 for v in t[:3])
    assert np.array_equal(I.align_atoms(x, *t[:3]), x)
    with pytest.raises(ValueError, match="ndim = 3"):
        I.get_align(np.zeros((3, 2)))


def _rotation(rng):
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def _write_xyz(path, frames, labels, scale=1.0):
    with open(path, "w") as f:
        for i, fr in enumerate(frames):
            f.write(f"  {len(fr)}\n frame {i} with a title longer than twenty-eight characters\n")
            for lab, row in zip(labels, fr):
                f.write(f"{lab} " + " ".join(f"{v * scale:.17e}" for v in row) + "\n")


def test_read_path_undoes_a_rigid_motion(tmp_path):
    """Frames = R base + t for one rotation R: read_path (first-frame angles, each frame's
    own atom1 removed) returns base_i - base_i[atom1] in the frame that aligns base_1."""
    rng = np.random.default_rng(5)
    natom, npath = 4, 6
    base = np.cumsum(rng.standard_normal((npath, natom, 3)) * 0.1, axis=0) + rng.standard_normal((natom, 3))
    Rm, t = _rotation(rng), rng.standard_normal(3) * 3
    moved = base @ Rm.T + t
    _write_xyz(tmp_path / "path.xyz", moved * 0.529177, ["C", "H", "H", "O"])
    pd = I.read_path(tmp_path / "path.xyz", npath, 3, natom, xunit=2)
    tb = I.get_align(base[0])
    want = np.stack([I.align_atoms(b, *tb[:3]) for b in base])
    assert np.allclose(pd.frames, want, atol=1e-11)
    # reaction coordinate: normalised cumulative distances between aligned frames
    d = np.linalg.norm((want[1:] - want[:-1]).reshape(npath - 1, -1), axis=1)
    assert np.allclose(pd.lampath, np.concatenate([[0], np.cumsum(d)]) / d.sum(), rtol=1e-13)
    # spline through every coordinate = scipy's natural cubic spline
    cs = scipy.interpolate.CubicSpline(pd.lampath, pd.frames, axis=0, bc_type="natural")
    for s in (0.0, 0.13, 0.5, 0.77, 1.0):
        assert np.allclose(pd.at(s), cs(s), atol=1e-12)
        assert np.allclose(pd.slope(s), cs(s, 1), atol=1e-10)
    beads = pd.beads(9)
    assert beads.shape == (3 * natom, 9)
    assert np.allclose(beads[:, 0], want[0].reshape(-1)) and np.allclose(beads[:, -1], want[-1].reshape(-1))


# ---------------------------------------------------------------- centring

def _peaked(xm, npath=41):
    lam = np.linspace(0.0, 1.0, npath)
    return lam, -((lam - xm) ** 2)


@pytest.mark.parametrize("xm", [0.42, 0.61, 0.5])
def test_findmiddle_and_centre(xm):
    lam, V = _peaked(xm)
    got = I.findmiddle(lam, V)
    assert abs(got - xm) < 2e-6
    new, x = I.centre_lampath(lam, V)
    assert x == got
    assert new[0] == pytest.approx(0.0, abs=1e-15) and new[-1] == pytest.approx(1.0, abs=1e-14)
    assert np.all(np.diff(new) > 0)
    # the maximum maps to 1/2: a s^2 + b s = xm at s = 1/2
    a, b = 2 - 4 * got, 4 * got - 1
    assert a * 0.25 + b * 0.5 == pytest.approx(got, abs=1e-15)
    if a == 0.0:
        assert np.array_equal(new, lam)


def test_findmiddle_needs_a_bracket():
    lam = np.linspace(0, 1, 11)
    with pytest.raises(ValueError, match="bracketed"):
        I.findmiddle(lam, lam ** 2)


# ---------------------------------------------------------------- restart files

def test_restart_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    natom, n = 3, 5
    x = rng.standard_normal((3 * natom, n)) * 1e3
    p = rng.standard_normal((3 * natom, n)) * 1e-7
    f = tmp_path / R.restart_filename(0, 7)
    assert f.name == "restart_proc0_7.xyz"
    R.write_restart(f, x, p, 123456, -0.1 / 3, ["O", "H", "H"], cursor=98765, kick_cursor=12)
    st = R.read_restart(f, n, natom)
    assert np.array_equal(st.x, x) and np.array_equal(st.p, p)
    assert st.dhdr == -0.1 / 3 and st.nmc == 123456
    assert st.cursor == 98765 and st.kick_cursor == 12
    # the record layout the reference reads: per bead natom / dHdr / natom x (label x y z),
    # then natom / ii / natom x (label px py pz)
    lines = f.read_text().splitlines()
    assert len(lines) == 2 * n * (2 + natom)
    assert lines[0].split() == ["3"] and float(lines[1]) == -0.1 / 3
    assert lines[2].split()[0] == "O" and len(lines[2].split()) == 4
    mom = lines[n * (2 + natom):]
    assert mom[1].split()[0] == "123456"


def test_restart_reads_reference_style_files(tmp_path):
    """A file as gfortran's list-directed write lays it out (leading blanks, E exponents),
    with no cursors: read as written, cursors unknown."""
    rows = []
    for i in range(2):
        rows += ["           1", "   1.2500000000000000     ",
                 f" H  {0.5 + i:.16E}  -2.0000000000000000E-003   3.0000000000000000     "]
    for i in range(2):
        rows += ["           1", "        4000", f" H   {1e-5 * (i + 1):.16E}   0.0000000000000000        -1.0000000000000000     "]
    f = tmp_path / "restart_proc0_1.xyz"
    f.write_text("\n".join(rows) + "\n")
    st = R.read_restart(f, 2, 1)
    assert st.dhdr == 1.25 and st.nmc == 4000 and st.cursor is None
    assert np.array_equal(st.x, [[0.5, 1.5], [-2e-3, -2e-3], [3.0, 3.0]])
    assert np.array_equal(st.p, [[1e-5, 2e-5], [0.0, 0.0], [-1.0, -1.0]])


def _reference_write_points(nmc, noutput, imin, thermostat):
    """The reference's loops (verletmodule.f90:202-207, 389-395) stepped one by one."""
    pts, count = [], 0
    for ii in range(1, nmc + 1):
        if thermostat == 2:
            count += 1
            if count >= noutput and ii > imin:
                count = 0
                pts.append((ii - 1, ii))
        elif ii % noutput == 0:
            pts.append((ii - 1, ii))
    return pts + [(nmc, nmc)]


@pytest.mark.parametrize("nmc, noutput, imin", [(100, 10, 0), (100, 10, 35), (97, 7, 3),
                                                (50, 60, 0), (40, 1, 5), (64, 8, 64)])
@pytest.mark.parametrize("thermostat", [1, 2])
def test_write_points(nmc, noutput, imin, thermostat):
    assert R.write_points(nmc, noutput, imin, thermostat) == \
        _reference_write_points(nmc, noutput, imin, thermostat)


# ---------------------------------------------------------------- deck set-up

class _HostPes:
    """A test double of Pes for the set-up code: V = sum (x - c)^2 on the host."""

    def __init__(self, ndof, c):
        self.ndof, self.c, self.V0 = ndof, np.asarray(c, float), 0.0

    def V_init(self):
        self.V0 = 0.0

    def set_V0(self, v):
        self.V0 = float(v)

    def V(self, x):
        x = np.asarray(x, float)
        if x.ndim == 1:
            return float(np.sum((x - self.c) ** 2)) - self.V0
        return np.sum((x - self.c[:, None]) ** 2, axis=0) - self.V0


def _deck_files(d, w1, w2, frames, labels, masses):
    with open(d / "masses.dat", "w") as f:
        for lab, m in zip(labels, masses):
            f.write(f"{lab} {m!r}\n")
    for name, w in (("well1.dat", w1), ("well2.dat", w2)):
        with open(d / name, "w") as f:
            for row in w:
                f.write(" ".join(repr(float(v)) for v in row) + "\n")
    _write_xyz(d / "path.xyz", frames, labels)


def test_pimd_deck_setup(tmp_path):
    rng = np.random.default_rng(11)
    natom = 3
    w1 = rng.standard_normal((natom, 3))
    w2 = w1 + 0.3 * rng.standard_normal((natom, 3))
    frames = np.stack([w1 + t * (w2 - w1) + 0.02 * np.sin(np.pi * t) for t in np.linspace(0, 1, 7)])
    _deck_files(tmp_path, w1, w2, frames, ["O", "H", "H"], [29156.9, 1837.2, 1837.2])
    pes = _HostPes(3 * natom, np.zeros(3 * natom))
    text = "&MCDATA n=16, beta=2000.0, npath=7, nintegral=4, nrep=2, natom=3 /"
    deck = driver.pimd_deck(tmp_path, pes, text=text)
    t = I.get_align(w1)
    w1a = I.align_atoms(w1, *t[:3])
    assert np.array_equal(deck.well1, w1a)
    assert np.array_equal(deck.well2, I.align_atoms(w2, *t[:3]))  # well1's angles
    assert pes.V0 == float(np.sum(w1a ** 2))  # V0 = V(well1) after alignment
    tt = driver.ti_tasks(deck)
    xi, w = ti.gauleg(0.0, 1.0, 4)
    assert np.array_equal(tt["xi"], xi) and len(tt["tasks"]) == 8
    assert [(q["node"], q["rep"]) for q in tt["tasks"]][:3] == [(0, 0), (0, 1), (1, 0)]
    cs = scipy.interpolate.CubicSpline(deck.path.lampath, deck.path.frames.reshape(7, -1),
                                       axis=0, bc_type="natural")
    assert np.allclose(tt["ends"], cs(xi), atol=1e-12)
    assert np.allclose(tt["grads"], cs(xi, 1), atol=1e-10)
    assert np.array_equal(tt["start"], deck.path.frames[0].reshape(-1))


def test_pimd_deck_centre_and_errors(tmp_path):
    # atom 2 moves out along x from atom 1 through a barrier in the distance r12 at 1.9;
    # the first frame is already aligned, so alignment leaves the frames as they are
    r = np.linspace(1.0, 3.0, 21)
    frames = np.stack([[[0.0, 0.0, 0.0], [ri, 0.0, 0.0], [0.0, 0.0, 1.0]] for ri in r])
    _deck_files(tmp_path, frames[0], frames[-1], frames, ["C", "H", "O"], [1.0, 1.0, 1.0])

    class Barrier(_HostPes):
        def V(self, x):
            x = np.asarray(x, float).reshape(-1, 3)
            return -(np.linalg.norm(x[1] - x[0]) - 1.9) ** 2 - self.V0

    pes = Barrier(9, np.zeros(9))
    deck = driver.pimd_deck(tmp_path, pes, text="&MCDATA npath=21, natom=3, centre=.true. /")
    assert np.array_equal(deck.path.frames, frames)
    xm = deck.path.extra["xmiddle"]
    assert abs(xm - 0.45) < 2e-6
    lam = deck.path.lampath
    assert lam[0] == 0.0 and lam[-1] == pytest.approx(1.0, abs=1e-14)
    # the frame at the barrier top (r = 1.9, frame 9 of 21, lambda 0.45) now sits at 1/2
    a, b = 2 - 4 * xm, 4 * xm - 1
    s9 = -0.5 * b / a + math.sqrt(0.45 / a + (0.5 * b / a) ** 2)
    assert lam[9] == pytest.approx(s9, rel=1e-14) and abs(s9 - 0.5) < 1e-5
    with pytest.raises(NotImplementedError):
        driver.pimd_deck(tmp_path, pes, text="&MCDATA npath=21, natom=3, readhess=.true. /")
    with pytest.raises(ValueError, match="dof"):
        driver.pimd_deck(tmp_path, _HostPes(6, np.zeros(6)), text="&MCDATA npath=21, natom=3 /")
    with pytest.raises(ValueError, match="alignment atoms"):
        driver.pimd_deck(tmp_path, pes, text="&MCDATA npath=21, natom=3, atom3=4 /")


def test_rpi_deck_linear_chain(tmp_path):
    (tmp_path / "masses.dat").write_text("H 1836.0\n")
    (tmp_path / "well1.dat").write_text("-0.7\n")
    (tmp_path / "well2.dat").write_text("0.7\n")
    pes = _HostPes(1, [0.0])
    deck = driver.rpi_deck(tmp_path, pes, text="&RPIDATA n=5, ndim=1, readpath=.false. /")
    assert np.allclose(deck.xtilde, [[-0.7, -0.35, 0.0, 0.35, 0.7]], atol=1e-16)
    assert pes.V0 == pytest.approx(0.49, rel=1e-15)


def test_orientation_ibeta_cap():
    # rpi_ser.f90:291-295: the 3-angle rows cap I(beta) at 1 once omega N > 1; the ndim = 2
    # rows (:341) keep tanh(omega N)
    assert driver.orientation_ibeta(0.125, 12, 3) == 1.0           # omega N = 1.5
    assert driver.orientation_ibeta(0.125, 12, 2) == math.tanh(1.5)  # 0.905...
    assert driver.orientation_ibeta(0.05, 12, 3) == math.tanh(0.05 * 12)
    assert driver.orientation_ibeta(0.25 / 3, 12, 3) == math.tanh(0.25 / 3 * 12)  # omega N <= 1


def test_orientation_grid_and_rotations():
    nml = I.rpidata("&RPIDATA npoints=3, cutoffphi=1.0 /")
    g = driver.orientation_grid(nml)
    assert len(g) == 27
    th, wth = ti.gauleg(0.0, min(2 * ti.PI_REF, 6.5), 3)
    ph, wph = ti.gauleg(0.0, 1.0, 3)
    assert g[5][0] == (th[0], ph[1], th[2]) and g[5][1] == wth[0] * wph[1] * wth[2]
    assert sum(w for _, w in g) == pytest.approx(2 * ti.PI_REF * 1.0 * 2 * ti.PI_REF, rel=1e-13)
    # rotation order of rpi_ser.f90:258-261: eta about x, then phi about z, then theta about x
    w = np.array([[0.3, -0.2, 0.9], [1.0, 0.5, -0.4]])
    ang = (0.4, 1.1, 2.3)

    def Rx(t):
        c, s = math.cos(t), math.sin(t)
        return np.array([[1, 0, 0], [0, c, s], [0, -s, c]])

    def Rz(t):
        c, s = math.cos(t), math.sin(t)
        return np.array([[c, s, 0], [-s, c, 0], [0, 0, 1]])

    want = w @ (Rx(ang[0]) @ Rz(ang[1]) @ Rx(ang[2])).T
    assert np.allclose(driver.rotated_well(w, ang), want, atol=1e-15)
