"""Input decks of the reference's drivers: namelists, masses.dat, well*.dat, path.xyz.

Everything here is host-side set-up that runs once per job at O(npath * ndof): it produces
the fixed ends, the TI endpoints b(xi) and db/dxi, and the initial chains that the device
engine consumes.  It follows the reference's readers statement by statement so that the
same deck gives the same numbers:

* ``read_namelist`` / ``mcdata`` / ``rpidata``: the MCDATA namelist with the defaults of
  pimd_par.f90:42-88 and the RPIDATA namelist of rpi_ser.f90:22-44 / rpi_par.f90:27-56;
* ``read_masses`` (pimd_par.f90:155-159), ``read_wells`` (pimd_par.f90:120-136);
* ``rotate_atoms`` / ``get_align`` / ``align_atoms`` (instantonmod.f90:222-312, 346-376);
* ``read_path`` (instantonmod.f90:895-1035 without its instanton branch, which the
  drivers run on the device) and ``centre_lampath`` / ``findmiddle``
  (instantonmod.f90:832-871, 985-998);
* ``spline`` / ``splint`` / ``splin_grad`` (instantonmod.f90:397-561) for the path.

Coordinates are kept as (natom, ndim) arrays, one row per atom as the files hold them; the
flat dof index used by the engine is atom * ndim + dim, i.e. ``atoms.reshape(-1)``.
Lengths are bohr after reading (xunit = 2 means the file is in angstrom,
pimd_par.f90:131-136).
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field

import numpy as np

BOHR_PER_ANGSTROM_DIV = 0.529177  # the reference divides angstrom values by this

# ---------------------------------------------------------------- namelists

# MCDATA (pimd_par.f90:42-45) and its defaults (pimd_par.f90:60-88).  npath has no default
# there; it must be given when a path is read.
MCDATA_DEFAULTS = {
    "n": 100, "beta": 100.0, "nmc": 5000000, "noutput": 100000, "dt": 1.0e-3,
    "iprint": False, "imin": 0, "tau": 1.0, "npath": None, "gamma": 1.0, "nintegral": 5,
    "nrep": 1, "use_mkl": False, "thermostat": 1, "ndim": 3, "natom": 1, "xunit": 1,
    "instapath": False, "centre": False, "dhdrlimit": -1.0, "readpath": True,
    "alignwell": False, "fixedends": True, "cayley": False, "readhess": False,
    "basename": "", "atom1": 1, "atom2": 2, "atom3": 3, "restart": 0,
}

# RPIDATA: the union of rpi_ser.f90:22-24 and rpi_par.f90:27-28, defaults from
# rpi_ser.f90:28-42.  rpi_ser leaves cutofftheta, cutoffphi, atom1-3 and basename
# unset; they take rpi_par.f90:43-51's values here.
RPIDATA_DEFAULTS = {
    "n": 100, "beta": 30.0, "ndim": 3, "natom": 1, "npath": 0, "xunit": 1,
    "angular": False, "npoints": 10, "cutofftheta": 6.5, "cutoffphi": 3.2,
    "output_instanton": False, "readpath": True, "alignwell": False, "fixedends": True,
    "alignpath": True, "checkhess": False, "basename": "", "atom1": 1, "atom2": 2,
    "atom3": 3,
}

# declared types where the default is None
_TYPES = {"npath": int}


def _strip_comments(text):
    out = []
    for line in text.splitlines():
        buf, q = [], None
        for ch in line:
            if q:
                buf.append(ch)
                if ch == q:
                    q = None
            elif ch in "'\"":
                q = ch
                buf.append(ch)
            elif ch == "!":
                break
            else:
                buf.append(ch)
        out.append("".join(buf))
    return "\n".join(out)


_TOKEN = re.compile(r"""\s*(?:'((?:[^']|'')*)'|"((?:[^"]|"")*)"|([^\s,=/'"]+)|(=)|(,)|(/))""")


def _group_body(text, group):
    """The text between '&group' (or '$group') and the closing '/' (or '$end')."""
    text = _strip_comments(text)
    m = re.search(r"[&$]" + re.escape(group) + r"\b", text, re.IGNORECASE)
    if not m:
        raise ValueError(f"namelist group &{group} not found")
    body = text[m.end():]
    q = None
    for i, ch in enumerate(body):
        if q:
            if ch == q:
                q = None
        elif ch in "'\"":
            q = ch
        elif ch == "/":
            return body[:i]
        elif ch in "&$" and re.match(r"[&$]end\b", body[i:], re.IGNORECASE):
            return body[:i]
    raise ValueError(f"namelist group &{group} is not terminated by '/'")


def _convert(name, raw, quoted, typ):
    if typ is str:
        return raw
    if quoted:
        raise ValueError(f"namelist item {name}: string given for a {typ.__name__}")
    if typ is bool:
        t = raw.lower().lstrip(".")
        if t[:1] == "t":
            return True
        if t[:1] == "f":
            return False
        raise ValueError(f"namelist item {name}: bad logical {raw!r}")
    if typ is int:
        if not re.fullmatch(r"[+-]?\d+", raw):
            raise ValueError(f"namelist item {name}: bad integer {raw!r}")
        return int(raw)
    if typ is float:
        try:
            return float(raw.replace("d", "e").replace("D", "e"))
        except ValueError:
            raise ValueError(f"namelist item {name}: bad real {raw!r}") from None
    raise TypeError(typ)


def read_namelist(text, group, defaults):
    """Fortran namelist input for ``group`` over ``defaults`` (names case-insensitive).

    Names absent from the group keep their defaults; a name the group does not declare, or
    a value of the wrong type, is an error, as it is for the reference's ``read(*, nml=)``.
    Integer items accept integer literals only (the reference's ``NMC= (5e6)`` is an
    assignment in code, not namelist input)."""
    body = _group_body(text, group)
    out = dict(defaults)
    toks = []
    pos = 0
    while pos < len(body):
        m = _TOKEN.match(body, pos)
        if not m or m.end() == pos:
            if body[pos:].strip() == "":
                break
            raise ValueError(f"namelist &{group}: cannot parse near {body[pos:pos + 20]!r}")
        pos = m.end()
        if m.group(1) is not None:
            toks.append(("str", m.group(1).replace("''", "'")))
        elif m.group(2) is not None:
            toks.append(("str", m.group(2).replace('""', '"')))
        elif m.group(3) is not None:
            toks.append(("word", m.group(3)))
        elif m.group(4):
            toks.append(("eq", "="))
    i = 0
    while i < len(toks):
        kind, name = toks[i]
        if kind != "word":
            raise ValueError(f"namelist &{group}: expected 'name = value' at {name!r}")
        if i + 1 >= len(toks) or toks[i + 1][0] != "eq":
            raise ValueError(f"namelist &{group}: expected '=' after {name!r}")
        key = name.lower()
        if key not in defaults:
            raise ValueError(f"namelist &{group}: unknown item {name!r}")
        if i + 2 >= len(toks) or toks[i + 2][0] == "eq":
            raise ValueError(f"namelist &{group}: missing value for {name!r}")
        vkind, raw = toks[i + 2]
        typ = _TYPES.get(key) or type(defaults[key])
        out[key] = _convert(key, raw, vkind == "str", typ)
        i += 3
    return out


def mcdata(text):
    """The MCDATA namelist of the PIMD-TI driver (pimd_par.f90:42-94), plus the derived
    betan = beta/(n+1) (pimd_par.f90:94)."""
    nml = read_namelist(text, "MCDATA", MCDATA_DEFAULTS)
    nml["betan"] = nml["beta"] / (nml["n"] + 1)
    return nml


def rpidata(text):
    """The RPIDATA namelist of the instanton driver, plus betan = beta/n (rpi_ser.f90:45)
    and the rule that angular integration forces fixed ends (rpi_ser.f90:53)."""
    nml = read_namelist(text, "RPIDATA", RPIDATA_DEFAULTS)
    nml["betan"] = nml["beta"] / nml["n"]
    if nml["angular"]:
        nml["fixedends"] = True
    return nml


# ---------------------------------------------------------------- list-directed files

def _fortran_float(tok):
    return float(tok.replace("d", "e").replace("D", "e"))


class ListReader:
    """Fortran list-directed reading over the lines of a text file: each ``read`` starts at
    a new record, takes values across records as needed, and discards what is left of its
    last record.  Separators are blanks and commas; ``r*c`` repeats c r times."""

    def __init__(self, lines, name="<file>"):
        self.lines = list(lines)
        self.pos = 0
        self.name = name

    @classmethod
    def open(cls, path):
        with open(path) as f:
            return cls(f.read().splitlines(), str(path))

    def _record(self):
        if self.pos >= len(self.lines):
            raise EOFError(f"{self.name}: end of file after {self.pos} records")
        line = self.lines[self.pos]
        self.pos += 1
        toks = []
        for t in re.split(r"[\s,]+", line.strip()):
            if not t:
                continue
            m = re.fullmatch(r"(\d+)\*(.+)", t)
            if m:
                toks += [m.group(2)] * int(m.group(1))
            else:
                toks.append(t)
        return toks

    def read(self, count):
        """The next ``count`` raw items of one read statement."""
        out = []
        while len(out) < count:
            out += self._record()
        return out[:count]

    def skip_record(self):
        """A formatted read of one record whose contents are thrown away."""
        self._record()


def read_masses(path, natom):
    """masses.dat: one ``label mass`` line per atom (pimd_par.f90:155-159).  The reference's
    labels are single characters (``character label(natom)``), so a label keeps only its
    first character, as the list-directed read into it does."""
    r = ListReader.open(path)
    labels, mass = [], np.zeros(natom)
    for j in range(natom):
        lab, m = r.read(2)
        labels.append(lab.strip("'\"")[:1])
        mass[j] = _fortran_float(m)
    return labels, mass


def _to_bohr(x, xunit):
    return x / BOHR_PER_ANGSTROM_DIV if xunit == 2 else x


def read_wells(path1, path2, ndim, natom, xunit=1):
    """well1.dat / well2.dat: ndim values per atom line (pimd_par.f90:120-136), converted
    from angstrom when xunit = 2.  Returns two (natom, ndim) arrays in bohr."""
    r1, r2 = ListReader.open(path1), ListReader.open(path2)
    w1, w2 = np.zeros((natom, ndim)), np.zeros((natom, ndim))
    for j in range(natom):
        w1[j] = [_fortran_float(t) for t in r1.read(ndim)]
        w2[j] = [_fortran_float(t) for t in r2.read(ndim)]
    return _to_bohr(w1, xunit), _to_bohr(w2, xunit)


def read_xyz_frames(path, nframes, natom, ndim, xunit=1):
    """``nframes`` xyz frames: a count record, a title record, then ``label x ...`` per atom
    with ndim coordinates read (instantonmod.f90:914-925).  Returns (nframes, natom, ndim)
    in bohr."""
    r = ListReader.open(path)
    out = np.zeros((nframes, natom, ndim))
    for i in range(nframes):
        r.read(1)
        r.skip_record()
        for j in range(natom):
            toks = r.read(1 + ndim)
            out[i, j] = [_fortran_float(t) for t in toks[1:]]
    return _to_bohr(out, xunit)


def write_xyz(path, frames, labels, titles=None, to_angstrom=True):
    """xyz frames as the reference writes aligned.xyz / instanton.xyz
    (instantonmod.f90:973-980, 1024-1031): coordinates times 0.529177 when converted."""
    frames = np.asarray(frames, dtype=np.float64)
    scale = BOHR_PER_ANGSTROM_DIV if to_angstrom else 1.0
    with open(path, "w") as f:
        for i, fr in enumerate(frames):
            f.write(f"{fr.shape[0]}\n")
            f.write(("" if titles is None else str(titles[i])) + "\n")
            for j, row in enumerate(fr):
                f.write(labels[j] + " " + " ".join(repr(float(v * scale)) for v in row) + "\n")


# ---------------------------------------------------------------- alignment

def rotate_atoms(atoms, axis, theta):
    """Rotation of every atom by theta about axis 1, 2 or 3 (instantonmod.f90:346-376): the
    matrix has R[a,a] = 1, R[j,j] = R[k,k] = cos, R[k,j] = -sin, R[j,k] = sin over the two
    other axes j < k, and angles within 1e-10 of zero leave the atoms untouched."""
    atoms = np.array(atoms, dtype=np.float64)
    if abs(theta) <= 1e-10:
        return atoms
    a = axis - 1
    j, k = [i for i in range(3) if i != a]
    R = np.zeros((3, 3))
    R[a, a] = 1.0
    R[j, j] = R[k, k] = math.cos(theta)
    R[k, j] = -math.sin(theta)
    R[j, k] = math.sin(theta)
    return atoms @ R.T


def _require_3d(atoms):
    if atoms.shape[1] != 3:
        raise ValueError("alignment needs ndim = 3 (instantonmod.f90:233-236 stops otherwise)")


def get_align(atoms, atom1=1, atom2=2, atom3=3):
    """Angles that put atom1 at the origin, atom2 on the +x axis and atom3 in the xz plane
    (instantonmod.f90:222-265): theta1 about z, then theta2 about y, then theta3 about x.
    Atom numbers are 1-based as in the namelist.  Returns (theta1, theta2, theta3, origin)."""
    atoms = np.asarray(atoms, dtype=np.float64)
    _require_3d(atoms)
    if not all(1 <= a <= len(atoms) for a in (atom1, atom2, atom3)):
        raise ValueError(f"alignment atoms {atom1}, {atom2}, {atom3} must be among the "
                         f"{len(atoms)} atoms of the frame")
    origin = atoms[atom1 - 1].copy()
    x = atoms - origin
    w = x[atom2 - 1] - x[atom1 - 1]
    t1 = math.atan2(w[1], w[0])
    x = rotate_atoms(x, 3, t1)
    w = x[atom2 - 1] - x[atom1 - 1]
    t2 = math.atan2(w[2], w[0])
    x = rotate_atoms(x, 2, t2)
    w = x[atom3 - 1] - x[atom1 - 1]
    t3 = -math.atan2(w[1], w[2])
    return t1, t2, t3, origin


def align_atoms(atoms, theta1, theta2, theta3, atom1=1):
    """Apply get_align's angles (instantonmod.f90:269-312).  As in the reference, the
    translation removes this frame's own atom1 (not the origin of get_align), and when all
    three angles are within 1e-10 of zero the frame is returned unchanged, untranslated."""
    atoms = np.asarray(atoms, dtype=np.float64)
    _require_3d(atoms)
    if not any(abs(t) > 1e-10 for t in (theta1, theta2, theta3)):
        return atoms.copy()
    x = atoms - atoms[atom1 - 1]
    x = rotate_atoms(x, 3, theta1)
    x = rotate_atoms(x, 2, theta2)
    return rotate_atoms(x, 1, theta3)


def eucliddist(x1, x2):
    """instantonmod.f90:876-890."""
    return math.sqrt(float(np.sum((np.asarray(x1) - np.asarray(x2)) ** 2)))


# ---------------------------------------------------------------- splines

def locate(xx, x):
    """Bisection index of x in the monotone table xx, 1-based, with the reference's end
    rules (instantonmod.f90:565-603)."""
    n = len(xx)
    ascnd = xx[-1] >= xx[0]
    jl, ju = 0, n + 1
    while ju - jl > 1:
        jm = (ju + jl) // 2
        if ascnd == (x >= xx[jm - 1]):
            jl = jm
        else:
            ju = jm
    if x == xx[0]:
        return 1
    if x == xx[-1]:
        return n - 1
    return jl


def spline(x, y):
    """Natural cubic spline second derivatives (instantonmod.f90:397-495 with the 1e31
    end-slope flags, which select the natural end conditions), by the reference's
    tridiagonal sweep."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    n = len(x)
    a, b, c, r, gam, y2 = (np.zeros(n) for _ in range(6))
    c[: n - 1] = x[1:] - x[:-1]
    r[: n - 1] = 6.0 * ((y[1:] - y[:-1]) / c[: n - 1])
    r[1: n - 1] = r[1: n - 1] - r[: n - 2].copy()
    a[1: n - 1] = c[: n - 2]
    b[1: n - 1] = 2.0 * (c[1: n - 1] + a[1: n - 1])
    b[0] = b[n - 1] = 1.0
    r[0] = c[0] = 0.0
    r[n - 1] = a[n - 1] = 0.0
    bet = b[0]
    y2[0] = r[0] / bet
    for j in range(1, n):
        gam[j] = c[j - 1] / bet
        bet = b[j] - a[j] * gam[j]
        y2[j] = (r[j] - a[j] * y2[j - 1]) / bet
    for j in range(n - 2, -1, -1):
        y2[j] = y2[j] - gam[j + 1] * y2[j + 1]
    return y2


def _bracket(xa, x):
    n = len(xa)
    klo = max(min(locate(xa, x), n - 1), 1) - 1
    khi = klo + 1
    h = xa[khi] - xa[klo]
    return klo, khi, h, (xa[khi] - x) / h, (x - xa[klo]) / h


def splint(xa, ya, y2a, x):
    """Spline value at x (instantonmod.f90:499-528); ya/y2a may carry trailing axes."""
    klo, khi, h, a, b = _bracket(xa, x)
    ya, y2a = np.asarray(ya), np.asarray(y2a)
    return a * ya[klo] + b * ya[khi] + ((a ** 3 - a) * y2a[klo] + (b ** 3 - b) * y2a[khi]) * h * h / 6.0


def splin_grad(xa, ya, y2a, x):
    """Spline slope at x (instantonmod.f90:532-561)."""
    klo, khi, h, a, b = _bracket(xa, x)
    ya, y2a = np.asarray(ya), np.asarray(y2a)
    return (ya[khi] - ya[klo]) / h + ((1.0 - 3.0 * a * a) * y2a[klo] + (3.0 * b * b - 1.0) * y2a[khi]) * h / 6.0


# ---------------------------------------------------------------- paths

@dataclass
class PathData:
    """A path of npath frames: lampath[npath] (normalised arc length), frames
    (npath, natom, ndim), the spline second derivatives of every coordinate, Vpath when a
    potential was given, and the alignment angles."""
    lampath: np.ndarray
    frames: np.ndarray
    y2: np.ndarray
    Vpath: np.ndarray | None = None
    angles: tuple = (0.0, 0.0, 0.0)
    extra: dict = field(default_factory=dict)

    @property
    def npath(self):
        return len(self.lampath)

    def flat(self):
        """(ndof, npath): the layout of rp_batch_set_path / rp_chain_set_path."""
        return np.ascontiguousarray(self.frames.reshape(self.npath, -1).T)

    def at(self, s):
        """Spline position of every coordinate at reaction coordinate s, (natom, ndim)."""
        return splint(self.lampath, self.frames, self.y2, s)

    def slope(self, s):
        return splin_grad(self.lampath, self.frames, self.y2, s)

    def beads(self, n):
        """xtilde: the n-bead chain at s = (k-1)/(n-1) (instantonmod.f90:945-947), as
        (ndof, n) in the engine's layout."""
        return np.stack([self.at(k / (n - 1)).reshape(-1) for k in range(n)], axis=1)


def arc_lampath(frames):
    """Cumulative eucliddist between consecutive frames, normalised to end at 1
    (instantonmod.f90:926-936)."""
    lam = np.zeros(len(frames))
    for i in range(1, len(frames)):
        lam[i] = lam[i - 1] + eucliddist(frames[i - 1], frames[i])
    return lam / lam[-1]


def spline_columns(lam, frames):
    """spline() of every coordinate of frames (npath, ...) over lam."""
    flat = frames.reshape(len(lam), -1)
    y2 = np.zeros_like(flat)
    for d in range(flat.shape[1]):
        y2[:, d] = spline(lam, flat[:, d])
    return y2.reshape(frames.shape)


def read_path(path, npath, ndim, natom, xunit=1, atom1=1, atom2=2, atom3=3, V=None,
              align=True):
    """path.xyz -> PathData (instantonmod.f90:895-949): frames converted to bohr, the
    alignment angles taken from the first frame and applied to every frame, the reaction
    coordinate the normalised cumulative distance between aligned frames, natural splines
    through every coordinate.  V (a callable on a flat ndof vector) fills Vpath.  The
    reference aligns unconditionally and so needs ndim = 3; align=False skips the
    alignment for models of other dimension."""
    raw = read_xyz_frames(path, npath, natom, ndim, xunit)
    if align:
        t1, t2, t3, _ = get_align(raw[0], atom1, atom2, atom3)
        frames = np.stack([align_atoms(fr, t1, t2, t3, atom1) for fr in raw])
    else:
        t1 = t2 = t3 = 0.0
        frames = raw
    lam = arc_lampath(frames)
    Vp = None if V is None else np.array([float(V(fr.reshape(-1))) for fr in frames])
    return PathData(lam, frames, spline_columns(lam, frames), Vp, (t1, t2, t3))


def path_from_chain(chain, well1=None, well2=None, V=None):
    """The path rebuilt from an instanton chain (ndof, n) (instantonmod.f90:958-990): with
    fixed ends the wells are prepended and appended (npath = n + 2), otherwise the beads
    alone (npath = n)."""
    chain = np.asarray(chain, dtype=np.float64)
    shape = (-1,) + (well1.shape if well1 is not None else (chain.shape[0],))
    frames = chain.T.reshape(shape)
    if well1 is not None:
        frames = np.concatenate([well1[None], frames, well2[None]])
    lam = arc_lampath(frames)
    Vp = None if V is None else np.array([float(V(fr.reshape(-1))) for fr in frames])
    return PathData(lam, frames, spline_columns(lam, frames), Vp)


def findmiddle(lampath, Vpath, x1=0.3, x2=0.7, xacc=1e-6, jmax=40):
    """Bisection for the zero of dV/dlambda on the spline of Vpath between x1 and x2
    (instantonmod.f90:832-871); errors where the reference stops."""
    y2 = spline(lampath, Vpath)
    fmid = splin_grad(lampath, Vpath, y2, x2)
    f = splin_grad(lampath, Vpath, y2, x1)
    if f * fmid >= 0.0:
        raise ValueError("findmiddle: root must be bracketed (the path's energy has no "
                         f"single maximum between {x1} and {x2})")
    if f < 0.0:
        rtb, dx = x1, x2 - x1
    else:
        rtb, dx = x2, x1 - x2
    for _ in range(jmax):
        dx *= 0.5
        xmid = rtb + dx
        fmid = splin_grad(lampath, Vpath, y2, xmid)
        if fmid <= 0.0:
            rtb = xmid
        if abs(dx) < xacc or fmid == 0.0:
            return rtb
    raise ValueError("findmiddle: too many bisections")


def centre_lampath(lampath, Vpath):
    """Re-map the reaction coordinate so that the energy maximum sits at 1/2
    (instantonmod.f90:985-998): with xm = findmiddle, a = 2 - 4 xm and b = 4 xm - 1, each
    lambda becomes the root of a s^2 + b s = lambda that runs from 0 to 1.  For xm = 1/2
    exactly (a = 0) the reference divides by zero; here the map takes its limit there,
    s = lambda.  Returns (new lampath, xm)."""
    xm = findmiddle(lampath, Vpath)
    a = 2.0 - 4.0 * xm
    b = 4.0 * xm - 1.0
    lam = np.asarray(lampath, dtype=np.float64)
    if a == 0.0:
        return lam.copy(), xm
    h = 0.5 * b / a
    root = np.sqrt(lam / a + h * h)
    out = -h + root if a >= 0.0 else -h - root
    return out, xm


def centred(pd: PathData):
    """PathData with the centred reaction coordinate and re-fitted splines."""
    if pd.Vpath is None:
        raise ValueError("centring needs Vpath (read the path with a potential)")
    lam, xm = centre_lampath(pd.lampath, pd.Vpath)
    out = PathData(lam, pd.frames, spline_columns(lam, pd.frames), pd.Vpath, pd.angles,
                   dict(pd.extra, xmiddle=xm))
    return out
