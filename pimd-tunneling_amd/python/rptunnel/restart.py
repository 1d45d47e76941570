"""Restart files of the PIMD-TI driver (write_restart, verletmodule.f90:162-185; the reader
of pimd_par.f90:327-370).

One text file per task, ``restart_proc<iproc>_<ii>.xyz``.  For every bead: a record with
natom, a record with the running TI **sum** dHdr, then natom records ``label x y z``.  Then
for every bead: natom, the step count, and ``label px py pz``.  The format has three
coordinate columns, so ndim must be 3.

Resume semantics (restart = 2): positions, momenta, the running sum and the step count
``restartnmc`` are read back; init_path is skipped; the new run keeps adding to the sum,
and the integrand is divided by NMC + restartnmc - imin (verletmodule.f90:247, 413).

One extension, invisible to the reference's reader: the step-count record may carry two
more integers, the Philox cursors of the normal and of the Poisson draws, after the count.
A list-directed read of one integer ignores the rest of its record, so the reference still
reads these files; this package reads the cursors back so that a resumed trajectory
continues its random stream instead of replaying it.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .inputs import ListReader, _fortran_float


def restart_filename(iproc, ii):
    """pimd_par.f90:336-352: "restart_proc" // iproc // "_" // ii // ".xyz" with minimal
    integer widths (the reference's format covers iproc < 1000 and ii < 1000 only)."""
    return f"restart_proc{int(iproc)}_{int(ii)}.xyz"


@dataclass
class RestartState:
    x: np.ndarray        # (ndof, n), dof = atom*3 + dim
    p: np.ndarray        # (ndof, n)
    dhdr: float          # running TI sum
    nmc: int             # restartnmc: the step count written with the momenta
    cursor: int | None = None        # Philox cursor of the normal draws
    kick_cursor: int | None = None   # Philox cursor of the Poisson draws


def _fmt(v):
    return repr(float(v))


def write_restart(path, x, p, ii, dhdr, labels, cursor=None, kick_cursor=None):
    """write_restart (verletmodule.f90:162-185) for one chain.  x, p: (3*natom, n).  The
    floats are written with round-trip precision, so reading the file back restores the
    state exactly."""
    x = np.asarray(x, dtype=np.float64)
    p = np.asarray(p, dtype=np.float64)
    natom = len(labels)
    if x.shape[0] != 3 * natom or p.shape != x.shape:
        raise ValueError("write_restart: the format needs ndim = 3 and x, p of shape "
                         f"(3*natom, n) = ({3 * natom}, n); got {x.shape}, {p.shape}")
    n = x.shape[1]
    tail = ""
    if cursor is not None:
        tail = f" {int(cursor)} {int(kick_cursor or 0)}"
    lines = []
    for i in range(n):
        lines.append(f"{natom}")
        lines.append(_fmt(dhdr))
        for j in range(natom):
            lines.append(f"{labels[j]} " + " ".join(_fmt(v) for v in x[3 * j:3 * j + 3, i]))
    for i in range(n):
        lines.append(f"{natom}")
        lines.append(f"{int(ii)}{tail}")
        for j in range(natom):
            lines.append(f"{labels[j]} " + " ".join(_fmt(v) for v in p[3 * j:3 * j + 3, i]))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_restart(path, n, natom):
    """The restart = 2 reader (pimd_par.f90:355-370): the dHdr and step count are those of
    the last bead block read, as the reference's loop leaves them."""
    r = ListReader.open(path)
    x = np.zeros((3 * natom, n))
    p = np.zeros((3 * natom, n))
    dhdr = 0.0
    for i in range(n):
        r.read(1)
        dhdr = _fortran_float(r.read(1)[0])
        for j in range(natom):
            toks = r.read(4)
            x[3 * j:3 * j + 3, i] = [_fortran_float(t) for t in toks[1:]]
    nmc, cursor, kcur = 0, None, None
    for i in range(n):
        r.read(1)
        rec = r._record()
        nmc = int(rec[0])
        if len(rec) >= 3:
            cursor, kcur = int(rec[1]), int(rec[2])
        for j in range(natom):
            toks = r.read(4)
            p[3 * j:3 * j + 3, i] = [_fortran_float(t) for t in toks[1:]]
    return RestartState(x, p, dhdr, nmc, cursor, kcur)


def write_points(nmc, noutput, imin, thermostat):
    """The steps at which the reference writes a restart file during one run, as
    (state after `done` steps, label) pairs, ending with the final write.

    Langevin (verletmodule.f90:389-395): a counter that reaches noutput on a step past imin
    writes, before that step runs, the state after ii-1 steps labelled ii.  Andersen
    (verletmodule.f90:205-207): every ii with ii mod noutput = 0, also before the step.
    Both write the state after nmc steps labelled nmc at the end (:246, :412); labels are
    offset by restartnmc by the caller."""
    if noutput <= 0 and thermostat != 2:
        raise ValueError("Andersen restart writes need noutput > 0 (mod(ii, Noutput))")
    every = max(int(noutput), 1)
    if thermostat == 2:
        # the counter reaches noutput at ii = noutput but cannot fire until ii > imin;
        # it restarts from zero after each write
        first = max(every, imin + 1)
    else:
        first = every
    pts = [(ii - 1, ii) for ii in range(first, nmc + 1, every)]
    pts.append((nmc, nmc))
    return pts
