"""The reference's two drivers over input decks: PIMD-TI (pimd_par.f90 / pimd_ser.f90) and
ring-polymer instanton rates (rpi_ser.f90 / rpi_par.f90).

A deck is a working directory with masses.dat, well1.dat, well2.dat and, when readpath is
set, path.xyz, plus the namelist text (the reference reads it from standard input).  The
set-up (alignment, V0, the path, the TI nodes) is host work done once; every trajectory,
action, gradient, Hessian and determinant runs in the device library, all tasks of the job
at once: the TI tasks of every (node, replica) pair are the polymers of one ``Batch``, and
the instanton chains (one, or one per orientation for angular integration) are the chains
of one ``Ring``.

Differences from the reference, all in set-up: alignment is skipped when ndim is not 3
(the reference stops there), a missing path with readpath = .false. in the TI driver is the
straight line between the wells (the reference leaves it unallocated), and readhess is
rejected (the reference's read_hess splines the path coordinates, not the Hessians,
instantonmod.f90:1065-1068).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import inputs as I
from . import restart as R
from . import ti

LANGEVIN, ANDERSEN = 2, 1


@dataclass
class Deck:
    nml: dict
    labels: list
    mass: np.ndarray              # natom
    well1: np.ndarray             # (natom, ndim), aligned, bohr
    well2: np.ndarray
    path: I.PathData | None = None
    xtilde: np.ndarray | None = None    # initial instanton chain (ndof, n)
    log: list = field(default_factory=list)

    @property
    def ndim(self):
        return self.nml["ndim"]

    @property
    def natom(self):
        return self.nml["natom"]

    @property
    def ndof(self):
        return self.ndim * self.natom


def _read_text(workdir, text, name):
    if text is not None:
        return text
    with open(os.path.join(workdir, name)) as f:
        return f.read()


def _base_deck(workdir, nml):
    p = lambda f: os.path.join(workdir, f)  # noqa: E731
    labels, mass = I.read_masses(p("masses.dat"), nml["natom"])
    w1, w2 = I.read_wells(p("well1.dat"), p("well2.dat"), nml["ndim"], nml["natom"],
                          nml["xunit"])
    return labels, mass, w1, w2


def _align_wells(nml, w1, w2, always):
    """pimd_par.f90:160-166 (always: well1's angles for both wells, or well2's own with
    alignwell) and rpi_ser.f90:73-80 (only with alignwell, each well its own angles)."""
    if nml["ndim"] != 3 or not (always or nml["alignwell"]):
        return w1, w2
    a = (nml["atom1"], nml["atom2"], nml["atom3"])
    t1, t2, t3, _ = I.get_align(w1, *a)
    w1 = I.align_atoms(w1, t1, t2, t3, a[0])
    if nml["alignwell"]:
        t1, t2, t3, _ = I.get_align(w2, *a)
    w2 = I.align_atoms(w2, t1, t2, t3, a[0])
    return w1, w2


def _V(pes):
    return lambda x: pes.V(np.asarray(x, dtype=np.float64).reshape(-1))


def _read_deck_path(workdir, nml, V):
    if nml["npath"] is None or nml["npath"] < 2:
        raise ValueError("readpath needs npath >= 2 frames in path.xyz")
    return I.read_path(os.path.join(workdir, "path.xyz"), nml["npath"], nml["ndim"],
                       nml["natom"], nml["xunit"], nml["atom1"], nml["atom2"], nml["atom3"],
                       V=V, align=nml["ndim"] == 3)


# ---------------------------------------------------------------- PIMD-TI

def pimd_deck(workdir, pes, text=None, ring_options=None):
    """pimd_par.f90:60-231: namelist, V_init, wells, masses, alignment, V0 = V(well1), the
    path (with the instanton and centring options) and the natural splines through it."""
    nml = I.mcdata(_read_text(workdir, text, "input"))
    if nml["readhess"]:
        raise NotImplementedError("readhess: the reference's read_hess does not spline the "
                                  "Hessians it computes; not supported")
    if pes.ndof != nml["ndim"] * nml["natom"]:
        raise ValueError(f"potential has {pes.ndof} dof, the deck {nml['ndim']}x{nml['natom']}")
    pes.V_init()
    labels, mass, w1, w2 = _base_deck(workdir, nml)
    w1, w2 = _align_wells(nml, w1, w2, always=True)
    pes.set_V0(pes.V(w1.reshape(-1)))
    deck = Deck(nml, labels, mass, w1, w2)
    V = _V(pes)
    if nml["readpath"]:
        pd = _read_deck_path(workdir, nml, V)
        if nml["instapath"]:
            pd = _instanton_path(deck, pes, pd, ring_options or {})
        if nml["centre"]:
            pd = I.centred(pd)
            deck.log.append(f"Centred, xmiddle = {pd.extra['xmiddle']!r}")
    else:
        pd = I.path_from_chain(np.stack([w1.reshape(-1), w2.reshape(-1)], axis=1), V=V)
    deck.path = pd
    return deck


def _instanton_path(deck, pes, pd, opts):
    """read_path's instanton branch (instantonmod.f90:951-990): optimise the n-bead chain
    interpolated on the path (device L-BFGS), then rebuild the path from it."""
    from .instanton import Ring

    nml = deck.nml
    n, betan = nml["n"], nml["betan"]
    fixed = nml["fixedends"]
    xt = pd.beads(n)
    ring = Ring(pes, deck.mass, [n], [betan], deck.well1.reshape(-1), deck.well2.reshape(-1),
                fixedends=fixed)
    xo, act, it, st = ring.instanton([xt], **opts)
    deck.log.append(f"Found instanton: U_M = {act[0]!r}, iterations {it[0]}, status {st[0]}")
    ring.close()
    return I.path_from_chain(xo[0], deck.well1 if fixed else None,
                             deck.well2 if fixed else None, V=_V(pes))


def ti_tasks(deck):
    """The task table of pimd_par.f90:237-255: Gauss-Legendre nodes on [0, 1], endpoints
    b(xi) = path(xi) and db/dxi, start = the path's first frame; task k = nrep*i + j."""
    nml = deck.nml
    xi, w = ti.gauleg(0.0, 1.0, nml["nintegral"])
    pd = deck.path
    ends = np.stack([pd.at(x).reshape(-1) for x in xi])
    grads = np.stack([pd.slope(x).reshape(-1) for x in xi])
    start = pd.frames[0].reshape(-1)
    tasks = []
    for k, (i, j) in enumerate(ti.task_table(nml["nintegral"], nml["nrep"])):
        tasks.append(dict(k=k, node=i, rep=j, xi=float(xi[i]), b=ends[i], dbdl=grads[i],
                          skip=ti.skip_task(ends[i])))
    return dict(xi=xi, weights=w, start=start, ends=ends, grads=grads, tasks=tasks)


def run_pimd(deck, pes, workdir=None, seed=0, iproc=0, nmc=None, launch=None):
    """The TI task loop and final quadrature of pimd_par.f90:316-424 on the batched engine.

    Every non-skipped task is one polymer.  restart >= 1 writes restart_proc<iproc>_<ii>.xyz
    (ii = task number from 1) at the reference's points of the run; restart = 2 resumes
    every task from its file.  Returns the quadrature dict of ti.quadrature plus the per-task
    integrands, nodes and weights."""
    from .batch import Batch

    nml = deck.nml
    n, betan = nml["n"], nml["betan"]
    NMC = int(nml["nmc"] if nmc is None else nmc)
    imin, restart, thermostat = nml["imin"], nml["restart"], nml["thermostat"]
    if thermostat not in (ANDERSEN, LANGEVIN):
        raise ValueError("Incorrect thermostat option.")  # pimd_par.f90:380-383
    if restart >= 1 and deck.ndim != 3:
        raise ValueError("restart files hold three coordinates per atom: ndim must be 3")
    tt = ti_tasks(deck)
    run = [t for t in tt["tasks"] if not t["skip"]]
    integrands = np.zeros(len(tt["tasks"]))
    fname = lambda t: os.path.join(workdir or ".", R.restart_filename(iproc, t["k"] + 1))  # noqa: E731
    if run:
        polys = [dict(n=n, betan=betan, xi=t["xi"], a=tt["start"], b=t["b"], dbdl=t["dbdl"],
                      stream=t["k"]) for t in run]
        b = Batch(pes, deck.mass, polys, nml["tau"], nml["dt"], gamma=nml["gamma"],
                  thermostat=thermostat, noutput=nml["noutput"], dhdrlimit=nml["dhdrlimit"],
                  cayley=nml["cayley"], seed=seed)
        if launch:
            b.set_launch(**launch)
        b.set_path(deck.path.lampath, deck.path.flat())
        dhdr0 = np.zeros(len(run))
        rnmc = np.zeros(len(run), dtype=np.int64)
        if restart < 2:
            b.init_path()
        else:
            cur = np.zeros(len(run), dtype=np.int64)
            kcur = np.zeros(len(run), dtype=np.int64)
            have = True
            for q, t in enumerate(run):
                st = R.read_restart(fname(t), n, deck.natom)
                b.set_state(q, st.x, st.p)
                dhdr0[q], rnmc[q] = st.dhdr, st.nmc
                if st.cursor is None:
                    have = False
                else:
                    cur[q], kcur[q] = st.cursor, st.kick_cursor
            if have:
                b.set_cursors(cur, kcur)
        if restart >= 1:
            done = 0
            for upto, label in R.write_points(NMC, nml["noutput"], imin, thermostat):
                if upto > done:
                    b.run(upto - done, imin=imin)
                    done = upto
                s, _, _ = b.ti_readout()
                cur, kcur = b.cursors(), b.kick_cursors()
                for q, t in enumerate(run):
                    x, p = b.get_state(q)
                    R.write_restart(fname(t), x, p, label + rnmc[q], dhdr0[q] + s[q],
                                    deck.labels, cur[q], kcur[q])
        else:
            b.run(NMC, imin=imin)
        s, cnt, nre = b.ti_readout()
        for q, t in enumerate(run):
            # verletmodule.f90:247, 413 and pimd_par.f90:387
            integrands[t["k"]] = (dhdr0[q] + s[q]) / float(NMC + rnmc[q] - imin) / betan ** 2
        b.close()
    res = ti.quadrature(integrands, tt["weights"], nml["nrep"], betan)
    res.update(integrands=integrands, nodes=tt["xi"], weights=tt["weights"])
    return res


# ---------------------------------------------------------------- instanton rates

def rpi_deck(workdir, pes, text=None):
    """rpi_ser.f90:22-170 set-up: namelist, masses, wells (aligned only with alignwell),
    V0 = V(well1), and the initial chain xtilde, from path.xyz or the straight line."""
    nml = I.rpidata(_read_text(workdir, text, "input"))
    if pes.ndof != nml["ndim"] * nml["natom"]:
        raise ValueError(f"potential has {pes.ndof} dof, the deck {nml['ndim']}x{nml['natom']}")
    pes.V_init()
    labels, mass, w1, w2 = _base_deck(workdir, nml)
    w1, w2 = _align_wells(nml, w1, w2, always=False)
    pes.set_V0(pes.V(w1.reshape(-1)))
    deck = Deck(nml, labels, mass, w1, w2)
    n = nml["n"]
    if nml["readpath"]:
        deck.path = _read_deck_path(workdir, nml, _V(pes))
        deck.xtilde = deck.path.beads(n)
    else:
        # rpi_ser.f90:141-147: xtilde_k = ((n-k) well1 + (k-1) well2)/(n-1)
        k = np.arange(1, n + 1)
        deck.xtilde = (np.outer(w1.reshape(-1), n - k) + np.outer(w2.reshape(-1), k - 1)) / (n - 1)
    return deck


def run_rpi(deck, pes, opts=None):
    """rpi_ser.f90:170-381: the instanton (device L-BFGS), the V0 reset when the chain dips
    below well 1, lndetJ0 at well 1 (all eigenvalues, single well), and then either the
    splitting (phi, theta, h_ij, Delta) or, with angular = .true., the orientation
    quadrature of rpi_ser.f90:241-345, every orientation's instanton in one batch."""
    from .instanton import Ring

    nml = deck.nml
    n, betan, D = nml["n"], nml["betan"], deck.ndof
    w1, w2 = deck.well1.reshape(-1), deck.well2.reshape(-1)
    fixed = nml["fixedends"]
    opts = dict(opts or {})
    out = {}
    if nml["angular"]:
        return _run_rpi_angular(deck, pes, opts)
    ring = Ring(pes, deck.mass, [n], [betan], w1, w2, fixedends=fixed)
    xo, act, it, st = ring.instanton([deck.xtilde], **opts)
    out.update(instanton=xo[0], iterations=int(it[0]), status=int(st[0]))
    Vpath = np.asarray(pes.V(xo[0]))
    if Vpath.min() < 0.0:
        # rpi_ser.f90:179-186: V0 from the lowest bead
        i = int(np.argmin(Vpath))
        pes.set_V0(0.0)
        pes.set_V0(pes.V(xo[0][:, i].copy()))
        Vpath = np.asarray(pes.V(xo[0]))
    out["Vpath"] = Vpath
    out["lndetj0"] = _lndetj0(pes, deck, n, betan, w1, w2, fixed)
    ring.set_x(xo)
    ld, nneg, eta = ring.logdet(singlewell=False, skip_lowest=True)
    skink = betan * float(ring.UM(xo)[0])
    phi, theta, delta = ti.rpi_splitting(betan, skink, float(ld[0]), out["lndetj0"])
    out.update(lndetj=float(ld[0]), nneg=int(nneg[0]), eta1=float(eta[0]), skink=skink,
               phi=phi, theta=theta, n_theta=theta * n, h_ij=theta / betan, delta=delta)
    ring.close()
    return out


def _lndetj0(pes, deck, n, betan, w1, w2, fixed):
    """rpi_ser.f90:218-237: every bead at well 1, singlewell Hessian, ln over all positive
    eigenvalues."""
    from .instanton import Ring

    r0 = Ring(pes, deck.mass, [n], [betan], w1, w2, fixedends=fixed)
    r0.set_x([np.repeat(w1[:, None], n, axis=1)])
    ld0, _, _ = r0.logdet(singlewell=True, skip_lowest=False)
    r0.close()
    return float(ld0[0])


def orientation_grid(nml):
    """Gauss-Legendre orientations of rpi_ser.f90:246-250 (ndim = 3: theta, phi, eta on
    [0, min(2 pi, cutofftheta)], [0, min(pi, cutoffphi)], [0, min(2 pi, cutofftheta)]) or
    :304 (ndim = 2: eta only), with the reference's pi literal.  Returns a list of
    (angles, weight)."""
    pi = ti.PI_REF
    m = nml["npoints"]
    th, wth = ti.gauleg(0.0, min(2.0 * pi, nml["cutofftheta"]), m)
    if nml["ndim"] == 3:
        ph, wph = ti.gauleg(0.0, min(pi, nml["cutoffphi"]), m)
        et, wet = ti.gauleg(0.0, min(2.0 * pi, nml["cutofftheta"]), m)
        return [((th[i], ph[j], et[k]), wth[i] * wph[j] * wet[k])
                for i in range(m) for j in range(m) for k in range(m)]
    if nml["ndim"] == 2:
        et, wet = ti.gauleg(0.0, min(2.0 * pi, nml["cutofftheta"]), m)
        return [((et[k],), wet[k]) for k in range(m)]
    raise ValueError("angular integration needs ndim = 2 or 3")


def orientation_ibeta(omega, n, ndim):
    """The I(beta) factor of one orientation: tanh(omega N), capped at 1 when omega N > 1 on
    the 3-angle grid (rpi_ser.f90:291-295); the 1-angle grid of ndim = 2 has no cap (:341)."""
    if ndim == 3 and omega * n > 1.0:
        return 1.0
    return math.tanh(omega * n)


def rotated_well(well2, angles):
    """rpi_ser.f90:258-261 (eta about x, phi about z, theta about x) and :307 (ndim = 2:
    eta about axis 1 of the plane)."""
    w = np.asarray(well2, dtype=np.float64)
    if len(angles) == 3:
        th, ph, et = angles
        w = I.rotate_atoms(w, 1, et)
        w = I.rotate_atoms(w, 3, ph)
        return I.rotate_atoms(w, 1, th)
    (et,) = angles
    if abs(et) <= 1e-10:
        return w.copy()
    # rotate_atoms about axis 1 on 2-D atoms: j = 2, k = 3 falls off the plane; the
    # reference's matmul with a 3x3 matrix on 2-vectors is not defined, so the planar
    # rotation R = [[c, s], [-s, c]] of the axis-3 case is used
    c, s = math.cos(et), math.sin(et)
    return w @ np.array([[c, s], [-s, c]]).T


def _run_rpi_angular(deck, pes, opts):
    from .instanton import Ring

    nml = deck.nml
    n, betan = nml["n"], nml["betan"]
    w1 = deck.well1.reshape(-1)
    grid = orientation_grid(nml)
    ends = [rotated_well(deck.well2, a).reshape(-1) for a, _ in grid]
    nc = len(grid)
    ring = Ring(pes, deck.mass, [n] * nc, [betan] * nc, np.repeat(w1[None], nc, axis=0),
                np.stack(ends), fixedends=True)
    xo, act, it, st = ring.instanton([deck.xtilde] * nc, **opts)
    lnd0 = _lndetj0(pes, deck, n, betan, w1, deck.well2.reshape(-1), True)
    ring.set_x(xo)
    ld, nneg, _ = ring.logdet(singlewell=False, skip_lowest=True)
    um = ring.UM(xo)
    ring.close()
    rows = []
    for c, (angles, wgt) in enumerate(grid):
        skink = betan * float(um[c])
        g = math.exp(0.5 * (float(ld[c]) - lnd0))
        omega = betan * math.exp(-skink) * math.sqrt(skink / (2.0 * ti.PI_REF)) / g
        rows.append(dict(angles=angles, weight=wgt, Ibeta=orientation_ibeta(omega, n, nml["ndim"]),
                         lndetj=float(ld[c]), skink=skink, status=int(st[c])))
    return dict(lndetj0=lnd0, rows=rows)
