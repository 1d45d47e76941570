"""The reference's own RPI driver (rpi_ser.f90) linked against the drop-in modules and run on
the device, with the device plugin and with a host plugin written to the reference's API.

oracle/ref/build_dropin.sh compiles rpi_ser.f90 from the reference's source (where the
reference is present) into oracle/_ref/rpi_dropin_device and rpi_dropin_host; the programs
travel with the tree.  Each reads an RPIDATA namelist, masses.dat, well1.dat, well2.dat and
the potential (rp_pes.dat), finds the instanton from a straight line (rpi_ser.f90:146-166,
with instanton's m = 8, factr = 1e6, pgtol = eps2 = 1e-5), writes it to instanton.xyz and
prints lndetJ, lndetJ0, s_kink and the splitting (:354-381).  Checks, on the 27-dof EVB:
* lndetJ0 (every bead at well 1, singlewell): the oracle's band, numpy's eigvalsh, 1e-9
  (device plugin; 1e-6 for the host plugin, whose Hessian is differenced as the
  reference's plugins difference it);
* the printed instanton: the oracle's U_M gives s_kink (1e-12), its gradient is at the
  L-BFGS tolerance, and the minimum Newton's method reaches from it lies below it by the
  quadratic model's drop (what the loose stopping rule leaves, here a few 1e-6);
* lndetJ at the printed instanton: eigvalsh of the oracle's band, 1e-8;
* the splitting: rpi_ser.f90:368-381's formula on the printed numbers, 1e-12;
* the host plugin's run agrees with the device plugin's.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from helpers import oracle_pes
from rptunnel import specs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = {k: os.path.join(ROOT, "oracle", "_ref", f"rpi_dropin_{k}") for k in ("device", "host")}
BOHR = 0.529177
N, BETA = 16, 4000.0


def _deck(tmp, sp):
    with open(os.path.join(tmp, "masses.dat"), "w") as f:
        for m in sp.mass:
            f.write(f"H {float(m)!r}\n")
    for name, w in (("well1.dat", sp.well1), ("well2.dat", sp.well2)):
        with open(os.path.join(tmp, name), "w") as f:
            for a in range(sp.natom):
                f.write(" ".join(repr(float(v)) for v in w[3 * a:3 * a + 3]) + "\n")
    with open(os.path.join(tmp, "rp_pes.dat"), "w") as f:
        f.write(f"{sp.kind} {sp.ndim} {sp.natom} {len(sp.dp)} {len(sp.ip)}\n")
        f.write(" ".join(repr(float(v)) for v in sp.dp) + "\n")
        f.write(" ".join(str(int(v)) for v in sp.ip) + "\n")
    return (f"&RPIDATA\n n={N}, beta={BETA!r}, ndim=3, natom={sp.natom}, readpath=.false.,"
            " fixedends=.true.\n/\n")


def _run(kind, tmp, sp):
    exe = BIN[kind]
    if not os.path.exists(exe):
        pytest.skip(f"{exe} is built only where the reference's sources are (oracle/ref/build_dropin.sh)")
    res = subprocess.run([exe], input=_deck(tmp, sp), cwd=tmp, capture_output=True, text=True,
                         timeout=300)
    assert res.returncode == 0, res.stdout[-3000:] + res.stderr[-3000:]
    out = res.stdout

    def num(tag, k=0):
        m = re.search(re.escape(tag) + r"\s*([-+0-9.EeDd]+)(?:\s+([-+0-9.EeDd]+))?", out)
        assert m, (tag, out[-2000:])
        return float(m.group(1 + k).replace("D", "E").replace("d", "e"))

    frames = []
    with open(os.path.join(tmp, "instanton.xyz")) as f:
        lines = f.read().split("\n")
    for i in range(N):
        blk = lines[i * (sp.natom + 2) + 2: (i + 1) * (sp.natom + 2)]
        frames.append(np.array([[float(t) for t in ln.split()[1:4]] for ln in blk]).reshape(-1) / BOHR)
    # V0 = V(well1) (rpi_ser.f90:95), or the lowest bead's V when the path dips below it
    # (:181-191)
    v0 = num("New V0=") if "New V0=" in out else None
    return dict(lndetj=num("lndetJ"), lndetj0=num("lndetJ", 1), skink=num("s_kink="),
                phi=num("phi="), delta=num("RPI delta="), x=np.array(frames).T, v0=v0)


@pytest.fixture(scope="module")
def evb():
    sp = specs.evb()
    return sp, oracle_pes(sp)


def _newton_action(sp, op, x, betan):
    """Newton on the exact U_M Hessian (dense; n * ndof = 432) from x -> (the minimum's
    action, the first step's Newton decrement g.H^-1 g / 2, the model's action drop)."""
    n, D = x.shape[1], x.shape[0]
    oc = O.Chain(n, sp.ndim, sp.natom, sp.mass, betan, True)
    m = np.repeat(sp.mass, sp.ndim)
    dec = None
    for _ in range(4):
        g, e = oc.UMforceenergy(op, x, sp.well1, sp.well2)
        H = np.zeros((n * D, n * D))
        for i in range(n):
            Hb = op.Vdoubleprime(x[:, i])
            idx = np.arange(D) * n + i
            H[np.ix_(idx, idx)] += Hb
        for p in range(D):
            k = m[p] / betan ** 2
            for i in range(n):
                H[p * n + i, p * n + i] += 2.0 * k
                if i + 1 < n:
                    H[p * n + i, p * n + i + 1] -= k
                    H[p * n + i + 1, p * n + i] -= k
        step = np.linalg.solve(H, g.reshape(-1))
        if dec is None:
            dec = 0.5 * float(g.reshape(-1) @ step)
        x = x - step.reshape(D, n)
    return oc.UM(op, x, sp.well1, sp.well2), dec


@pytest.mark.parametrize("kind", ["device", "host"])
def test_reference_rpi_driver_on_the_device(kind, evb, tmp_path):
    sp, op = evb
    r = _run(kind, str(tmp_path), sp)
    betan = BETA / N  # rpi_ser.f90:45
    oc = O.Chain(N, sp.ndim, sp.natom, sp.mass, betan, True)
    # lndetJ0: every bead at well 1, singlewell (rpi_ser.f90:221-237)
    xw = np.repeat(sp.well1[:, None], N, axis=1)
    ev0 = O.detj_eigs(oc.UMhessian(op, xw, True))
    tol0 = 1e-9 if kind == "device" else 1e-6
    assert abs(r["lndetj0"] - O.lndetj0(ev0)) <= tol0 * abs(O.lndetj0(ev0)), (r["lndetj0"], O.lndetj0(ev0))
    # the instanton it printed, with the driver's V0
    op = oracle_pes(sp, r["v0"] if r["v0"] is not None else op.V(sp.well1))
    x = r["x"]
    um = oc.UM(op, x, sp.well1, sp.well2)
    assert abs(r["skink"] - betan * um) <= 1e-12 * abs(betan * um) + 1e-12
    g, _ = oc.UMforceenergy(op, x, sp.well1, sp.well2)
    assert np.abs(g).max() <= 1e-4, np.abs(g).max()
    # the minimum lies below by what instanton's stopping rule (factr = 1e6, pgtol = 1e-5)
    # leaves: the quadratic model's drop from the printed chain
    s_star, dec = _newton_action(sp, op, x.copy(), betan)
    assert s_star <= um and abs(um - s_star) <= 1e-5 * abs(s_star), (um, s_star)
    assert abs((um - s_star) - dec) <= 0.1 * dec + 1e-15, (um - s_star, dec)
    # lndetJ at that instanton (rpi_ser.f90:354-365)
    ev = O.detj_eigs(oc.UMhessian(op, x, False))
    tol = 1e-8 if kind == "device" else 1e-6
    assert abs(r["lndetj"] - O.lndetj(ev)) <= tol * abs(O.lndetj(ev)), (r["lndetj"], O.lndetj(ev))
    # the splitting from the printed numbers (rpi_ser.f90:366-381)
    phi, theta, delta = O.rpi_splitting(betan, r["skink"], r["lndetj"], r["lndetj0"])
    assert r["phi"] == pytest.approx(phi, rel=1e-12)
    assert r["delta"] == pytest.approx(delta, rel=1e-12)


def test_host_and_device_plugins_agree(evb, tmp_path):
    sp, _ = evb
    (tmp_path / "d").mkdir()
    (tmp_path / "h").mkdir()
    rd = _run("device", str(tmp_path / "d"), sp)
    rh = _run("host", str(tmp_path / "h"), sp)
    assert rh["skink"] == pytest.approx(rd["skink"], rel=1e-7)
    assert rh["lndetj0"] == pytest.approx(rd["lndetj0"], rel=1e-6)
    assert rh["lndetj"] == pytest.approx(rd["lndetj"], rel=1e-6)
