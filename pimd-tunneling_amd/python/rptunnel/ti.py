"""Thermodynamic-integration driver arithmetic (host side of pimd_par.f90).

Gauss-Legendre nodes (verletmodule.f90:124-160), the task table and the final
quadrature / error / q/q0 (pimd_par.f90:243-308, 321-424), and the splitting
formulas (rpi_ser.f90:368-381; multiwell_waterdimer.py:6-16).  The trajectories
themselves run in the batched engine (``Batch``); this module only places the
lambda nodes and reduces the integrands, as the reference's driver does.
"""
from __future__ import annotations

import math

import numpy as np

PI_REF = 3.14159265358979  # instantonmod.f90:4


def gauleg(x1, x2, n):
    """Gauss-Legendre nodes and weights on [x1, x2] (verletmodule.f90:124-160)."""
    EPS = 3.0e-14
    x = np.zeros(n)
    w = np.zeros(n)
    m = (n + 1) // 2
    xm, xl = 0.5 * (x2 + x1), 0.5 * (x2 - x1)
    for i in range(1, m + 1):
        z = math.cos(PI_REF * (i - 0.25) / (n + 0.5))
        while True:
            p1, p2 = 1.0, 0.0
            for j in range(1, n + 1):
                p3 = p2
                p2 = p1
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j
            pp = n * (z * p1 - p2) / (z * z - 1.0)
            z1 = z
            z = z1 - p1 / pp
            if abs(z - z1) <= EPS:
                break
        x[i - 1] = xm - xl * z
        x[n - i] = xm + xl * z
        w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp)
        w[n - i] = w[i - 1]
    return x, w


def task_table(nintegral, nrep):
    """lambda-major task order of pimd_par.f90:245-255: task k = nrep*(i-1) + j."""
    return [(i, j) for i in range(nintegral) for j in range(nrep)]


def skip_task(endpoint):
    """pimd_par.f90:322-325: a node whose endpoint is within 1e-10 of the origin in every
    coordinate is not run and contributes integrand 0."""
    return bool(np.all(np.abs(endpoint) < 1e-10))


def quadrature(integrands, weights, nrep, betan):
    """pimd_par.f90:397-424: mean and population variance over the nrep replicas of each
    node, Delta A = sum w_i mean_i, sigma_A^2 = sum w_i^2 var_i, q/q0 = exp(-Delta A betan)
    and its error betan sqrt(sigma_A^2/nrep) q/q0."""
    I = np.asarray(integrands, dtype=np.float64).reshape(len(weights), nrep)
    mean = I.sum(axis=1) / nrep
    var = (I ** 2).sum(axis=1) / nrep - mean ** 2
    answer = float(np.sum(np.where(np.isnan(mean), 0.0, np.asarray(weights) * mean)))
    sigmaA = float(np.sum(var * np.asarray(weights) ** 2))
    q = math.exp(-answer * betan)
    return {"deltaA": answer, "sigmaA": sigmaA, "q_ratio": q,
            "q_err": betan * math.sqrt(sigmaA / nrep) * q}


def splitting_two_beta(I1, I2, beta1, beta2):
    """multiwell_waterdimer.py:6-10: Delta = 2 (artanh I2 - artanh I1) / (beta2 - beta1)."""
    return 2.0 * (math.atanh(I2) - math.atanh(I1)) / (beta2 - beta1)


def splitting_single_beta(I, beta):
    """Two-level limit of the above at one temperature: Delta = 2 artanh(I) / beta."""
    return 2.0 * math.atanh(I) / beta


def rpi_splitting(betan, skink, lndetj, lndetj0):
    """rpi_ser.f90:368-381 -> (phi, theta, delta) with the reference's pi literal."""
    phi = math.exp(0.5 * (lndetj - lndetj0))
    theta = betan * math.exp(-skink) * math.sqrt(skink / (2.0 * PI_REF)) / phi
    return phi, theta, 2.0 * theta / betan


def run_ti(batch_factory, spec, n, beta, nintegral, nrep, nsteps, imin, path=None, **kw):
    """PIMD-TI on the batched engine with the reference driver's semantics.

    batch_factory(polymers) -> Batch.  Every (node, replica) task of pimd_par.f90 becomes
    one polymer; the integrand is dHdr/betan^2 with dHdr the running TI mean
    (pimd_par.f90:379).  Returns the quadrature dict plus the raw integrands."""
    from .specs import arc_lampath, linear_path

    betan = beta / (n + 1)
    if path is None:
        lampath, pth = linear_path(spec.well1, spec.well2, npath=2)
    else:
        pth = np.asarray(path, dtype=np.float64)
        lampath = arc_lampath(pth)
    xi, w = gauleg(0.0, 1.0, nintegral)
    # endpoints b(xi) and db/dxi from the path spline (pimd_par.f90:214-221)
    ends, grads = _spline_nodes(lampath, pth, xi)
    start = pth[:, 0]
    polymers, keep = [], []
    for k, (i, j) in enumerate(task_table(nintegral, nrep)):
        if skip_task(ends[i]):
            continue
        polymers.append(dict(n=n, betan=betan, xi=float(xi[i]), a=start, b=ends[i], dbdl=grads[i],
                             stream=k))
        keep.append(k)
    integrands = np.zeros(nintegral * nrep)
    if polymers:
        b = batch_factory(polymers)
        b.set_path(lampath, pth)
        b.init_path()
        b.run(nsteps, imin=imin)
        sums, cnt, _ = b.ti_readout()
        for k, s in zip(keep, sums):
            integrands[k] = s / (nsteps - imin) / betan ** 2
    res = quadrature(integrands, w, nrep, betan)
    res["integrands"] = integrands
    res["nodes"] = xi
    res["weights"] = w
    return res


def run_ti_plain(pes, spec, n, beta, nintegral, nrep, nsteps, imin, tau, dt, gamma, seed,
                 path=None):
    """The same TI tasks as run_ti, one trajectory at a time on the plain per-trajectory
    path (rp_chain), as the reference's task loop runs them (pimd_par.f90:321-381): for each
    task, init_nm with the node's ends, init_path from the task's Philox stream (seed, k),
    propagate_pimd_pile with TI after imin steps.  With the batched engine's seed the two
    paths draw the same numbers, so their integrands agree to rounding.  Returns the
    quadrature dict plus the integrands and the wall seconds of the trajectories."""
    import time

    from .specs import arc_lampath, linear_path
    from .verlet import Chain

    betan = beta / (n + 1)
    if path is None:
        lampath, pth = linear_path(spec.well1, spec.well2, npath=2)
    else:
        pth = np.asarray(path, dtype=np.float64)
        lampath = arc_lampath(pth)
    xi, w = gauleg(0.0, 1.0, nintegral)
    ends, grads = _spline_nodes(lampath, pth, xi)
    start = pth[:, 0]
    integrands = np.zeros(nintegral * nrep)
    wall = 0.0
    for k, (i, j) in enumerate(task_table(nintegral, nrep)):
        if skip_task(ends[i]):
            continue
        ch = Chain(pes, n, spec.mass, betan, tau, dt, gamma, start, ends[i])
        ch.set_path(lampath, pth)
        ch.set_rng(seed, k)
        t0 = time.perf_counter()
        ch.init_path(float(xi[i]))
        dh, _ = ch.propagate_pimd_pile(nsteps, grads[i], imin=imin, xi=float(xi[i]))
        wall += time.perf_counter() - t0
        ch.close()
        integrands[k] = dh / betan ** 2
    res = quadrature(integrands, w, nrep, betan)
    res["integrands"] = integrands
    res["nodes"] = xi
    res["weights"] = w
    res["seconds"] = wall
    return res


def _spline_nodes(lampath, path, xi):
    """splint / splin_grad of every dof at the nodes (instantonmod.f90:499-561);
    path is (ndof, npath)."""
    from .inputs import spline_columns, splin_grad, splint

    P = np.ascontiguousarray(np.asarray(path, dtype=np.float64).T)
    y2 = spline_columns(lampath, P)
    ends = np.stack([splint(lampath, P, y2, x) for x in xi])
    grads = np.stack([splin_grad(lampath, P, y2, x) for x in xi])
    return ends, grads
