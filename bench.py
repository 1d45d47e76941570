"""Headline benchmark: bead-steps/s of PIMD-TI on one MI355X, batched vs plain GPU path.

Workload (the 27-DOF EVB double-well batched-TI configuration): 9 atoms, 27 dof,
N in {32, 64, 96, 128, 192, 256} beads, 512 trajectories x 11 Gauss-Legendre lambda
nodes at every N (33,792 polymers of uneven size), beta = 2000, tau = 20, dt = 1,
gamma = 1e-3, Langevin (PILE) thermostat, FP64 throughout, synthetic seeded inputs.
A "step" is one MD step of every polymer (B A O A + TI accumulation).

Reported beside the batched value:
  * the plain per-trajectory GPU path (the drop-in, one trajectory per call sequence)
    timed on a bounded sample of the same trajectories (every N);
  * the roofline of the dominant fused-kernel launch (FP64 MFMA, 78.6 TFLOP/s), from
    hipEvent timing of the launches inside the timed region;
  * the CPU baseline: the oracle's restatement of the reference algorithm (dense
    transforms, reference step order) on the host cores, bounded sample at every N.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pimd-tunneling_amd", "python"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

METRIC = "bead-steps/s, batched vs plain GPU, at equal TI splitting error (1 GPU)"
NS = [32, 64, 96, 128, 192, 256]
NTRAJ, NLAMBDA = 512, 11
BETA, TAU, DT, GAMMA = 2000.0, 20.0, 1.0, 1e-3
PEAK_FP64 = 78.6e12


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def flops_per_bead_step(n, D):
    """FP64 work the kernel executes per bead-step (no credit for padding).  Transforms:
    for even N the folded DST-I (two N/2 x N/2 halves per transform, 4 N D for the two
    transforms of the 2D state columns) plus its pre-fold and recombination (4 D); for odd
    N the dense N x N transforms (8 N D).  Elementwise kick / rotations / O-step / shifts:
    21 D.  EVB potential: two wells, each y = W^T (x - xw) and W h (2 x 2 D^2) plus ~13 D."""
    w_pes = 8 * D * D + 26 * D
    transforms = 4 * n * D + 4 * D if n % 2 == 0 else 8 * n * D
    return transforms + 21 * D + w_pes


def build_polymers(sp, ns, ntraj, nlam):
    from rptunnel import specs

    xi, _ = gauleg(nlam)
    lampath, path = specs.linear_path(sp.well1, sp.well2, npath=2)
    pols = []
    stream = 0
    for n in ns:
        for il, x in enumerate(xi):
            b = sp.well1 + x * (sp.well2 - sp.well1)
            for r in range(ntraj):
                pols.append(dict(n=n, betan=BETA / (n + 1), xi=float(x), a=sp.well1, b=b,
                                 dbdl=sp.well2 - sp.well1, stream=stream))
                stream += 1
    return pols, lampath, path


def gauleg(n):
    from rptunnel import ti
    return ti.gauleg(0.0, 1.0, n)


def plain_sample(pes, sp, ns, steps):
    """Seconds per polymer-step of the plain path at each N (one trajectory, host loop
    over per-step launch sequences, as the reference driver steps one trajectory)."""
    from rptunnel import Chain

    out = {}
    dbdl = sp.well2 - sp.well1
    for n in ns:
        betan = BETA / (n + 1)
        b = sp.well1 + 0.5 * (sp.well2 - sp.well1)
        ch = Chain(pes, n, sp.mass, betan, TAU, DT, GAMMA, sp.well1, b)
        t = (np.arange(n) + 1.0) / (n + 1.0)
        x = np.outer(sp.well1, 1 - t) + np.outer(b, t)
        ch.set_state(x, np.zeros_like(x))
        ch.set_rng(7, n)
        ch.propagate_pimd_pile(5, dbdl)  # warm-up
        t0 = time.perf_counter()
        ch.propagate_pimd_pile(steps, dbdl)
        out[n] = (time.perf_counter() - t0) / steps
        ch.close()
    return out


MKL_RT = "/opt/conda/lib/libmkl_rt.so"


def host_cores():
    """The host threads this process may use: its CPU affinity, capped by OMP_NUM_THREADS
    (which a shared GPU host sets to the job's share)."""
    cores = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        cores = min(cores, int(env))
    return cores


def cpu_sample(sp, ns, budget_s, threads):
    """The oracle's restatement of the reference step on the host cores: seconds per
    polymer-step at each N.  It runs the -O3 build, with the transforms' mat-vec through
    MKL's sequential dsymv -- the call the reference makes (verletmodule.f90:259, 279,
    makefile:6) -- when the MKL runtime is on the host, else through the restatement's loop."""
    import ctypes

    import oracle as O

    blas = None
    if os.path.exists(MKL_RT):
        os.environ["MKL_THREADING_LAYER"] = "SEQUENTIAL"
        os.environ["MKL_NUM_THREADS"] = "1"
        blas = ctypes.cast(ctypes.CDLL(MKL_RT).dsymv_, ctypes.c_void_p).value
    O.use_blas_dsymv(blas)
    op = O.Pes(sp.kind, sp.ndim, sp.natom, sp.dp, sp.ip)
    out, detail = {}, {}
    for n in ns:
        betan = BETA / (n + 1)
        b = sp.well1 + 0.5 * (sp.well2 - sp.well1)
        t = (np.arange(n) + 1.0) / (n + 1.0)
        x0 = np.outer(sp.well1, 1 - t) + np.outer(b, t)
        # calibrate: one trajectory, few steps
        # calibrate on one trajectory: double the steps until the set-up no longer dominates
        cal = 4
        while True:
            t1, _ = O.bench_pile(op, n, sp.ndim, sp.natom, sp.mass, betan, TAU, DT, GAMMA,
                                 sp.well1, b, sp.well2 - sp.well1, x0, 1, cal, 1, 1, fast=True)
            if t1 > 0.05 * budget_s / len(ns) or cal >= 1 << 20:
                break
            cal *= 4
        per = t1 / cal
        steps = max(4, int(budget_s / len(ns) / per))
        wall, _ = O.bench_pile(op, n, sp.ndim, sp.natom, sp.mass, betan, TAU, DT, GAMMA,
                               sp.well1, b, sp.well2 - sp.well1, x0, threads, steps, threads, 1,
                               fast=True)
        out[n] = wall / (threads * steps)  # seconds per polymer-step, all cores busy
        detail[n] = {"traj": threads, "steps": steps, "wall_s": round(wall, 3)}
    O.use_blas_dsymv(None)
    return out, detail, ("mkl_dsymv" if blas else "loop")


def _mkl_dsymv():
    """Address of MKL's sequential dsymv_ (the reference's transform call), or None."""
    import ctypes

    if not os.path.exists(MKL_RT):
        return None
    os.environ["MKL_THREADING_LAYER"] = "SEQUENTIAL"
    os.environ["MKL_NUM_THREADS"] = "1"
    return ctypes.cast(ctypes.CDLL(MKL_RT).dsymv_, ctypes.c_void_p).value


def quartic_full_sample(threads, nrep=8, nsteps=100000, imin=1000, n=64):
    """The 1D quartic double well in full (V0 = 0.005, a = 0.7, m = 1836), N = 64,
    beta = 2000, 11 Gauss-Legendre nodes x nrep replicas x nsteps Langevin steps, on the GPU
    (the batched engine: init_path, TI after imin, the reference driver's quadrature) and
    on the host cores (the oracle's restatement through MKL dsymv, from the straight line
    with resampled momenta, TI from the first step).  Wall time to solution on each side
    and both Delta A +- sigma_A, q/q0: equal within their error bars."""
    from concurrent.futures import ThreadPoolExecutor

    import oracle as O
    from rptunnel import Batch, Pes, specs, ti

    sp = specs.quartic()
    pes = Pes(sp)
    betan = BETA / (n + 1)

    def factory(pols):
        b = Batch(pes, sp.mass, pols, TAU, DT, GAMMA, seed=11)
        b.set_launch(64)
        return b

    rp_sync()
    t0 = time.perf_counter()
    g = ti.run_ti(factory, sp, n, BETA, NLAMBDA, nrep, nsteps, imin)
    rp_sync()
    tg = time.perf_counter() - t0
    ntask = int(np.count_nonzero(g["integrands"]))
    # host cores: the same tasks on the oracle
    O.use_blas_dsymv(_mkl_dsymv())
    op = O.Pes(sp.kind, sp.ndim, sp.natom, sp.dp, sp.ip)
    lampath, path = specs.linear_path(sp.well1, sp.well2, npath=2)
    xi, w = ti.gauleg(0.0, 1.0, NLAMBDA)
    ends, grads = ti._spline_nodes(lampath, path, xi)
    t = (np.arange(n) + 1.0) / (n + 1.0)
    per_call = max(1, min(nrep, threads))

    def node(i):
        if ti.skip_task(ends[i]):
            return np.zeros(nrep)
        x0 = np.outer(sp.well1, 1 - t) + np.outer(ends[i], t)
        _, dh = O.bench_pile(op, n, sp.ndim, sp.natom, sp.mass, betan, TAU, DT, GAMMA, sp.well1,
                             ends[i], grads[i], x0, nrep, nsteps, per_call, 1000 + i, fast=True)
        return dh / betan ** 2

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=max(1, threads // per_call)) as ex:
        cpu_int = np.concatenate(list(ex.map(node, range(NLAMBDA))))
    tc = time.perf_counter() - t0
    O.use_blas_dsymv(None)
    c = ti.quadrature(cpu_int, w, nrep, betan)
    keys = ("deltaA", "sigmaA", "q_ratio", "q_err")
    return {"workload": f"1D quartic double well TI in full: N = {n}, beta = {BETA:g}, "
                        f"{NLAMBDA} lambda x {nrep} replicas x {nsteps} steps ({ntask} trajectories; "
                        "the middle node has b = 0 and is skipped, pimd_par.f90:322-325)",
            "gpu": {"wall_s": tg, "bead_steps_per_s": ntask * n * nsteps / tg,
                    **{k: g[k] for k in keys}},
            "cpu": {"wall_s": tc, "threads": threads,
                    "bead_steps_per_s": ntask * n * nsteps / tc, **{k: c[k] for k in keys}},
            "deltaA_difference_in_sigma": abs(g["deltaA"] - c["deltaA"])
                                          / math.sqrt(g["sigmaA"] / nrep + c["sigmaA"] / nrep)}


def cpu_workloads_sample(threads, budget_s=3.0):
    """The host-core baselines of the other workloads (oracle restatement; MKL dsymv for the
    transforms): coupled-2D and water-cluster bead-steps/s on a bounded window of all cores
    (as the EVB TI baseline), and the EVB instanton figures at N = 64, 256, 1024:
    U_M + gradient bead-evaluations/s (one chain per thread), the UMhessian band builds and
    the DSBEVD eigenvalue solve (LAPACK through scipy's eig_banded, the call detJ makes)."""
    from concurrent.futures import ThreadPoolExecutor

    import scipy.linalg

    import oracle as O
    from rptunnel import specs

    out = {}
    for name, sp in (("coupled2d", specs.coupled2d()),
                     ("water6", specs.water_cluster(nmol=6))):
        tc, detail, symv = cpu_sample(sp, [128], budget_s, threads)
        out[name] = {"bead_steps_per_s": 128 / tc[128], "sample": detail[128], "symv": symv}
    sp = specs.evb()
    op = O.Pes(sp.kind, sp.ndim, sp.natom, sp.dp, sp.ip)
    inst = {}
    for n in (64, 256, 1024):
        betan = 4000.0 / n
        ch = O.Chain(n, sp.ndim, sp.natom, sp.mass, betan, True)
        k = np.arange(1, n + 1)
        x = (np.outer(sp.well1, n - k) + np.outer(sp.well2, k - 1)) / (n - 1)
        ch.UMforceenergy(op, x, sp.well1, sp.well2)
        t1 = time.perf_counter()
        ch.UMforceenergy(op, x, sp.well1, sp.well2)
        per = max(time.perf_counter() - t1, 1e-6)
        reps = max(1, int(budget_s / 3 / per))

        def work(_):
            for _ in range(reps):
                ch.UMforceenergy(op, x, sp.well1, sp.well2)

        t1 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=threads) as ex:
            list(ex.map(work, range(threads)))
        tu = time.perf_counter() - t1
        t1 = time.perf_counter()
        band = ch.UMhessian(op, x, False)
        th = time.perf_counter() - t1
        te = None
        if n <= 256:  # the O((nD)^2) eigensolve takes over a minute at N = 1024
            t1 = time.perf_counter()
            scipy.linalg.eig_banded(band, lower=True, eigvals_only=True)
            te = time.perf_counter() - t1
        inst[n] = {"um_grad_bead_evals_per_s": threads * reps * n / tu,
                   "hessian_band_bead_blocks_per_s": n / th,
                   "dsbevd_bead_blocks_per_s": n / te if te else None}
    out["evb_instantons"] = {"per_N": inst, "threads_for_um": threads,
                                     "hessian_and_eigensolve": "one core each; the eigensolve "
                                     "only up to N = 256"}
    return out


def copy_probe(mib):
    """One device-to-device copy of a known size: rocprofv3's FETCH_SIZE / WRITE_SIZE of
    this dispatch against its 2 x mib MiB of algorithmic traffic calibrate the counters."""
    import torch

    n = mib * (1 << 20) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_mib": mib, "bytes_read": n * 8, "bytes_written": n * 8}))


def instanton_sample(pes, sp, nchain=256, reps=10, seed=4):
    """The batched instanton workload on the same EVB surface: nchain open chains with N
    uniform in [64, 1024] and beta in [1000, 8000] (RPI betan = beta/N, rpi_ser.f90:45).

    Throughput: bead-evaluations per second of U_M with its adjoint gradient and of the HVP,
    log-determinant bead blocks per second (lndetJ0 at well 1); each figure includes the
    host copies of its inputs/outputs (the gradient and HVP figures move n*ndof doubles each
    way per call), except the energy-only U_M figure, which returns one double per chain,
    and the *_device figures, whose vectors stay in device memory (rp_ring_*_device).

    Solution (rpi_ser.f90:95, 146-191, 221-237, 354-381): V0 = V(well1); every chain's
    instanton from a straight line by the device L-BFGS with the reference's settings
    (m = 8, factr = 1e6, pgtol = eps2 = 1e-5, no iteration cap, instantonmod.f90:741-767);
    V0 moved down to the instanton's lowest bead where it dips below V(well1);
    lndetJ at each instanton, lndetJ0 with every bead at well 1, and the splitting
    Delta = 2 theta / betan.  Reported: converged chains, iterations, wall time of each stage
    and in total, and per chain S = betan U_M, lndetJ, lndetJ0 and Delta (cm^-1)."""
    from rptunnel import Ring

    rng = np.random.default_rng(seed)
    ns = [int(v) for v in rng.integers(64, 1025, nchain)]
    betas = rng.uniform(1000.0, 8000.0, nchain)
    betan = [b / n for b, n in zip(betas, ns)]
    ring = Ring(pes, sp.mass, ns, betan, sp.well1, sp.well2)
    chains = []
    for n in ns:
        k = np.arange(1, n + 1)
        chains.append((np.outer(sp.well1, n - k) + np.outer(sp.well2, k - 1)) / (n - 1))
    ring.set_x(chains)
    beads = sum(ns)
    ring.UMforceenergy()
    t0 = time.perf_counter()
    for _ in range(reps):
        ring.UMforceenergy()
    t_um = (time.perf_counter() - t0) / reps
    ring.UM()
    t0 = time.perf_counter()
    for _ in range(reps):
        ring.UM()
    t_e = (time.perf_counter() - t0) / reps
    v = ring.unpack(rng.standard_normal(ring.off[-1]))
    ring.hvp(v)
    t0 = time.perf_counter()
    for _ in range(reps):
        ring.hvp(v)
    t_hvp = (time.perf_counter() - t0) / reps
    # the same two calls with the chains resident in device memory (rp_ring_*_device, as the
    # torch.autograd wrapper runs them): no host copies of full vectors
    import torch

    dev = torch.device("cuda", torch.cuda.current_device())
    xd = torch.as_tensor(np.concatenate([c.reshape(-1) for c in chains]), device=dev)
    gd = torch.empty_like(xd)
    vd = torch.as_tensor(np.concatenate([c.reshape(-1) for c in v]), device=dev)
    hd = torch.empty_like(vd)
    torch.cuda.synchronize()
    ring.set_x_device(xd.data_ptr())
    ring.um_device(gd.data_ptr())
    t0 = time.perf_counter()
    for _ in range(reps):
        ring.um_device(gd.data_ptr())
    t_um_dev = (time.perf_counter() - t0) / reps
    ring.hvp_device(vd.data_ptr(), hd.data_ptr())
    t0 = time.perf_counter()
    for _ in range(reps):
        ring.hvp_device(vd.data_ptr(), hd.data_ptr())
    t_hvp_dev = (time.perf_counter() - t0) / reps
    del xd, gd, vd, hd
    wells = [np.repeat(sp.well1[:, None], n, axis=1) for n in ns]
    ring.set_x(wells)
    ring.logdet(singlewell=True, skip_lowest=False)
    t0 = time.perf_counter()
    for _ in range(max(1, reps // 5)):
        ring.logdet(singlewell=True, skip_lowest=False)
    t_ld = (time.perf_counter() - t0) / max(1, reps // 5)

    # the solution, as rpi_ser.f90 computes it for every chain
    pes.set_V0(pes.V(sp.well1))
    t0 = time.perf_counter()
    xo, act, it, st = ring.instanton(chains, pgtol=1e-5, factr=1e6, m=8, maxiter=0)
    t_opt = time.perf_counter() - t0
    # the same search preconditioned by the springs (rp_ring_set_preconditioner), same tests
    ring.set_preconditioner(0.05)
    t0 = time.perf_counter()
    xp, actp, itp, stp = ring.instanton(chains, pgtol=1e-5, factr=1e6, m=8, maxiter=0)
    t_pc = time.perf_counter() - t0
    ring.set_preconditioner(0.0)
    pc_ds = np.asarray(betan) * (actp - act)  # S = betan U_M of the two searches
    # rpi_ser.f90:181-191: when the instanton dips below V(well1), V0 moves to its lowest
    # bead, which lowers every bead's V - V0 (and U_M by n times the dip)
    for c, x in enumerate(xo):
        vmin = float(np.min(pes.V(x)))
        if vmin < 0.0:
            act[c] -= ns[c] * vmin
    ring.set_x(xo)
    t0 = time.perf_counter()
    ldj, nneg, _ = ring.logdet(singlewell=False, skip_lowest=True)
    t_ldj = time.perf_counter() - t0
    ring.set_x(wells)
    t0 = time.perf_counter()
    ldj0, _, _ = ring.logdet(singlewell=True, skip_lowest=False)
    t_ldj0 = time.perf_counter() - t0
    pes.set_V0(0.0)
    ring.close()
    bn = np.asarray(betan)
    skink = bn * act
    phi = np.exp(0.5 * (ldj - ldj0))
    theta = bn * np.exp(-skink) * np.sqrt(skink / (2.0 * 3.14159265358979)) / phi
    delta_cm = 2.0 * theta / bn * 219475.0

    def r10(a):
        return [float(f"{float(x):.10g}") for x in a]

    return {"workload": f"evb27 instantons: {nchain} chains, N in [64, 1024] (total {beads} beads), "
                        "beta in [1000, 8000]",
            "um_grad_bead_evals_per_s": beads / t_um,
            "um_energy_only_bead_evals_per_s": beads / t_e,
            "hvp_bead_evals_per_s": beads / t_hvp,
            "um_grad_device_bead_evals_per_s": beads / t_um_dev,
            "hvp_device_bead_evals_per_s": beads / t_hvp_dev,
            "logdet_bead_blocks_per_s": beads / t_ld,
            "solution": {
                "settings": "V0 = V(well1); L-BFGS m = 8, factr = 1e6, pgtol = 1e-5, no iteration cap",
                "converged": int(np.sum(st == 0)), "chains": nchain,
                "status_counts": {str(k): int(np.sum(st == k)) for k in np.unique(st)},
                "iterations": {"max": int(np.max(it)), "mean": float(np.mean(it))},
                "seconds": {"lbfgs": t_opt, "lndetJ": t_ldj, "lndetJ0": t_ldj0,
                            "total": t_opt + t_ldj + t_ldj0},
                "lbfgs_ms_per_iteration": 1e3 * t_opt / max(int(np.max(it)), 1),
                "preconditioned": {"sigma": 0.05, "converged": int(np.sum(stp == 0)),
                                   "iterations": {"max": int(np.max(itp)), "mean": float(np.mean(itp))},
                                   "seconds_lbfgs": t_pc,
                                   "max_abs_S_difference": float(np.max(np.abs(pc_ds))),
                                   "chains_lower_S_than_plain": int(np.sum(pc_ds < 0.0))},
                "negative_eigenvalues_at_instanton": {str(k): int(np.sum(nneg == k)) for k in np.unique(nneg)},
                "per_chain": {"n": ns, "beta": r10(betas), "S": r10(skink), "lndetJ": r10(ldj),
                              "lndetJ0": r10(ldj0), "delta_cm-1": r10(delta_cm)}}}


def water_sample(steps=16, warmup=8, ntraj=1024, n=128, nchain=256):
    """The 6-water flexible cluster (18 atoms, 54 dof): batched PIMD-TI at N = 128 with
    ntraj trajectories x 11 lambda nodes (Langevin, the same step parameters as the
    headline), and the adjoint sensitivities dU_M/d(qH, De) of nchain instanton chains of
    N beads.  Timed on the device with the inputs resident; bead-steps and bead-evaluations
    per second."""
    from rptunnel import Batch, Pes, Ring, specs

    from rptunnel import ti

    sp = specs.water_cluster(nmol=6)
    pes = Pes(sp)
    # the path turns molecule 0's H pair about its bisector (a straight line to the
    # exchanged well would run the two H atoms through each other)
    path = np.ascontiguousarray(sp.extra["path"].T)
    lampath = specs.arc_lampath(path)
    xi, _ = gauleg(NLAMBDA)
    ends, grads = ti._spline_nodes(lampath, path, xi)
    pols = [dict(n=n, betan=BETA / (n + 1), xi=float(x), a=path[:, 0], b=ends[il],
                 dbdl=grads[il], stream=il * ntraj + r)
            for il, x in enumerate(xi) for r in range(ntraj)]
    b = Batch(pes, sp.mass, pols, TAU, DT, GAMMA, seed=7)
    b.set_path(lampath, path)
    b.init_path()
    b.set_launch(8)
    b.run(warmup)
    rp_sync()
    t0 = time.perf_counter()
    b.run(steps)
    rp_sync()
    wall = time.perf_counter() - t0
    ms, launches = b.last_kernel_ms()
    sums, _, _ = b.ti_readout()
    assert np.all(np.isfinite(sums)), "non-finite TI sums (water)"
    b.close()
    k = np.arange(1, n + 1)
    chain = (np.outer(sp.well1, n - k) + np.outer(sp.well2, k - 1)) / (n - 1)
    ring = Ring(pes, sp.mass, [n] * nchain, [BETA / n] * nchain, sp.well1, sp.well2)
    ring.set_x([chain] * nchain)
    ring.param_grad()
    t1 = time.perf_counter()
    for _ in range(5):
        pg = ring.param_grad()
    tpg = (time.perf_counter() - t1) / 5
    assert np.all(np.isfinite(pg))
    ring.close()
    return {"workload": f"water6 (54 dof) batched TI: N = {n}, {ntraj} traj x {NLAMBDA} lambda "
                        f"({len(pols)} polymers), Langevin; dU_M/d(qH, De) of {nchain} chains "
                        f"of N = {n}",
            "bead_steps_per_s": len(pols) * n * steps / wall,
            "ms_per_step": 1e3 * wall / steps,
            "avg_launch_ms": round(float(ms[0]), 4), "steps_per_launch": 8,
            "param_grad_bead_evals_per_s": nchain * n / tpg}


def coupled2d_sample(steps=16, warmup=8, ntraj=4096, nlam=16, n=128, plain_steps=20):
    """The 2D coupled double well (V0 = 0.005, a = 0.7, w = 0.01, c = 0.3, mass
    1836), PIMD-TI at N = 128 with ntraj trajectories x nlam lambda nodes (65,536 polymers),
    Langevin, the headline's beta / tau / dt / gamma, on the batched engine (inputs resident,
    device-timed) and on the plain per-trajectory path (one trajectory, plain_steps steps)."""
    from rptunnel import Batch, Pes, specs

    sp = specs.coupled2d()
    pes = Pes(sp)
    xi, _ = gauleg(nlam)
    lampath, path = specs.linear_path(sp.well1, sp.well2, npath=2)
    pols = [dict(n=n, betan=BETA / (n + 1), xi=float(x), a=sp.well1,
                 b=sp.well1 + x * (sp.well2 - sp.well1), dbdl=sp.well2 - sp.well1,
                 stream=il * ntraj + r)
            for il, x in enumerate(xi) for r in range(ntraj)]
    b = Batch(pes, sp.mass, pols, TAU, DT, GAMMA, seed=5)
    b.set_path(lampath, path)
    b.init_path()
    b.set_launch(8)
    b.run(warmup)
    rp_sync()
    t0 = time.perf_counter()
    b.run(steps)
    rp_sync()
    wall = time.perf_counter() - t0
    ms, _ = b.last_kernel_ms()
    shape = b.family_shapes()[0]
    sums, _, _ = b.ti_readout()
    assert np.all(np.isfinite(sums)), "non-finite TI sums (coupled 2D)"
    b.close()
    tp = plain_sample(pes, sp, [n], plain_steps)[n]
    batched = len(pols) * n * steps / wall
    return {"workload": f"coupled 2D double well batched TI: N = {n}, {ntraj} traj x {nlam} "
                        f"lambda ({len(pols)} polymers), Langevin",
            "bead_steps_per_s": batched, "ms_per_step": 1e3 * wall / steps,
            "avg_launch_ms": round(float(ms[0]), 4), "steps_per_launch": 8,
            "kernel_shape": {"waves": shape[1], "units_per_wave": shape[2], "folded": shape[3]},
            "plain_gpu_bead_steps_per_s": n / tp, "batched_over_plain": batched * tp / n}


def equal_splitting_sample(pes, sp, n=32, nrep=4, nsteps=400, imin=100, seed=77):
    """The metric's "at equal TI splitting error" clause, measured: the same TI tasks
    (11 Gauss-Legendre nodes x nrep replicas of the 27-dof EVB at one N, init_path, nsteps
    Langevin steps, TI after imin) on the batched engine and on the plain per-trajectory
    path, drawing the same Philox streams.  Reports both paths' integrands' largest
    relative difference, Delta A, sigma_A, q/q0 and the two-level splitting 2 artanh(q/q0)/beta
    from each, and the plain path's bead-steps/s over all its trajectories."""
    from rptunnel import Batch, ti

    def factory(pols):
        return Batch(pes, sp.mass, pols, TAU, DT, GAMMA, seed=seed)

    rb = ti.run_ti(factory, sp, n, BETA, NLAMBDA, nrep, nsteps, imin)
    rp = ti.run_ti_plain(pes, sp, n, BETA, NLAMBDA, nrep, nsteps, imin, TAU, DT, GAMMA, seed)
    ib, ip = rb["integrands"], rp["integrands"]
    diff = float(np.max(np.abs(ib - ip) / np.maximum(np.abs(ip), 1e-300)))

    def summary(r):
        q = r["q_ratio"]
        return {"deltaA": r["deltaA"], "sigmaA": r["sigmaA"], "q_ratio": q, "q_err": r["q_err"],
                "splitting_cm-1": (2.0 * np.arctanh(q) / BETA * 219475.0) if 0.0 < q < 1.0 else None}

    ntraj = NLAMBDA * nrep
    return {"workload": f"evb27 TI at N = {n}: {NLAMBDA} nodes x {nrep} replicas, {nsteps} steps "
                        f"(TI after {imin}), same Philox streams on both paths",
            "max_rel_integrand_difference": diff,
            "batched": summary(rb), "plain": summary(rp),
            "plain_bead_steps_per_s": ntraj * n * nsteps / rp["seconds"],
            "plain_trajectories_timed": ntraj}


def andersen_sample(pes, sp, ns=(64, 256), steps=16, ntraj=NTRAJ):
    """The Andersen thermostat (propagate_pimd_nm, verletmodule.f90:190-250) on the same
    27-dof EVB batch shape at two bead counts: 512 trajectories x 11 lambda per N, Poisson
    resampling with mean interval 5 steps, 8 steps per launch.  Bead-steps/s and each
    family's average launch time."""
    from rptunnel import Batch
    from rptunnel.batch import ANDERSEN

    pols, lampath, path = build_polymers(sp, list(ns), ntraj, NLAMBDA)
    b = Batch(pes, sp.mass, pols, TAU, DT, thermostat=ANDERSEN, noutput=5, seed=2025)
    b.set_path(lampath, path)
    b.init_path()
    b.set_launch(8)
    b.run(8)
    rp_sync()
    t0 = time.perf_counter()
    b.run(steps)
    rp_sync()
    wall = time.perf_counter() - t0
    ms, _ = b.last_kernel_ms()
    sums, _, _ = b.ti_readout()
    assert np.all(np.isfinite(sums)), "non-finite TI sums (Andersen)"
    fams = [{"n": n, "waves": w, "units_per_wave": u, "avg_launch_ms": round(float(m), 4)}
            for (n, w, u, _), m in zip(b.family_shapes(), ms)]
    b.close()
    return {"workload": f"evb27 batched TI, Andersen (noutput 5): N = {list(ns)}, {ntraj} traj x "
                        f"{NLAMBDA} lambda each ({len(pols)} polymers)",
            "bead_steps_per_s": sum(p["n"] for p in pols) * steps / wall,
            "ms_per_step": 1e3 * wall / steps, "steps_per_launch": 8, "families": fams}


def rp_sync():
    import rptunnel
    rptunnel.lib().rp_finalize()  # device synchronize


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--ntraj", type=int, default=NTRAJ)
    ap.add_argument("--plain-steps", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-instanton", action="store_true",
                    help="skip the batched-instanton sample reported beside the headline")
    ap.add_argument("--no-equal", action="store_true",
                    help="skip the same-TI-on-both-paths sample reported beside the headline")
    ap.add_argument("--no-andersen", action="store_true",
                    help="skip the Andersen-thermostat sample reported beside the headline")
    ap.add_argument("--no-others", action="store_true",
                    help="skip the quartic TI in full and the host baselines of the coupled-2D, "
                         "EVB-instanton and water workloads")
    ap.add_argument("--no-coupled2d", action="store_true",
                    help="skip the coupled-2D sample reported beside the headline")
    ap.add_argument("--no-water", action="store_true",
                    help="skip the water-cluster sample reported beside the headline")
    ap.add_argument("--steps-per-launch", type=int, default=8,
                    help="steps fused into one kernel launch (warm-up and timed steps are "
                         "multiples of it, so every launch of a family has the same shape)")
    ap.add_argument("--autotune", action="store_true",
                    help="pick steps per launch by a timed trial instead")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (the HBM byte-counter "
                         "calibration for the traffic figure) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): every trajectory "
              "is independent, so the tasks would be dealt out to the GPUs (each its own batch "
              "of whole polymers) with only the TI sums gathered at the end.", file=sys.stderr)
        sys.exit(2)

    # torch first, as under pytest: the library then shares the HIP runtime torch loads, and
    # the instanton sample can hand it device tensors
    import torch  # noqa: F401

    import rptunnel
    from rptunnel import Batch, Pes, specs

    sp = specs.evb()
    D = sp.ndof
    pes = Pes(sp)
    pols, lampath, path = build_polymers(sp, NS, args.ntraj, NLAMBDA)
    log(f"creating batch of {len(pols)} polymers (N = {NS}, {args.ntraj} traj x {NLAMBDA} lambda)")
    batch = Batch(pes, sp.mass, pols, TAU, DT, GAMMA, seed=2024)
    batch.set_path(lampath, path)
    batch.init_path()
    fams = batch.families()
    if args.autotune:
        best_steps, _ = batch.autotune(trial_steps=16)
        log(f"autotune: {best_steps} steps per launch")
    else:
        batch.set_launch(args.steps_per_launch)
    log(f"warm-up {args.warmup} steps")
    batch.run(args.warmup)
    rptunnel.lib().rp_finalize()  # device synchronize
    log(f"timing {args.steps} steps")
    t0 = time.perf_counter()
    batch.run(args.steps)
    rptunnel.lib().rp_finalize()
    wall = time.perf_counter() - t0
    beads_per_step = sum(p["n"] for p in pols)
    value = beads_per_step * args.steps / wall
    fam_ms, launches = batch.last_kernel_ms()
    sums, cnt, _ = batch.ti_readout()
    assert np.all(np.isfinite(sums)), "non-finite TI sums"
    log(f"batched: {value:.4e} bead-steps/s ({1e3 * wall / args.steps:.3f} ms/step)")

    # roofline of the dominant launch (the family with the largest summed launch time)
    steps_per_launch = int(np.ceil(args.steps / launches)) if launches else args.steps
    per_launch_steps = args.steps / launches
    fam_rows = []
    for (n, count), ms in zip(fams, fam_ms):
        fl = per_launch_steps * count * n * flops_per_bead_step(n, D)
        fam_rows.append({"n": n, "polymers": count, "avg_launch_ms": round(ms, 4),
                         "tflops": round(fl / (ms * 1e-3) / 1e12, 3)})
    dom = max(range(len(fam_rows)), key=lambda i: fam_ms[i])
    dn, dcount = fams[dom]
    dflops = per_launch_steps * dcount * dn * flops_per_bead_step(dn, D)
    achieved = dflops / (fam_ms[dom] * 1e-3)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "batch_evb27_n256_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    log(f"plain path sample ({args.plain_steps} steps per N)")
    tp = plain_sample(pes, sp, NS, args.plain_steps)
    polys_per_n = args.ntraj * NLAMBDA
    plain_value = sum(n * polys_per_n for n in NS) / sum(tp[n] * polys_per_n for n in NS)

    eqs = None
    if not args.no_equal:
        log("equal-splitting sample: the same TI on the batched and the plain path")
        eqs = equal_splitting_sample(pes, sp)

    andersen = None
    if not args.no_andersen:
        log("Andersen sample (N = 64, 256)")
        andersen = andersen_sample(pes, sp)

    inst = None
    if not args.no_instanton:
        log("instanton sample (256 chains)")
        inst = instanton_sample(pes, sp)

    coupled = None
    if not args.no_coupled2d:
        log("coupled 2D sample (N = 128, 65536 polymers)")
        coupled = coupled2d_sample()

    water = None
    if not args.no_water:
        log("water-cluster sample (N = 128, 11264 polymers)")
        water = water_sample()

    quartic = cpu_others = None
    if not args.no_cpu and not args.no_others:
        threads = host_cores()
        log(f"the quartic TI in full on the GPU and on {threads} host threads")
        quartic = quartic_full_sample(threads)
        log("host baselines of the coupled-2D, EVB-instanton and water workloads")
        cpu_others = cpu_workloads_sample(threads)

    cpu = None
    if not args.no_cpu:
        threads = host_cores()
        log(f"cpu baseline: oracle on {threads} threads, ~{args.cpu_seconds:.0f} s")
        tc, detail, symv = cpu_sample(sp, NS, args.cpu_seconds, threads)
        cpu_value = sum(n for n in NS) / sum(tc[n] for n in NS)
        calib = None
        cfile = os.path.join(ROOT, "oracle", "cpu_calibration.json")
        if os.path.exists(cfile):
            with open(cfile) as f:
                c = json.load(f)
            calib = {"restatement_loop_over_mkl_dsymv_mix": round(c["mix_ratio_restatement_over_mkl"], 3),
                     "per_N": {k: round(v["ratio"], 3) for k, v in c["sec_per_polymer_step"].items()},
                     "measured_on": f"{c['cpu']}, {c['cores']} cores (oracle/calibrate_cpu.py)"}
        cpu = {"value": cpu_value, "unit": "bead-steps/s", "cores": threads, "kind": "port",
               "sample": "oracle restatement of the reference PILE step (10 dense transforms "
                         f"per step through {'MKL sequential dsymv' if symv == 'mkl_dsymv' else 'the restatement loop'}, "
                         "reference step order; -O3 build), one trajectory per thread, "
                         f"at every N of the mix: {detail}; equal polymer counts per N",
               "symv": symv, "calibration": calib}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "bead-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "evb27-batched-ti", "natom": 9, "ndof": D, "N": NS,
                   "traj_per_N": args.ntraj, "lambda_nodes": NLAMBDA, "polymers": len(pols),
                   "beta": BETA, "tau": TAU, "dt": DT, "gamma": GAMMA,
                   "thermostat": "langevin", "steps_per_launch": steps_per_launch},
        "plain_gpu": {"value": plain_value, "unit": "bead-steps/s",
                      "sec_per_polymer_step": {str(n): tp[n] for n in NS},
                      "sample": f"one trajectory per N, {args.plain_steps} steps, same step"},
        "speedup_vs_plain": value / plain_value,
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64, "traffic": traffic,
                     "kernel": f"batch_kernel N={dn} family ({dcount} polymers)",
                     "flops_per_bead_step": flops_per_bead_step(dn, D),
                     "families": fam_rows},
        "cpu_baseline": cpu,
        "equal_splitting": eqs,
        "andersen": andersen,
        "instanton": inst,
        "coupled2d": coupled,
        "quartic_full": quartic,
        "cpu_other_workloads": cpu_others,
        "water": water,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
