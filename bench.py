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


def cpu_sample(sp, ns, budget_s, threads):
    """Oracle restatement on the host cores: seconds per polymer-step at each N."""
    import oracle as O

    op = O.Pes(sp.kind, sp.ndim, sp.natom, sp.dp, sp.ip)
    out, detail = {}, {}
    for n in ns:
        betan = BETA / (n + 1)
        b = sp.well1 + 0.5 * (sp.well2 - sp.well1)
        t = (np.arange(n) + 1.0) / (n + 1.0)
        x0 = np.outer(sp.well1, 1 - t) + np.outer(b, t)
        # calibrate: one trajectory, few steps
        t1, _ = O.bench_pile(op, n, sp.ndim, sp.natom, sp.mass, betan, TAU, DT, GAMMA, sp.well1,
                             b, sp.well2 - sp.well1, x0, 1, 4, 1, 1)
        per = t1 / 4
        steps = max(4, int(budget_s / len(ns) / per))
        wall, _ = O.bench_pile(op, n, sp.ndim, sp.natom, sp.mass, betan, TAU, DT, GAMMA,
                               sp.well1, b, sp.well2 - sp.well1, x0, threads, steps, threads, 1)
        out[n] = wall / (threads * steps)  # seconds per polymer-step, all cores busy
        detail[n] = {"traj": threads, "steps": steps, "wall_s": round(wall, 3)}
    return out, detail


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


def instanton_sample(pes, sp, nchain=256, reps=10, maxiter=300, seed=4):
    """The batched instanton workload on the same EVB surface: nchain open chains with N
    uniform in [64, 1024] and beta in [1000, 8000] (RPI betan = beta/N, rpi_ser.f90:45).
    Reports bead-evaluations per second of U_M with its adjoint gradient and of the HVP,
    log-determinant bead blocks per second (lndetJ0 at well 1), and one batched L-BFGS
    run from straight lines; each figure includes the host copies of its inputs/outputs
    (the gradient and HVP figures move n*ndof doubles each way per call), except the
    energy-only U_M figure, which returns one double per chain.  The L-BFGS search keeps
    its vectors on the device: all chains iterate in lockstep, so its wall time follows
    the largest iteration count."""
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
    ring.set_x([np.repeat(sp.well1[:, None], n, axis=1) for n in ns])
    ring.logdet(singlewell=True, skip_lowest=False)
    t0 = time.perf_counter()
    for _ in range(max(1, reps // 5)):
        ring.logdet(singlewell=True, skip_lowest=False)
    t_ld = (time.perf_counter() - t0) / max(1, reps // 5)
    t0 = time.perf_counter()
    _, act, it, st = ring.instanton(chains, pgtol=1e-5, factr=1e6, m=8, maxiter=maxiter)
    t_opt = time.perf_counter() - t0
    ring.close()
    return {"workload": f"evb27 instantons: {nchain} chains, N in [64, 1024] (total {beads} beads), "
                        "beta in [1000, 8000]",
            "um_grad_bead_evals_per_s": beads / t_um,
            "um_energy_only_bead_evals_per_s": beads / t_e,
            "hvp_bead_evals_per_s": beads / t_hvp,
            "logdet_bead_blocks_per_s": beads / t_ld,
            "lbfgs": {"seconds": t_opt, "max_iterations": int(np.max(it)),
                      "ms_per_iteration": 1e3 * t_opt / max(int(np.max(it)), 1),
                      "mean_iterations": float(np.mean(it)),
                      "converged": int(np.sum(st == 0)), "chains": nchain,
                      "pgtol": 1e-5, "factr": 1e6, "iteration_cap": maxiter}}


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


def rp_sync():
    import rptunnel
    rptunnel.lib().rp_finalize()  # device synchronize


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--ntraj", type=int, default=NTRAJ)
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-instanton", action="store_true",
                    help="skip the batched-instanton sample reported beside the headline")
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

    inst = None
    if not args.no_instanton:
        log("instanton sample (256 chains)")
        inst = instanton_sample(pes, sp)

    water = None
    if not args.no_water:
        log("water-cluster sample (N = 128, 11264 polymers)")
        water = water_sample()

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline: oracle on {threads} threads, ~{args.cpu_seconds:.0f} s")
        tc, detail = cpu_sample(sp, NS, args.cpu_seconds, threads)
        cpu_value = sum(n for n in NS) / sum(tc[n] for n in NS)
        cpu = {"value": cpu_value, "unit": "bead-steps/s", "cores": threads, "kind": "port",
               "sample": "oracle restatement of the reference PILE step (10 dense transforms "
                         "per step, reference step order), one trajectory per thread, "
                         f"at every N of the mix: {detail}; equal polymer counts per N"}

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
        "instanton": inst,
        "water": water,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
