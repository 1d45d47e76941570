"""CPU-baseline calibration (container only): the restatement's PILE step with its own
mat-vec loop against the same step with MKL's sequential dsymv, the BLAS call the reference
makes (verletmodule.f90:259, 279; linked -mkl=sequential, makefile:6).

Both variants run oracle/_build/liboracle_fast.so (the build bench.py's cpu_baseline leg
times) on the 27-dof EVB workload at every N of the benchmark mix, one trajectory per core
on all the cores of this machine.  Writes oracle/cpu_calibration.json: per N the seconds
per polymer-step of each variant and their ratio (restatement / MKL: above 1 means MKL is
faster), and the same for the whole mix.  TEST AND BASELINE INFRASTRUCTURE ONLY.

Usage: python oracle/calibrate_cpu.py [--mkl /opt/conda/lib/libmkl_rt.so] [--seconds 4]
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [HERE, os.path.join(HERE, "..", "pimd-tunneling_amd", "python")]

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402

NS = [32, 64, 96, 128, 192, 256]
BETA, TAU, DT, GAMMA = 2000.0, 20.0, 1.0, 1e-3


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mkl", default="/opt/conda/lib/libmkl_rt.so")
    ap.add_argument("--seconds", type=float, default=4.0, help="per N and variant")
    args = ap.parse_args()
    os.environ["MKL_THREADING_LAYER"] = "SEQUENTIAL"
    os.environ["MKL_NUM_THREADS"] = "1"
    mkl = C.CDLL(args.mkl)
    dsymv = C.cast(mkl.dsymv_, C.c_void_p).value
    from rptunnel import specs

    sp = specs.evb()
    op = O.Pes(sp.kind, sp.ndim, sp.natom, sp.dp, sp.ip)
    cores = len(os.sched_getaffinity(0))
    rows = {}
    for n in NS:
        betan = BETA / (n + 1)
        b = sp.well1 + 0.5 * (sp.well2 - sp.well1)
        t = (np.arange(n) + 1.0) / (n + 1.0)
        x0 = np.outer(sp.well1, 1 - t) + np.outer(b, t)
        res = {}
        for name, fn in (("restatement", None), ("mkl_dsymv", dsymv)):
            O.use_blas_dsymv(fn)
            t1, _ = O.bench_pile(op, n, sp.ndim, sp.natom, sp.mass, betan, TAU, DT, GAMMA, sp.well1,
                                 b, sp.well2 - sp.well1, x0, 1, 4, 1, 1, fast=True)
            steps = max(4, int(args.seconds * cores / cores / (t1 / 4)))
            wall, _ = O.bench_pile(op, n, sp.ndim, sp.natom, sp.mass, betan, TAU, DT, GAMMA,
                                   sp.well1, b, sp.well2 - sp.well1, x0, cores, steps, cores, 1,
                                   fast=True)
            res[name] = wall / (cores * steps)  # seconds per polymer-step, all cores busy
        O.use_blas_dsymv(None)
        res["ratio"] = res["restatement"] / res["mkl_dsymv"]
        rows[str(n)] = res
        print(n, res, flush=True)
    mix = {k: sum(n for n in NS) / sum(rows[str(n)][k] for n in NS) for k in ("restatement", "mkl_dsymv")}
    out = {"what": "PILE step of the 27-dof EVB, restatement loop vs MKL sequential dsymv, "
                   "liboracle_fast.so (-O3 -march=x86-64-v3)",
           "cpu": platform.processor() or platform.machine(), "cores": cores,
           "mkl": os.path.realpath(args.mkl),
           "sec_per_polymer_step": rows,
           "mix_bead_steps_per_s": mix,
           "mix_ratio_restatement_over_mkl": mix["mkl_dsymv"] / mix["restatement"]}
    with open(os.path.join(HERE, "cpu_calibration.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out["mix_bead_steps_per_s"]), out["mix_ratio_restatement_over_mkl"])


if __name__ == "__main__":
    main()
