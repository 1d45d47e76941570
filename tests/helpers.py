"""Shared helpers of the tests (parity metrics, spec-to-oracle plumbing)."""
import numpy as np

import oracle as O


def rel(a, b):
    """Normwise relative difference ||a-b|| / ||b|| (||b|| floored at tiny)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def oracle_pes(spec, v0=0.0):
    return O.Pes(spec.kind, spec.ndim, spec.natom, spec.dp, spec.ip, v0)


def rand_chain(spec, n, seed, scale=0.1):
    """A chain (ndof, n) spread from well1 to well2 with Gaussian jitter."""
    rng = np.random.default_rng(seed)
    t = (np.arange(n) + 1.0) / (n + 1.0)
    x = np.outer(spec.well1, 1 - t) + np.outer(spec.well2, t)
    return x + scale * rng.standard_normal(x.shape)


def spec_by_name(name):
    from rptunnel import specs
    return {"quartic": specs.quartic, "coupled2d": specs.coupled2d, "evb": specs.evb,
            "water": lambda: specs.water_cluster(nmol=2),
            "harmonic": lambda: specs.harmonic([1e-3, 2e-3], [0.1, -0.2], [1000.0, 2000.0])}[name]()
