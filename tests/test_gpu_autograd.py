"""The ring-polymer action as a torch.autograd.Function (rptunnel.autograd): first
derivatives from the device adjoint, second derivatives from the device HVP, checked
against central finite differences (torch.autograd.gradcheck / gradgradcheck) and
against the oracle's UMprime and its dense Hessian (instantonmod.f90:50-89, 155-217
without the band quirk)."""
import numpy as np
import pytest
import torch

import oracle as O
from helpers import oracle_pes, rand_chain, rel, spec_by_name
from rptunnel import Pes, Ring
from rptunnel.autograd import ring_action

pytestmark = pytest.mark.gpu


def _setup(name, ns, betan, seed=3, scale=0.05):
    sp = spec_by_name(name)
    pes = Pes(sp)
    ring = Ring(pes, sp.mass, ns, betan, sp.well1, sp.well2)
    xs = [rand_chain(sp, n, seed + k, scale) for k, n in enumerate(ns)]
    x = torch.tensor(np.concatenate([c.reshape(-1) for c in xs]), dtype=torch.float64,
                     requires_grad=True)
    return sp, pes, ring, xs, x


@pytest.mark.parametrize("name", ["quartic", "coupled2d", "evb"])
def test_gradcheck_and_gradgradcheck(name):
    ns = [5, 3] if name != "evb" else [3]
    sp, pes, ring, xs, x = _setup(name, ns, [9.0] * len(ns))
    f = lambda v: ring_action(v, ring)  # noqa: E731
    assert torch.autograd.gradcheck(f, (x,), eps=1e-6, atol=1e-7, rtol=1e-5)
    assert torch.autograd.gradgradcheck(f, (x,), eps=1e-6, atol=1e-7, rtol=1e-5)


@pytest.mark.parametrize("name", ["coupled2d", "evb"])
def test_autograd_against_oracle(name):
    ns, betan = [7, 4], [11.0, 6.0]
    sp, pes, ring, xs, x = _setup(name, ns, betan, seed=8)
    op = oracle_pes(sp)
    w = torch.tensor([0.7, -1.3], dtype=torch.float64)
    s = (ring_action(x, ring) * w).sum()
    (gx,) = torch.autograd.grad(s, x, create_graph=True)
    off = 0
    want = []
    for k, n in enumerate(ns):
        oc = O.Chain(n, sp.ndim, sp.natom, sp.mass, betan[k], True)
        want.append(float(w[k]) * oc.UMprime(op, xs[k], sp.well1, sp.well2).reshape(-1))
    assert rel(gx.detach().numpy(), np.concatenate(want)) <= 1e-12
    # Hessian-vector product of the weighted sum against the oracle's dense Hessians
    v = torch.tensor(np.random.default_rng(4).standard_normal(x.numel()), dtype=torch.float64)
    (hv,) = torch.autograd.grad(gx, x, grad_outputs=v)
    D = sp.ndof
    for k, n in enumerate(ns):
        oc = O.Chain(n, sp.ndim, sp.natom, sp.mass, betan[k], True)
        H = np.zeros((n * D, n * D))
        m = np.repeat(oc.mass, sp.ndim)
        for i in range(n):
            Hb = op.Vdoubleprime(xs[k][:, i])
            for p in range(D):
                for q in range(D):
                    H[p * n + i, q * n + i] += Hb[p, q]
        for p in range(D):
            kk = m[p] / betan[k] ** 2
            for i in range(n):
                H[p * n + i, p * n + i] += 2 * kk
                if i + 1 < n:
                    H[p * n + i, p * n + i + 1] -= kk
                    H[p * n + i + 1, p * n + i] -= kk
        seg = slice(off, off + n * D)
        assert rel(hv[seg].numpy(), float(w[k]) * H @ v[seg].numpy()) <= 1e-11
        off += n * D


@pytest.mark.parametrize("name", ["coupled2d", "evb"])
def test_device_tensors_match_host_tensors(name):
    """Tensors on the GPU go through the device-memory entry points (no host copies of the
    chains); every derivative equals the host-tensor path's, and both second-derivative
    paths (d/dx and d/dw of the weighted gradient) run on the device."""
    ns, betan = [7, 4, 12], [11.0, 6.0, 9.0]
    sp, pes, ring, xs, x = _setup(name, ns, betan, seed=12)
    dev = torch.device("cuda", 0)
    xd = x.detach().to(dev).requires_grad_(True)
    w = torch.tensor([0.7, -1.3, 0.4], dtype=torch.float64, requires_grad=True)
    wd = w.detach().to(dev).requires_grad_(True)
    v = torch.tensor(np.random.default_rng(5).standard_normal(x.numel()), dtype=torch.float64)
    out = []
    for xx, ww, vv in ((x, w, v), (xd, wd, v.to(dev))):
        e = ring_action(xx, ring)
        (gx,) = torch.autograd.grad((e * ww).sum(), xx, create_graph=True)
        hv, dw = torch.autograd.grad(gx, (xx, ww), grad_outputs=vv)
        out.append([t.detach().cpu().numpy() for t in (e, gx, hv, dw)])
        if xx.is_cuda:
            assert all(t.device == dev for t in (e, gx, hv, dw))
    for h, d in zip(*out):
        assert rel(d, h) <= 1e-14
    # d/dw_c of sum_c w_c grad U_c . v is grad U_c . v
    g, _ = ring.UMforceenergy(xs)
    want = [float(np.dot(g[c].reshape(-1), v.numpy()[ring.off[c]:ring.off[c + 1]])) for c in range(3)]
    assert rel(out[1][3], np.array(want)) <= 1e-13
