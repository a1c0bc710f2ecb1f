# -*- coding: utf-8 -*-
"""Parameter maps (Hazen-Williams resistance, emitter coefficient) on the
GPU against float64 torch-on-CPU autograd of the same expressions."""
import numpy as np
import pytest
import torch

from wdsgpu_amd.autodiff import emitter_ke, hw_resistance
from wdsgpu_amd.network import LPS_PER_CFS, M_PER_FT

pytestmark = pytest.mark.gpu


def test_hw_resistance_and_gradient(cuda):
    g = np.random.default_rng(3)
    B, L = 5, 257
    C = torch.tensor(g.uniform(80, 150, (B, L)), requires_grad=True)
    ln = torch.tensor(g.uniform(100, 3000, L))
    dm = torch.tensor(g.uniform(0.3, 1.5, L))
    mask = torch.tensor(g.random(L) < 0.9)
    ref = torch.where(mask, 4.727 * ln / C ** 1.852 / dm ** 4.871, torch.zeros(()))
    w = torch.tensor(g.normal(size=(B, L)))
    (ref * w).sum().backward()
    Cg = C.detach().to(cuda).requires_grad_(True)
    r = hw_resistance(Cg, ln.to(cuda), dm.to(cuda), mask=mask.to(cuda))
    (r * w.to(cuda)).sum().backward()
    assert torch.allclose(r.detach().cpu(), ref.detach(), rtol=1e-14, atol=0)
    assert torch.allclose(Cg.grad.cpu(), C.grad, rtol=1e-13, atol=0)
    assert (r.detach().cpu()[:, ~mask] == 0).all()


def test_emitter_coefficient_and_gradient(cuda):
    g = np.random.default_rng(4)
    c = g.uniform(0.0, 0.5, 300)
    c[::7] = 0.0
    Cc = torch.tensor(c, requires_grad=True)
    coef = LPS_PER_CFS ** 2 / M_PER_FT
    pos = Cc > 0
    ref = torch.where(pos, coef / torch.where(pos, Cc, torch.ones(())) ** 2, torch.zeros(()))
    w = torch.tensor(g.normal(size=300))
    (ref * w).sum().backward()
    Cg = Cc.detach().to(cuda).requires_grad_(True)
    ke = emitter_ke(Cg)
    (ke * w.to(cuda)).sum().backward()
    assert torch.allclose(ke.detach().cpu(), ref.detach(), rtol=1e-14, atol=0)
    assert torch.allclose(Cg.grad.cpu(), Cc.grad, rtol=1e-13, atol=0)


@pytest.mark.parametrize("B", [1, 17, 1024])
def test_batch_sums_match_column_sums(cuda, B):
    # the one-launch column sums of the batch-shared gradients: four arrays
    # of ragged widths (not multiples of the 16-column block), against
    # float64 sums on the host, and the same bits on a second call
    import ctypes as C
    from wdsgpu_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(B)
    widths = [5, 784, 908, 33]
    src = [torch.tensor(rng.normal(size=(B, w)), device=cuda) for w in widths]
    dst = [torch.empty(w, dtype=torch.float64, device=cuda) for w in widths]
    arr = C.c_void_p * 4
    n = (C.c_int64 * 4)(*widths)

    def run():
        _lib.check(lib.wds_batch_sums(None, B, 4, arr(*[_lib.ptr(t) for t in src]), n,
                                      arr(*[_lib.ptr(t) for t in dst])), "wds_batch_sums")
        torch.cuda.synchronize()
        return [d.clone() for d in dst]

    first = run()
    for t, d in zip(src, first):
        ref = t.cpu().numpy().sum(axis=0)
        assert np.allclose(d.cpu().numpy(), ref, rtol=1e-13, atol=1e-13 * B)
    assert all(torch.equal(a, b) for a, b in zip(first, run()))
