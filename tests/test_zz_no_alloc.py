# -*- coding: utf-8 -*-
"""wds_solve_noalloc (the fused forward with the library's own scratch
instead of a caller workspace) against wds_solve, through the C ABI: two
calls in a row, the second with fewer scenarios, bit for bit; the size
limit and the argument checks raise before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

from wdsgpu_amd import _lib
from wdsgpu_amd.generators import planar_network, scenario_batch
from wdsgpu_amd.solver import GGASolver

pytestmark = pytest.mark.gpu

OUTS = ("head", "flow", "emitter", "st", "relerr", "status", "diag_ratio", "resid", "factor")


def _run(s, d, rh, noalloc, polish):
    a, out, keep, single, B = s._args(torch.tensor(d), torch.tensor(rh), None, None, None, 200,
                                      1e-6, True, polish_steps=polish,
                                      want_factor=polish >= 0)
    if noalloc:
        a.workspace = None
        fn = s._lib.wds_solve_noalloc
    else:
        fn = s._lib.wds_solve
    with torch.cuda.device(s.device):
        _lib.check(fn(s._plan, s._stream(), C.byref(a)), "solve")
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("polish", [-1, 3])
def test_noalloc_equals_wds_solve(cuda, polish):
    net = planar_network(1)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    s.cudss_plan(1)                      # uploads the plan (an allocation)
    per = int(s._lib.wds_workspace_bytes(s._plan, 1))
    bmax = (32 << 20) // per
    print("workspace %d bytes per scenario: at most %d scenarios per call" % (per, bmax),
          flush=True)
    for B, seed in ((min(200, bmax), 3), (37, 4)):
        d, rh = scenario_batch(net, B, seed=seed)
        ref = _run(s, d, rh, False, polish)
        got = _run(s, d, rh, True, polish)
        for k in OUTS:
            if ref[k] is not None:
                assert torch.equal(ref[k], got[k]), (B, k)
        assert int(got["info"].abs().sum()) == 0
        print("B=%d polish %d: bitwise equal" % (B, polish), flush=True)


def test_noalloc_refuses_what_it_cannot_serve(cuda):
    net = planar_network(1)
    s = GGASolver(net, device=cuda, dense_status_machine=True)
    d, rh = scenario_batch(net, 2, seed=1)
    a, out, keep, single, B = s._args(torch.tensor(d), torch.tensor(rh), None, None, None, 200,
                                      1e-6, True)
    st = s._stream()
    # plan not uploaded yet on this solver: refused, nothing launched
    a.workspace = None
    s2 = GGASolver(net, device=cuda, dense_status_machine=True)
    with pytest.raises(ValueError, match="not uploaded"):
        _lib.check(s2._lib.wds_solve_noalloc(s2._plan, st, C.byref(a)), "noalloc")
    s.cudss_plan(1)
    # a caller workspace is refused
    a.workspace = C.c_void_p(keep[6].data_ptr())
    with pytest.raises(ValueError, match="workspace must be NULL"):
        _lib.check(s._lib.wds_solve_noalloc(s._plan, st, C.byref(a)), "noalloc")
    # more scenarios than the scratch holds
    a.workspace = None
    a.B = 100000
    with pytest.raises(ValueError, match="more than the"):
        _lib.check(s._lib.wds_solve_noalloc(s._plan, st, C.byref(a)), "noalloc")
