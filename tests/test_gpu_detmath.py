# -*- coding: utf-8 -*-
"""The kernels' deterministic pow / log equal the CPU checker's bit for bit."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle.detmath import det_log, det_pow
from wdsgpu_amd import _lib

from test_detmath import CASES

pytestmark = pytest.mark.gpu


def _eval(op, x, y, dev):
    lib = _lib.load()
    xt = torch.tensor(x, device=dev)
    yt = torch.tensor(y, device=dev)
    out = torch.empty_like(xt)
    st = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _lib.check(lib.wds_detmath_eval(st, op, xt.numel(), _lib.ptr(xt), _lib.ptr(yt),
                                    _lib.ptr(out)), "wds_detmath_eval")
    return out.cpu().numpy()


def test_kernel_pow_and_log_are_the_checkers_bits(cuda):
    g = np.random.default_rng(11)
    xs, ys = [], []
    for y, lo, hi in CASES:
        x = 10.0 ** g.uniform(lo, hi, 20000)
        xs.append(x)
        ys.append(np.full_like(x, y))
    # pump speed ratios and emitter flows near zero and one
    xs.append(np.concatenate([g.uniform(0.5, 1.5, 5000), 10.0 ** g.uniform(-30, -10, 5000)]))
    ys.append(g.uniform(-1.0, 2.0, 10000))
    x = np.concatenate(xs)
    y = np.concatenate(ys)
    got = _eval(0, x, y, cuda)
    ref = det_pow(x, y)
    bad = np.nonzero(got != ref)[0]
    assert bad.size == 0, (x[bad[:5]], y[bad[:5]], got[bad[:5]], ref[bad[:5]])
    lx = np.concatenate([10.0 ** g.uniform(-300, 300, 50000), 1.0 + g.uniform(-1e-3, 1e-3, 5000)])
    assert np.array_equal(_eval(1, lx, lx, cuda), det_log(lx))
