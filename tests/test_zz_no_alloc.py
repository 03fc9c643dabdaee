"""The allocation-free adjoint contraction sdc_jtvec_contract_noalloc, called through the C ABI,
against sdc_jtvec_contract (per_column = 0) on the same inputs: equal bit for bit on two calls in
a row (the second with fewer columns and a column pairing), across the multi-pass case where the
column groups' partial sums exceed the library's fixed scratch array, and the size limit."""

import ctypes as C

import numpy as np
import pytest
import torch

from simpeg_amd import _lib

pytestmark = pytest.mark.gpu

SCRATCH_DOUBLES = 4 << 20


def _engine(shape):
    from simpeg_amd.engine import NodalEngine

    widths = [np.full(n, 10.0) for n in shape]
    eng = NodalEngine(widths, np.zeros(3), bc_type="Robin")
    sigma = np.exp(np.random.default_rng(0).normal(np.log(0.01), 0.5, eng.nC))
    eng.set_sigma(sigma)
    return eng


def _both(eng, u, lam, ncol, ucol):
    lib = eng.lib
    uc = None if ucol is None else torch.as_tensor(np.asarray(ucol, np.int32)).to(eng.device)
    ref = torch.full((eng.nC,), np.nan, dtype=torch.float64, device=eng.device)
    got = torch.full((eng.nC,), np.nan, dtype=torch.float64, device=eng.device)
    _lib.check(lib.sdc_jtvec_contract(eng.h, _lib.ptr(u), _lib.ptr(lam), ncol,
                                      _lib.ptr(uc) if uc is not None else None, 0, _lib.ptr(ref),
                                      eng.nC, eng._stream()), "sdc_jtvec_contract")
    _lib.check(lib.sdc_jtvec_contract_noalloc(eng.h, _lib.ptr(u), _lib.ptr(lam), ncol,
                                              _lib.ptr(uc) if uc is not None else None,
                                              _lib.ptr(got), eng._stream()),
               "sdc_jtvec_contract_noalloc")
    torch.cuda.synchronize(eng.device)
    return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_contract_matches_allocating_entry_point(gpu):
    shape = (64, 64, 48)
    eng = _engine(shape)
    nC = eng.nC
    g = torch.Generator(device=eng.device).manual_seed(3)
    ncol = 200  # 25 groups of 8 columns; 21 fit in the scratch at this size: two passes
    assert (ncol + 7) // 8 > SCRATCH_DOUBLES // nC
    u = torch.randn((ncol, eng.ld), dtype=torch.float64, device=eng.device, generator=g)
    lam = torch.randn((ncol, eng.ld), dtype=torch.float64, device=eng.device, generator=g)
    print("call 1: mesh", shape, "ncol", ncol, flush=True)
    ref, got = _both(eng, u, lam, ncol, None)
    assert np.isfinite(ref).all()
    assert np.array_equal(ref, got)
    # second call, smaller: 9 columns paired with arbitrary u columns (one pass, two groups)
    ucol = np.random.default_rng(4).integers(0, ncol, 9)
    print("call 2: ncol 9 with pairing", flush=True)
    ref2, got2 = _both(eng, u, lam[:9].contiguous(), 9, ucol)
    assert np.array_equal(ref2, got2)
    assert not np.array_equal(ref2, ref)
    # one group: written directly
    ref3, got3 = _both(eng, u, lam, 5, None)
    assert np.array_equal(ref3, got3)
    eng.close()


def test_noalloc_contract_rejects_meshes_beyond_the_scratch(gpu):
    from simpeg_amd.engine import NodalEngine

    shape = (162, 162, 160)  # 4,199,040 cells > 4,194,304
    eng = NodalEngine([np.full(n, 10.0) for n in shape], np.zeros(3), bc_type="Neumann")
    assert eng.nC > SCRATCH_DOUBLES
    buf = torch.zeros(8, dtype=torch.float64, device=eng.device)
    rc = eng.lib.sdc_jtvec_contract_noalloc(eng.h, _lib.ptr(buf), _lib.ptr(buf), 1, None,
                                            _lib.ptr(buf), eng._stream())
    assert rc == _lib.SDC_ERR_ARG
    assert b"scratch" in eng.lib.sdc_last_error()
    eng.close()
