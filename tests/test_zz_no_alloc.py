"""ppmg_p2m_noalloc: the allocation-free P2M, through the C ABI.

Its binning scratch comes from one fixed device array per GPU instead of a
plan's hipMalloc'd buffers; the kernels are ppmg_p2m's.  Checked here:
* two calls in a row, the second with a smaller input (fewer particles on a
  smaller mesh), each bitwise equal to ppmg_p2m on the same input (the tiled
  3D M'4 path in FP64 and FP32c, and the general BSP2 path), while every plan
  allocation is refused (PPMG_FAIL_ALLOC_ABOVE=0), so no buffer came from
  hipMalloc;
* an input too large for the fixed scratch is rejected before any launch
  (field untouched), and a shortfall found only on the general path fails
  with ppm_err_alloc before the field is written.
"""
import ctypes

import numpy as np
import pytest
import torch

from ppm_legacy_amd import workloads as wl

pytestmark = pytest.mark.gpu

MP4, BSP2 = 2, 1


@pytest.fixture(scope="module")
def core():
    from ppm_legacy_amd import core as c
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return c


def _noalloc(core, mesh, x, u, field, lo, kernel, prec):
    """ppmg_p2m_noalloc through ctypes; returns (info, code)."""
    from ppm_legacy_amd import _lib
    L = _lib.lib()
    L.ppmg_p2m_noalloc.restype = ctypes.c_int
    f = core.field_desc(mesh, field, lo)
    info = L.ppmg_p2m_noalloc(mesh.ref(), kernel, prec, ctypes.c_int64(x.shape[0]), core._ptr(x), u.shape[1],
                              core._ptr(u), ctypes.byref(f), core._stream())
    code = _lib.last_error()[1] if info != 0 else 0
    return info, code


def _case(core, N, n_keep, lda, prec, seed):
    """Advected-like input: a jittered lattice in shuffled order (the binning's
    atomic path), n_keep of its particles."""
    xp = wl.jittered_lattice(N, seed=seed)
    rng = np.random.default_rng(seed)
    idx = rng.permutation(len(xp))[:n_keep]
    up = wl.strengths(seed + 1, len(xp), lda)[idx]
    vdt = torch.float64 if prec == core.PREC_F64 else torch.float32
    return core.Mesh(3, N + 1), torch.as_tensor(xp[idx]).cuda(), torch.as_tensor(up).to(vdt).cuda()


@pytest.mark.parametrize("kernel,prec", [(MP4, 0), (MP4, 1), (BSP2, 0)])
def test_noalloc_matches_p2m_twice(core, monkeypatch, kernel, prec):
    cases = [(48, 48 ** 3, 3, 11), (20, 5000, 3, 12)]   # the second call is smaller
    lo = [0, 0, 0] if kernel == MP4 else [-2, -2, -2]
    inputs, refs = [], []
    for N, n, lda, seed in cases:
        mesh, x, u = _case(core, N, n, lda, prec, seed)
        shape = mesh.compact_shape(lda) if kernel == MP4 else mesh.padded_shape(lda, [2, 2, 2])
        ref = torch.empty(shape, dtype=u.dtype, device="cuda")
        core.p2m(mesh, x, u, ref, lo, kernel, prec, core.Plan())
        inputs.append((mesh, x, u, shape))
        refs.append(ref)
    torch.cuda.synchronize()
    # from here on every plan allocation would fail: the allocation-free
    # entry point must not need one
    monkeypatch.setenv("PPMG_FAIL_ALLOC_ABOVE", "0")
    for (mesh, x, u, shape), ref in zip(inputs, refs):
        print(f"noalloc kernel {kernel} prec {prec}: {x.shape[0]} particles, mesh {mesh.cells}", flush=True)
        out = torch.full(shape, 3.0, dtype=u.dtype, device="cuda")
        info, code = _noalloc(core, mesh, x, u, out, lo, kernel, prec)
        assert info == 0, code
        torch.cuda.synchronize()
        assert torch.equal(out, ref)
    monkeypatch.delenv("PPMG_FAIL_ALLOC_ABOVE")


def test_noalloc_rejects_what_does_not_fit(core, monkeypatch):
    from ppm_legacy_amd import _lib
    # 6M particles need 72 MB for keys, slots and the permutation alone
    N = 64
    mesh = core.Mesh(3, N + 1)
    x = torch.rand(6_000_000, 3, dtype=torch.float64, device="cuda")
    u = torch.ones(6_000_000, 1, dtype=torch.float64, device="cuda")
    field = torch.full(mesh.compact_shape(1), 2.0, dtype=torch.float64, device="cuda")
    info, code = _noalloc(core, mesh, x, u, field, [0, 0, 0], MP4, 0)
    assert info == -2 and code == _lib.ERR_ARGUMENT
    torch.cuda.synchronize()
    assert torch.all(field == 2.0)
    # the general path's sorted weights (48 B per particle for BSP2 in 3D,
    # FP64) do not fit beside the binning of 1.5M particles: ppm_err_alloc,
    # field untouched (the tiled deposit, which keeps no sorted copies, would
    # fit, so the general path is forced)
    n = 1_500_000
    mesh2 = core.Mesh(3, N + 1)
    x2 = torch.rand(n, 3, dtype=torch.float64, device="cuda")
    u2 = torch.ones(n, 3, dtype=torch.float64, device="cuda")
    f2 = torch.full(mesh2.padded_shape(3, [2, 2, 2]), 2.0, dtype=torch.float64, device="cuda")
    monkeypatch.setenv("PPMG_GENERAL", "1")
    info, code = _noalloc(core, mesh2, x2, u2, f2, [-2, -2, -2], BSP2, 0)
    monkeypatch.delenv("PPMG_GENERAL")
    assert info == -1 and code == _lib.ERR_ALLOC
    torch.cuda.synchronize()
    assert torch.all(f2 == 2.0)
    # on the tiled deposit the same call fits and matches ppmg_p2m
    ref2 = torch.empty_like(f2)
    core.p2m(mesh2, x2, u2, ref2, [-2, -2, -2], BSP2, 0, core.Plan())
    info, code = _noalloc(core, mesh2, x2, u2, f2, [-2, -2, -2], BSP2, 0)
    assert info == 0, code
    torch.cuda.synchronize()
    assert torch.equal(f2, ref2)
    # and the entry point still works afterwards
    mesh3, x3, u3 = _case(core, 16, 16 ** 3, 1, 0, 21)
    ref = torch.empty(mesh3.compact_shape(1), dtype=torch.float64, device="cuda")
    core.p2m(mesh3, x3, u3, ref, [0, 0, 0], MP4, 0, core.Plan())
    out = torch.empty_like(ref)
    info, code = _noalloc(core, mesh3, x3, u3, out, [0, 0, 0], MP4, 0)
    assert info == 0, code
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
