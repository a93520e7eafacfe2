"""The allocation-free differentiable forward (trvg_diff_bispec_diag_noalloc)
against trvg_diff_bispec_diag, through the C ABI: two calls in a row on one
device, the second on a smaller mesh, in FP64 (1e-10 of the largest |B|)
and FP32 (1e-5), mode counts bit-exact; inputs beyond the scratch array
are refused before any device work."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib():
    from triumvirate_amd import _lib as L
    return L.load()


def _params(ng, L, prec):
    from triumvirate_amd import _abi
    p = _abi.Params()
    for a in range(3):
        p.ngrid[a] = ng
        p.boxsize[a] = L
    p.assignment_order = 2
    p.interlace = 0
    p.precision = prec
    p.device = 0
    p.verbose = 20
    return p


def _mesh(ng, prec, seed):
    import torch
    g = torch.Generator(device='cpu').manual_seed(seed)
    m = torch.randn((ng, ng, ng), generator=g, dtype=torch.float64)
    return m.to('cuda', torch.float32 if prec == 32 else torch.float64)


def _both(lib, ng, prec, seed, nb=8):
    from triumvirate_amd import _abi
    from triumvirate_amd.dataobjs import Binning
    L = 1000.
    kb = Binning('fourier', 'lin', bin_min=0.01, bin_max=0.01 + 0.012 * nb,
                 num_bins=nb)
    kbv = _abi.BinningView(kb.bin_edges, kb.bin_centres)
    p = _params(ng, L, prec)
    mesh = _mesh(ng, prec, seed)
    import torch
    torch.cuda.synchronize()
    out = {}
    for name in ('trvg_diff_bispec_diag', 'trvg_diff_bispec_diag_noalloc'):
        bk = np.zeros(nb)
        nm = np.zeros(nb, dtype=np.int32)
        st = getattr(lib, name)(C.byref(p), C.byref(kbv.struct), C.c_double(1e9),
                                C.c_void_p(mesh.data_ptr()), _abi.dptr(bk),
                                nm.ctypes.data_as(C.POINTER(C.c_int)))
        assert st == 0, lib.trvg_last_error().decode()
        out[name] = (bk, nm)
    return out


@pytest.mark.parametrize('prec,tol', [(64, 1e-10), (32, 1e-5)])
def test_noalloc_matches_allocating_forward(prec, tol):
    lib = _lib()
    for ng, seed in ((64, 1), (32, 2)):        # second call: smaller input
        print(f"noalloc fp{prec} {ng}^3 seed {seed}", flush=True)
        out = _both(lib, ng, prec, seed)
        ref, nref = out['trvg_diff_bispec_diag']
        got, ngot = out['trvg_diff_bispec_diag_noalloc']
        assert np.array_equal(ngot, nref)
        fin = np.isfinite(ref)
        assert np.array_equal(np.isfinite(got), fin)
        err = np.max(np.abs(got[fin] - ref[fin])) / np.max(np.abs(ref[fin]))
        print(f"  rel err {err:.2e}", flush=True)
        assert err < tol


def test_noalloc_refuses_oversized_mesh():
    from triumvirate_amd import _abi
    from triumvirate_amd.dataobjs import Binning
    lib = _lib()
    assert lib.trvg_noalloc_scratch_bytes() == 64 << 20
    kb = Binning('fourier', 'lin', bin_min=0.01, bin_max=0.05, num_bins=4)
    kbv = _abi.BinningView(kb.bin_edges, kb.bin_centres)
    p = _params(256, 1000., 64)                # 3 x 256 MiB of FP64 meshes
    bk = np.zeros(4)
    # The size check precedes every device access: a small valid buffer
    # stands in for the mesh and is never read.
    import torch
    small = torch.zeros(8, dtype=torch.float64, device='cuda')
    st = lib.trvg_diff_bispec_diag_noalloc(C.byref(p), C.byref(kbv.struct),
                                           C.c_double(1.),
                                           C.c_void_p(small.data_ptr()),
                                           _abi.dptr(bk), None)
    assert st == _abi.TRVG_ERR_PARAM
    assert 'too large' in lib.trvg_last_error().decode()
