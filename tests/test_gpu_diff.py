"""GPU checks of the differentiable periodic-box B_000 diag path
(triumvirate_amd.autodiff over trvg_diff_*):

* forward == the engine's compute_bispec_in_gpp_box bk_raw (FP64: 1e-10;
  FP32: 1e-5 relative to the largest |B|);
* mesh gradient == the numpy adjoint of tests/diff_reference.py (FP64,
  1e-9 relative to the largest component);
* position and weight gradients == central finite differences of the GPU
  forward (FP64; eps = 1e-3 Mpc/h for positions, 1e-3 for weights;
  agreement 1e-5 relative to the largest component);
* FP32 gradients against FP64 gradients (relative to the largest
  component, bound stated in the test).
"""

import numpy as np
import pytest

import diff_reference as DR
import parity as P

pytestmark = pytest.mark.gpu

L = 1000.


def _setup(n, ng, nbins=6, seed=21, kmax=None):
    import torch
    from triumvirate_amd import threept as T
    from triumvirate_amd.dataobjs import Binning
    from triumvirate_amd.parameters import ParameterSet
    d = P.uniform_box_catalogue(n, L, seed=seed)
    kn = np.pi * ng / L
    b = Binning('fourier', 'lin', bin_min=0.01,
                bin_max=kmax or 0.8 * kn, num_bins=nbins)
    ps = ParameterSet(param_dict={
        'boxsize': {'x': L, 'y': L, 'z': L},
        'ngrid': {'x': ng, 'y': ng, 'z': ng}, 'assignment': 'cic',
        'norm_convention': 'particle',
        'degrees': {'ell1': None, 'ell2': None, 'ELL': 0}})
    from triumvirate_amd import _lib
    ref = T._compute_3pt_stats_sim_like(
        T.CLibBackend(_lib.load()), 'bispec', d, paramset=ps,
        degrees=(0, 0, 0), binning=b, form='diag')
    x, y, z, _, _, _ = d.columns()          # re-centred by the driver
    pos = torch.tensor(np.stack([x, y, z], axis=1), device='cuda')
    return ref, b, pos, ref['_norm_factors'][0]


@pytest.mark.parametrize('prec,tol', [('fp64', 1e-10), ('fp32', 1e-5)])
def test_forward_matches_engine(prec, tol):
    from triumvirate_amd.autodiff import DifferentiableBispectrum
    print(f"diff forward {prec}", flush=True)
    ref, b, pos, norm = _setup(20000, 64)
    B = DifferentiableBispectrum(L, 64, b, 'cic', prec, norm)
    bk = B(pos).cpu().numpy()
    assert np.array_equal(B.nmodes, ref['nmodes_1'])
    err = np.max(np.abs(bk - ref['bk_raw'].real)) / np.max(np.abs(ref['bk_raw'].real))
    print(f"  rel err {err:.2e}", flush=True)
    assert err < tol


def test_mesh_gradient_matches_numpy_adjoint():
    import torch
    from triumvirate_amd.autodiff import DifferentiableBispectrum
    print("diff mesh grad vs numpy", flush=True)
    ref, b, pos, norm = _setup(5000, 32, nbins=4)
    B = DifferentiableBispectrum(L, 32, b, 'cic', 'fp64', norm)
    delta = B.mesh(pos).detach().clone().requires_grad_(True)
    gB = torch.tensor(np.random.default_rng(2).standard_normal(4), device='cuda')
    (B.from_mesh(delta) * gB).sum().backward()
    got = delta.grad.cpu().numpy()
    want = DR.adjoint(delta.detach().cpu().numpy(), L, b.bin_edges, norm,
                      gB.cpu().numpy())
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    print(f"  rel err {err:.2e}", flush=True)
    assert err < 1e-9


def _loss_fn(B, wb):
    def f(pos, w=None):
        return float((B(pos, w) * wb).sum())
    return f


def test_position_and_weight_gradients_match_finite_differences():
    import torch
    from triumvirate_amd.autodiff import DifferentiableBispectrum
    print("diff FD positions/weights", flush=True)
    ref, b, pos, norm = _setup(20000, 64)
    rng = np.random.default_rng(5)
    wb = torch.tensor(rng.standard_normal(b.num_bins), device='cuda')
    B = DifferentiableBispectrum(L, 64, b, 'cic', 'fp64', norm)
    w = torch.tensor(rng.uniform(0.5, 1.5, pos.shape[0]), device='cuda')
    p = pos.clone().requires_grad_(True)
    ww = w.clone().requires_grad_(True)
    (B(p, ww) * wb).sum().backward()
    gpos, gw = p.grad.cpu().numpy(), ww.grad.cpu().numpy()
    f = _loss_fn(B, wb)
    eps = 1e-3
    idx = rng.choice(pos.shape[0], 12, replace=False)
    scale_p = np.max(np.abs(gpos))
    for i in idx:
        for a in range(3):
            pp, pm = pos.clone(), pos.clone()
            pp[i, a] += eps
            pm[i, a] -= eps
            fd = (f(pp, w) - f(pm, w)) / (2 * eps)
            assert abs(fd - gpos[i, a]) < 1e-5 * scale_p, (i, a, fd, gpos[i, a])
    scale_w = np.max(np.abs(gw))
    for i in idx[:6]:
        wp, wm = w.clone(), w.clone()
        wp[i] += 1e-3
        wm[i] -= 1e-3
        fd = (f(pos, wp) - f(pos, wm)) / 2e-3
        assert abs(fd - gw[i]) < 1e-5 * scale_w, (i, fd, gw[i])


def test_fp32_gradients_against_fp64():
    import torch
    from triumvirate_amd.autodiff import DifferentiableBispectrum
    print("diff fp32 vs fp64 gradients", flush=True)
    ref, b, pos, norm = _setup(100000, 128, nbins=8)
    wb = torch.tensor(np.random.default_rng(7).standard_normal(8), device='cuda')
    grads = {}
    for prec in ('fp64', 'fp32'):
        B = DifferentiableBispectrum(L, 128, b, 'cic', prec, norm)
        p = pos.clone().requires_grad_(True)
        (B(p) * wb).sum().backward()
        grads[prec] = p.grad.cpu().numpy()
    err = np.max(np.abs(grads['fp32'] - grads['fp64'])) / np.max(np.abs(grads['fp64']))
    print(f"  rel err {err:.2e}", flush=True)
    assert err < 1e-4
