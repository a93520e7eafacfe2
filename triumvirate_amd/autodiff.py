"""Differentiable periodic-box bispectrum B_000 (diag form) on the GPU.

The forward chain is the reference's ``compute_bispec_in_gpp_box``
(threept.py:1624 -> threept.cpp:1473-2188) for degrees (0, 0, 0) and the
``diag`` form, with the mesh built from particle positions by the
reference's assignment rules (field.cpp:569-1112) and the field fluctuation
``delta = n - sum(w) / V`` (field.cpp:1235-1243; ``sum(w)`` is the particle
count for unit weights, as in the reference).  The values are the real part
of ``bk_raw`` (the imaginary part is round-off for a real field).

Gradients come from hand-written adjoints in the engine
(``trvg_diff_assign_backward``, ``trvg_diff_bispec_diag_backward``): with
respect to the mesh field, the particle positions and the particle weights.
Shell fields are recomputed in the backward pass rather than stored.

Only the shot-noise-free part (``bk_raw``) is differentiated; the reference
has no adjoint of its shot-noise terms.

Example::

    B = DifferentiableBispectrum(boxsize=1000., ngrid=256, binning=kbin,
                                 assignment='cic', precision='fp32',
                                 norm_factor=norm)
    bk = B(positions, weights)        # torch float64 (num_bins,)
    loss = (wb * bk).sum()
    loss.backward()                   # positions.grad, weights.grad
"""

import ctypes as C

import numpy as np
import torch

from . import _abi, _lib

_ORDERS = {'ngp': 1, 'cic': 2, 'tsc': 3, 'pcs': 4}


def _check(lib, status):
    if status != _abi.TRVG_OK:
        from .threept import CLibBackend
        CLibBackend(lib)._check(status)


def _sync():
    # The engine runs on its own stream: make torch's pending work on the
    # inputs visible first.
    torch.cuda.current_stream().synchronize()


def _check_positions(pos):
    """The engine reads positions as an (n, 3) row-major float64 device
    array; anything else would be an out-of-bounds device access."""
    if not isinstance(pos, torch.Tensor):
        raise ValueError("positions must be a torch tensor.")
    if not pos.is_cuda:
        raise ValueError("positions must be a CUDA (HIP) tensor.")
    if pos.dtype != torch.float64:
        raise ValueError(f"positions must be float64, not {pos.dtype}.")
    if pos.ndim != 2 or pos.shape[1] != 3:
        raise ValueError("positions must have shape (n, 3), not "
                         f"{tuple(pos.shape)}.")


def _check_weights(w, pos):
    if w is None:
        return
    if not isinstance(w, torch.Tensor):
        raise ValueError("weights must be a torch tensor or None.")
    if not w.is_cuda or w.device != pos.device:
        raise ValueError("weights must be a CUDA (HIP) tensor on the device "
                         "of the positions.")
    if w.dtype != torch.float64:
        raise ValueError(f"weights must be float64, not {w.dtype}.")
    if w.ndim != 1 or w.shape[0] != pos.shape[0]:
        raise ValueError(f"weights must have shape ({pos.shape[0]},), not "
                         f"{tuple(w.shape)}.")


def _check_mesh(mesh, cfg, what='mesh'):
    if not isinstance(mesh, torch.Tensor):
        raise ValueError(f"{what} must be a torch tensor.")
    if not mesh.is_cuda:
        raise ValueError(f"{what} must be a CUDA (HIP) tensor.")
    if tuple(mesh.shape) != tuple(cfg.ngrid):
        raise ValueError(f"{what} must have shape {tuple(cfg.ngrid)}, not "
                         f"{tuple(mesh.shape)}.")
    if mesh.is_complex() or not mesh.is_floating_point():
        raise ValueError(f"{what} must be a real floating-point tensor.")


def _params_on(cfg, t):
    """The ABI parameters with the device of tensor `t`."""
    p = _abi.Params.from_buffer_copy(cfg.params)
    p.device = t.device.index if t.device.index is not None \
        else torch.cuda.current_device()
    return p


class _Config:
    def __init__(self, boxsize, ngrid, binning, assignment, precision,
                 norm_factor):
        bs = [float(boxsize)] * 3 if np.isscalar(boxsize) else \
            [float(v) for v in boxsize]
        ng = [int(ngrid)] * 3 if np.isscalar(ngrid) else [int(v) for v in ngrid]
        if assignment not in _ORDERS:
            raise ValueError(f"Unknown assignment scheme: {assignment!r}.")
        if precision not in ('fp32', 'fp64'):
            raise ValueError("precision must be 'fp32' or 'fp64'.")
        self.boxsize, self.ngrid = bs, ng
        self.volume = bs[0] * bs[1] * bs[2]
        self.dtype = torch.float32 if precision == 'fp32' else torch.float64
        self.norm = float(norm_factor)
        p = _abi.Params()
        for i in range(3):
            p.boxsize[i] = bs[i]
            p.ngrid[i] = ng[i]
        p.assignment_order = _ORDERS[assignment]
        p.interlace = 0
        p.ell1 = p.ell2 = p.ELL = 0
        p.shape = _abi.SHAPE_CODES['diag']
        p.idx_bin = 0
        p.precision = _abi.TRVG_PREC_FP32 if precision == 'fp32' \
            else _abi.TRVG_PREC_FP64
        p.verbose = 20
        p.device = 0
        self.params = p
        self.binning = binning
        self.bview = _abi.BinningView(binning.bin_edges, binning.bin_centres)
        self.nmodes = None


class _Assign(torch.autograd.Function):
    """positions (n, 3) float64, weights (n,) float64 or None -> mesh n(x)."""

    @staticmethod
    def forward(ctx, pos, w, cfg):
        _check_positions(pos)
        _check_weights(w, pos)
        lib = _lib.load()
        pos = pos.detach().contiguous()
        wd = None if w is None else w.detach().contiguous()
        mesh = torch.empty(cfg.ngrid, dtype=cfg.dtype, device=pos.device)
        _sync()
        _check(lib, lib.trvg_diff_assign(
            C.byref(_params_on(cfg, pos)), pos.data_ptr(),
            None if wd is None else wd.data_ptr(), pos.shape[0],
            mesh.data_ptr(), 0))
        ctx.cfg = cfg
        ctx.has_w = w is not None
        ctx.save_for_backward(pos, *(() if wd is None else (wd,)))
        return mesh

    @staticmethod
    def backward(ctx, gmesh):
        lib = _lib.load()
        cfg = ctx.cfg
        saved = ctx.saved_tensors
        pos = saved[0]
        wd = saved[1] if ctx.has_w else None
        _check_mesh(gmesh, cfg, 'mesh gradient')
        gmesh = gmesh.detach().to(device=pos.device,
                                  dtype=cfg.dtype).contiguous()
        gpos = torch.empty_like(pos) if ctx.needs_input_grad[0] else None
        gw = torch.empty_like(wd) if (ctx.has_w and ctx.needs_input_grad[1]) \
            else None
        _sync()
        _check(lib, lib.trvg_diff_assign_backward(
            C.byref(_params_on(cfg, pos)), pos.data_ptr(),
            None if wd is None else wd.data_ptr(), pos.shape[0],
            gmesh.data_ptr(), None if gpos is None else gpos.data_ptr(),
            None if gw is None else gw.data_ptr()))
        return gpos, gw, None


class _BispecDiag(torch.autograd.Function):
    """real mesh delta(x) -> Re B_000(k_b, k_b) for every bin b."""

    @staticmethod
    def forward(ctx, mesh, cfg):
        _check_mesh(mesh, cfg)
        lib = _lib.load()
        mesh = mesh.detach().to(cfg.dtype).contiguous()
        nb = cfg.binning.num_bins
        bk = np.zeros(nb)
        nmodes = (C.c_int * nb)()
        _sync()
        _check(lib, lib.trvg_diff_bispec_diag(
            C.byref(_params_on(cfg, mesh)), C.byref(cfg.bview.struct), cfg.norm,
            mesh.data_ptr(), _abi.dptr(bk), nmodes))
        cfg.nmodes = np.array(nmodes[:], dtype=np.int64)
        ctx.cfg = cfg
        ctx.save_for_backward(mesh)
        return torch.from_numpy(bk).to(mesh.device)

    @staticmethod
    def backward(ctx, gbk):
        lib = _lib.load()
        cfg = ctx.cfg
        (mesh,) = ctx.saved_tensors
        g = np.ascontiguousarray(gbk.detach().double().cpu().numpy())
        grad = torch.empty_like(mesh)
        _sync()
        _check(lib, lib.trvg_diff_bispec_diag_backward(
            C.byref(_params_on(cfg, mesh)), C.byref(cfg.bview.struct), cfg.norm,
            mesh.data_ptr(), _abi.dptr(g), grad.data_ptr()))
        return grad, None


class DifferentiableBispectrum:
    """Periodic-box B_000 diag form with gradients.

    Parameters
    ----------
    boxsize, ngrid : float / int or length-3 sequences
    binning : :class:`~triumvirate_amd.dataobjs.Binning` (Fourier space)
    assignment : {'ngp', 'cic', 'tsc', 'pcs'}
    precision : {'fp32', 'fp64'}
        Mesh and transform precision; bin sums are FP64 in both.
    norm_factor : float
        Multiplies the values (e.g. the particle normalisation
        ``V^2 / N^3`` of calc_bispec_normalisation_from_particles with
        ``nz = N / V``).
    """

    def __init__(self, boxsize, ngrid, binning, assignment='cic',
                 precision='fp32', norm_factor=1.):
        self.cfg = _Config(boxsize, ngrid, binning, assignment, precision,
                           norm_factor)

    @property
    def nmodes(self):
        """Mode counts of the bins (after the first evaluation)."""
        return self.cfg.nmodes

    def mesh(self, positions, weights=None):
        """Field fluctuation delta(x) = n(x) - sum(w) / V (differentiable)."""
        _check_positions(positions)
        _check_weights(weights, positions)
        n = _Assign.apply(positions, weights, self.cfg)
        total = (weights.sum() if weights is not None
                 else float(positions.shape[0]))
        return n - total / self.cfg.volume

    def from_mesh(self, delta):
        """B_000 diag values of a real mesh (differentiable)."""
        _check_mesh(delta, self.cfg)
        return _BispecDiag.apply(delta, self.cfg)

    def __call__(self, positions, weights=None):
        return self.from_mesh(self.mesh(positions, weights))
