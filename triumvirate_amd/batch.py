"""Batched periodic-box bispectra: many catalogues of one geometry (e.g.
lognormal mocks of uneven size) through one engine plan.

Every result equals ``compute_bispec_in_gpp_box`` (reference
threept.py:1624) of that catalogue with the default particle normalisation:
the catalogues are periodised exactly as the reference's simulation-box
driver does (threept.py:1312-1623).  The plan computes all catalogues of a
batch in the same kernels, keeps every buffer between batches, times
candidate launch shapes on its first batch (autotuning) and replays the
whole batch as one captured HIP graph.

Example::

    bb = BatchedBispectrum(boxsize=1000., ngrid=128, binning=kbin,
                           multipoles=[((0, 0, 0), 'full'),
                                       ((2, 0, 2), 'full')],
                           assignment='tsc', precision='fp32', batch=32)
    results = bb.compute([(x0, y0, z0), (x1, y1, z1), ...])
    results[i][q]   # dict of multipole q of catalogue i
"""

import ctypes as C

import numpy as np

from . import _abi, _lib
from .catalogue import ParticleCatalogue
from .parameters import ParameterSet


class BatchedBispectrum:
    """Plan for batches of periodic-box bispectra.

    Parameters
    ----------
    boxsize, ngrid : float / int or length-3 sequences
    binning : :class:`~triumvirate_amd.dataobjs.Binning` (Fourier space)
    multipoles : sequence of (degrees, form) or (degrees, form, idx_bin)
    assignment : {'ngp', 'cic', 'tsc', 'pcs'}
    precision : {'fp32', 'fp64'}
        Shell fields and pair products; bin sums, normalisation and shot
        noise are FP64 in both.
    batch : int
        Catalogues per batch (device memory grows linearly with it).
    capacity : int, optional
        Particles per batch; by default the largest batch of the first
        :meth:`compute` call (a later, larger batch rebuilds the plan).
    autotune, graph : bool
        Time candidate launch shapes on the first batch; capture the batch
        into a HIP graph and replay it.
    """

    def __init__(self, boxsize, ngrid, binning, multipoles=(((0, 0, 0), 'full'),),
                 assignment='tsc', precision='fp32', batch=32, capacity=None,
                 autotune=True, graph=True):
        from .threept import _multipole_structs
        bs = [float(boxsize)] * 3 if np.isscalar(boxsize) else \
            [float(v) for v in boxsize]
        ng = [int(ngrid)] * 3 if np.isscalar(ngrid) else [int(v) for v in ngrid]
        self.boxsize = bs
        self.binning = binning
        self.multipoles = list(multipoles)
        ps = ParameterSet(param_dict={
            'boxsize': dict(zip('xyz', bs)), 'ngrid': dict(zip('xyz', ng)),
            'assignment': assignment, 'interlace': 'false',
            'norm_convention': 'particle', 'precision': precision,
            'degrees': {'ell1': 0, 'ell2': 0, 'ELL': 0},
            'form': 'diag', 'catalogue_type': 'sim',
            'statistic_type': 'bispec',
            'range': [binning.bin_min, binning.bin_max],
            'num_bins': binning.num_bins})
        self._params = _multipole_structs(ps, self.multipoles)
        for p in self._params:
            p.device = 0
        self.batch = int(batch)
        self.capacity = capacity
        self.autotune = bool(autotune)
        self.graph = bool(graph)
        self._bview = _abi.BinningView(binning.bin_edges, binning.bin_centres)
        self._plan = None
        self._lib = _lib.load()

    def __del__(self):
        if getattr(self, '_plan', None) is not None:
            self._lib.trvg_batch_plan_free(self._plan)
            self._plan = None

    def _check(self, st):
        if st != _abi.TRVG_OK:
            from .threept import CLibBackend
            CLibBackend(self._lib)._check(st)

    def _create(self, capacity):
        n = len(self._params)
        arr = (_abi.Params * n)(*self._params)
        plan = C.c_void_p()
        self._check(self._lib.trvg_batch_plan_create(
            arr, n, C.byref(self._bview.struct), self.batch, int(capacity),
            int(self.autotune), int(self.graph), C.byref(plan)))
        self._plan = plan
        self.capacity = int(capacity)

    @property
    def tuning(self):
        """Launch shapes chosen by the autotuner (after the first batch)."""
        if self._plan is None:
            return None
        t = (C.c_int * 5)()
        self._check(self._lib.trvg_batch_plan_info(self._plan, t))
        return {'assign_blocks': t[0], 'pair_blocks': t[1],
                'stat_blocks': t[2], 'pair_row_block_cap': t[3],
                'graph': bool(t[4])}

    def _positions(self, cat):
        if isinstance(cat, ParticleCatalogue):
            pc = cat
        else:
            x, y, z = cat
            pc = ParticleCatalogue(np.asarray(x, float), np.asarray(y, float),
                                   np.asarray(z, float))
        pc.periodise(self.boxsize)
        x, y, z, _, _, _ = pc.columns()
        return x, y, z

    def prepare(self, catalogues):
        """Periodise the catalogues (as the reference's simulation-box driver)
        and pack them into batches: list of (counts, x, y, z)."""
        pos = [self._positions(c) for c in catalogues]
        out = []
        for i in range(0, len(pos), self.batch):
            g = pos[i:i + self.batch]
            out.append((np.array([len(p[0]) for p in g], dtype=np.int64),
                        np.ascontiguousarray(np.concatenate([p[0] for p in g])),
                        np.ascontiguousarray(np.concatenate([p[1] for p in g])),
                        np.ascontiguousarray(np.concatenate([p[2] for p in g]))))
        return out

    def upload(self, batches):
        """Device copies (torch tensors on cuda:0) of prepared batches, for
        :meth:`run` with catalogues resident in device memory."""
        import torch
        return [(c, torch.from_numpy(x).to('cuda'), torch.from_numpy(y).to('cuda'),
                 torch.from_numpy(z).to('cuda')) for c, x, y, z in batches]

    def run(self, batches):
        """Results of prepared batches (host arrays or uploaded tensors)."""
        need = max(int(b[0].sum()) for b in batches)
        if self._plan is not None and need > self.capacity:
            # A larger batch than the plan holds: rebuild it (its tuning
            # and captured graph are redone on the next batch).
            self._lib.trvg_batch_plan_free(self._plan)
            self._plan = None
            self.capacity = need
        if self._plan is None:
            self._create(max(self.capacity or 0, need))
        nmp = len(self._params)
        out = []
        for counts, x, y, z in batches:
            nc = len(counts)
            bufs = [[_abi.ResultBuffers(_abi.result_dim(p.shape,
                                                        self.binning.num_bins,
                                                        p.idx_bin), 'bispec')
                     for p in self._params] for _ in range(nc)]
            arr = (_abi.BispecResult * (nc * nmp))(
                *[b.struct for row in bufs for b in row])
            cp = counts.ctypes.data_as(C.POINTER(C.c_longlong))
            if isinstance(x, np.ndarray):
                self._check(self._lib.trvg_batch_run(
                    self._plan, nc, cp, _abi.dptr(x), _abi.dptr(y), _abi.dptr(z), arr))
            else:
                self._check(self._lib.trvg_batch_run_device(
                    self._plan, nc, cp, x.data_ptr(), y.data_ptr(), z.data_ptr(), arr))
            for i, row in enumerate(bufs):
                res = []
                for q, b in enumerate(row):
                    b.struct.dim = arr[i * nmp + q].dim
                    res.append(b.as_dict())
                out.append(res)
        return out

    def compute(self, catalogues):
        """Bispectra of every catalogue: list (per catalogue) of lists (per
        multipole) of result dicts, as ``compute_bispec_in_gpp_box``."""
        return self.run(self.prepare(catalogues))
