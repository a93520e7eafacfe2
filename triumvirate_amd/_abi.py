"""ctypes mirror of the C ABI declared in ``include/trv_gpu.h``.

The structures here are plain views over caller-owned numpy arrays; nothing
is copied when a catalogue, binning or result buffer crosses the boundary.
"""

import ctypes as C

import numpy as np

TRVG_OK = 0
TRVG_ERR_PARAM = 1
TRVG_ERR_DATA = 2
TRVG_ERR_DEVICE = 3
TRVG_ERR_ALLOC = 4
TRVG_ERR_INTERNAL = 5

SHAPE_CODES = {'diag': 0, 'off-diag': 1, 'row': 2, 'full': 3, 'triu': 4}

TRVG_PREC_FP64 = 64
TRVG_PREC_FP32 = 32

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Catalogue(C.Structure):
    _fields_ = [
        ('ntotal', C.c_longlong), ('stride', C.c_longlong),
        ('x', _dp), ('y', _dp), ('z', _dp),
        ('nz', _dp), ('ws', _dp), ('wc', _dp),
    ]


class Params(C.Structure):
    _fields_ = [
        ('boxsize', C.c_double * 3), ('ngrid', C.c_int * 3),
        ('assignment_order', C.c_int), ('interlace', C.c_int),
        ('ell1', C.c_int), ('ell2', C.c_int), ('ELL', C.c_int),
        ('shape', C.c_int), ('idx_bin', C.c_int), ('precision', C.c_int),
        ('verbose', C.c_int), ('device', C.c_int),
    ]


class Binning(C.Structure):
    _fields_ = [('num_bins', C.c_int), ('bin_edges', _dp), ('bin_centres', _dp)]


class BispecResult(C.Structure):
    _fields_ = [
        ('dim', C.c_int),
        ('k1_bin', _dp), ('k2_bin', _dp), ('k1_eff', _dp), ('k2_eff', _dp),
        ('nmodes_1', _ip), ('nmodes_2', _ip),
        ('bk_raw', _dp), ('bk_shot', _dp),
    ]


class ThreePCFResult(C.Structure):
    _fields_ = [
        ('dim', C.c_int),
        ('r1_bin', _dp), ('r2_bin', _dp), ('r1_eff', _dp), ('r2_eff', _dp),
        ('npairs_1', _ip), ('npairs_2', _ip),
        ('zeta_raw', _dp), ('zeta_shot', _dp),
    ]


class PowspecResult(C.Structure):
    _fields_ = [('dim', C.c_int), ('kbin', _dp), ('keff', _dp),
                ('nmodes', _ip), ('pk_raw', _dp), ('pk_shot', _dp)]


class TwoPCFResult(C.Structure):
    _fields_ = [('dim', C.c_int), ('rbin', _dp), ('reff', _dp),
                ('npairs', _ip), ('xi', _dp)]


def dptr(a):
    """Pointer to a C-contiguous float64 array (or NULL for None)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags['C_CONTIGUOUS']
    return a.ctypes.data_as(_dp)


def iptr(a):
    assert a.dtype == np.intc and a.flags['C_CONTIGUOUS']
    return a.ctypes.data_as(_ip)


class CatalogueView:
    """Keeps the column arrays alive while the C view is in use."""

    def __init__(self, x, y, z, nz=None, ws=None, wc=None):
        cols = [np.ascontiguousarray(c, dtype=np.float64) if c is not None
                else None for c in (x, y, z, nz, ws, wc)]
        n = len(cols[0])
        for c in cols:
            if c is not None and len(c) != n:
                raise ValueError(
                    "Inconsistent particle data dimensions (source=extdata).")
        self._cols = cols
        self.struct = Catalogue(n, 1, *[dptr(c) for c in cols])

    @property
    def ntotal(self):
        return self.struct.ntotal


class BinningView:
    def __init__(self, edges, centres):
        self._edges = np.ascontiguousarray(edges, dtype=np.float64)
        self._centres = np.ascontiguousarray(centres, dtype=np.float64)
        self.struct = Binning(len(self._centres), dptr(self._edges),
                              dptr(self._centres))


class ResultBuffers:
    """Caller-owned output arrays for one bispectrum / 3PCF call."""

    def __init__(self, dim, kind):
        self.kind = kind
        self.dim = dim
        self.b1 = np.zeros(dim)
        self.b2 = np.zeros(dim)
        self.e1 = np.zeros(dim)
        self.e2 = np.zeros(dim)
        self.n1 = np.zeros(dim, dtype=np.intc)
        self.n2 = np.zeros(dim, dtype=np.intc)
        self.raw = np.zeros(2 * dim)
        self.shot = np.zeros(2 * dim)
        cls = BispecResult if kind == 'bispec' else ThreePCFResult
        self.struct = cls(
            dim, dptr(self.b1), dptr(self.b2), dptr(self.e1), dptr(self.e2),
            iptr(self.n1), iptr(self.n2), dptr(self.raw), dptr(self.shot))

    def as_dict(self):
        dim = self.struct.dim
        raw = self.raw[:2 * dim].view(np.complex128)
        shot = self.shot[:2 * dim].view(np.complex128)
        if self.kind == 'bispec':
            return {
                'k1_bin': self.b1[:dim].copy(), 'k2_bin': self.b2[:dim].copy(),
                'k1_eff': self.e1[:dim].copy(), 'k2_eff': self.e2[:dim].copy(),
                'nmodes_1': self.n1[:dim].astype(np.int64),
                'nmodes_2': self.n2[:dim].astype(np.int64),
                'bk_raw': raw.copy(), 'bk_shot': shot.copy(),
            }
        return {
            'r1_bin': self.b1[:dim].copy(), 'r2_bin': self.b2[:dim].copy(),
            'r1_eff': self.e1[:dim].copy(), 'r2_eff': self.e2[:dim].copy(),
            'npairs_1': self.n1[:dim].astype(np.int64),
            'npairs_2': self.n2[:dim].astype(np.int64),
            'zeta_raw': raw.copy(), 'zeta_shot': shot.copy(),
        }


class TwoPtBuffers:
    """Caller-owned output arrays for one power spectrum ('powspec') or
    correlation function ('2pcf') call."""

    def __init__(self, num_bins, kind):
        self.kind = kind
        self.bin = np.zeros(num_bins)
        self.eff = np.zeros(num_bins)
        self.cnt = np.zeros(num_bins, dtype=np.intc)
        self.raw = np.zeros(2 * num_bins)
        self.shot = np.zeros(2 * num_bins)
        if kind == 'powspec':
            self.struct = PowspecResult(num_bins, dptr(self.bin), dptr(self.eff),
                                        iptr(self.cnt), dptr(self.raw),
                                        dptr(self.shot))
        else:
            self.struct = TwoPCFResult(num_bins, dptr(self.bin), dptr(self.eff),
                                       iptr(self.cnt), dptr(self.raw))

    def as_dict(self):
        n = self.struct.dim
        raw = self.raw[:2 * n].view(np.complex128).copy()
        if self.kind == 'powspec':
            return {'kbin': self.bin[:n].copy(), 'keff': self.eff[:n].copy(),
                    'nmodes': self.cnt[:n].astype(np.int64), 'pk_raw': raw,
                    'pk_shot': self.shot[:2 * n].view(np.complex128).copy()}
        return {'rbin': self.bin[:n].copy(), 'reff': self.eff[:n].copy(),
                'npairs': self.cnt[:n].astype(np.int64), 'xi': raw}


def result_dim(shape_code, num_bins, idx_bin):
    """Data-vector length per shape (compute_bispec dv_dim switch)."""
    if shape_code in (0, 2):
        return num_bins
    if shape_code == 1:
        return num_bins - abs(idx_bin)
    if shape_code == 3:
        return num_bins * num_bins
    if shape_code == 4:
        return num_bins * (num_bins + 1) // 2
    raise ValueError(f"Unknown shape code {shape_code}.")


def declare_common(lib, prefix):
    """Attach argtypes/restypes for the shared call shapes."""
    P = C.POINTER
    f = getattr(lib, prefix + 'compute_bispec')
    f.argtypes = [P(Catalogue), P(Catalogue), _dp, _dp, P(Params), P(Binning),
                  C.c_double, P(BispecResult)]
    f.restype = C.c_int
    f = getattr(lib, prefix + 'compute_3pcf')
    f.argtypes = [P(Catalogue), P(Catalogue), _dp, _dp, P(Params), P(Binning),
                  C.c_double, P(ThreePCFResult)]
    f.restype = C.c_int
    f = getattr(lib, prefix + 'compute_bispec_in_gpp_box')
    f.argtypes = [P(Catalogue), P(Params), P(Binning), C.c_double,
                  P(BispecResult)]
    f.restype = C.c_int
    f = getattr(lib, prefix + 'compute_3pcf_in_gpp_box')
    f.argtypes = [P(Catalogue), P(Params), P(Binning), C.c_double,
                  P(ThreePCFResult)]
    f.restype = C.c_int
    f = getattr(lib, prefix + 'calc_bispec_normalisation_from_particles')
    f.argtypes = [P(Catalogue), C.c_double, _dp]
    f.restype = C.c_int
    f = getattr(lib, prefix + 'calc_bispec_normalisation_from_mesh')
    f.argtypes = [P(Catalogue), P(Params), C.c_double, _dp]
    f.restype = C.c_int
    f = getattr(lib, prefix + 'compute_3pcf_window')
    f.argtypes = [P(Catalogue), _dp, P(Params), P(Binning), C.c_double,
                  C.c_double, C.c_int, C.c_int, C.c_int, P(ThreePCFResult)]
    f.restype = C.c_int
    for name in ('calc_powspec_normalisation_from_particles',
                 'calc_powspec_shotnoise_from_particles'):
        f = getattr(lib, prefix + name)
        f.argtypes = [P(Catalogue), C.c_double, _dp]
        f.restype = C.c_int
    f = getattr(lib, prefix + 'calc_powspec_normalisation_from_mesh')
    f.argtypes = [P(Catalogue), P(Params), C.c_double, _dp]
    f.restype = C.c_int
    f = getattr(lib, prefix + 'calc_powspec_normalisation_from_meshes')
    f.argtypes = [P(Catalogue), P(Catalogue), P(Params), C.c_double, _dp]
    f.restype = C.c_int
    f = getattr(lib, prefix + 'compute_powspec')
    f.argtypes = [P(Catalogue), P(Catalogue), _dp, _dp, P(Params), P(Binning),
                  C.c_double, P(PowspecResult)]
    f.restype = C.c_int
    f = getattr(lib, prefix + 'compute_corrfunc')
    f.argtypes = [P(Catalogue), P(Catalogue), _dp, _dp, P(Params), P(Binning),
                  C.c_double, P(TwoPCFResult)]
    f.restype = C.c_int
    f = getattr(lib, prefix + 'compute_powspec_in_gpp_box')
    f.argtypes = [P(Catalogue), P(Params), P(Binning), C.c_double,
                  P(PowspecResult)]
    f.restype = C.c_int
    f = getattr(lib, prefix + 'compute_corrfunc_in_gpp_box')
    f.argtypes = [P(Catalogue), P(Params), P(Binning), C.c_double,
                  P(TwoPCFResult)]
    f.restype = C.c_int
    f = getattr(lib, prefix + 'compute_corrfunc_window')
    f.argtypes = [P(Catalogue), _dp, P(Params), P(Binning), C.c_double,
                  C.c_double, P(TwoPCFResult)]
    f.restype = C.c_int
    f = getattr(lib, prefix + 'result_dim')
    f.argtypes = [C.c_int, C.c_int, C.c_int, _ip]
    f.restype = C.c_int
    f = getattr(lib, prefix + 'last_error')
    f.argtypes = []
    f.restype = C.c_char_p


def declare_extensions(lib):
    """Attach argtypes for engine entry points beyond the reference's
    interface (present only in the HIP library)."""
    P = C.POINTER
    vp = C.c_void_p
    lib.trvg_catalogue_upload.argtypes = [P(Catalogue), _dp, C.c_int, P(vp)]
    lib.trvg_catalogue_upload.restype = C.c_int
    lib.trvg_catalogue_upload_observer.argtypes = [P(Catalogue), _dp, C.c_int,
                                                   P(vp)]
    lib.trvg_catalogue_upload_observer.restype = C.c_int
    lib.trvg_compute_bispec_multi_observer.argtypes = [
        P(Catalogue), P(Catalogue), _dp, P(Params), C.c_int, P(Binning), _dp,
        P(BispecResult)]
    lib.trvg_compute_bispec_multi_observer.restype = C.c_int
    lib.trvg_catalogue_free.argtypes = [vp]
    lib.trvg_catalogue_free.restype = C.c_int
    lib.trvg_calc_bispec_normalisation_from_mesh_dev.argtypes = [
        vp, P(Params), C.c_double, _dp]
    lib.trvg_calc_bispec_normalisation_from_mesh_dev.restype = C.c_int
    lib.trvg_compute_bispec_dev.argtypes = [vp, vp, P(Params), P(Binning),
                                            C.c_double, P(BispecResult)]
    lib.trvg_compute_bispec_dev.restype = C.c_int
    lib.trvg_compute_3pcf_dev.argtypes = [vp, vp, P(Params), P(Binning),
                                          C.c_double, P(ThreePCFResult)]
    lib.trvg_compute_3pcf_dev.restype = C.c_int
    lib.trvg_release_workspace.argtypes = []
    lib.trvg_release_workspace.restype = C.c_int
    lib.trvg_mesh_assign.argtypes = [P(Catalogue), _dp, P(Params), C.c_int,
                                     C.c_int, C.c_int, _dp]
    lib.trvg_mesh_assign.restype = C.c_int
    lib.trvg_fft3d.argtypes = [P(Params), _dp, C.c_int]
    lib.trvg_fft3d.restype = C.c_int
    lib.trvg_compute_bispec_multi_dev.argtypes = [
        vp, vp, P(Params), C.c_int, P(Binning), _dp, P(BispecResult)]
    lib.trvg_compute_bispec_multi_dev.restype = C.c_int
    lib.trvg_compute_bispec_multi.argtypes = [
        P(Catalogue), P(Catalogue), _dp, _dp, P(Params), C.c_int, P(Binning),
        _dp, P(BispecResult)]
    lib.trvg_compute_bispec_multi.restype = C.c_int
    lib.trvg_profile_enable.argtypes = [C.c_int]
    lib.trvg_profile_enable.restype = C.c_int
    _vp = C.c_void_p
    lib.trvg_batch_plan_create.argtypes = [P(Params), C.c_int, P(Binning), C.c_int,
                                           C.c_longlong, C.c_int, C.c_int,
                                           P(C.c_void_p)]
    lib.trvg_batch_plan_create.restype = C.c_int
    lib.trvg_batch_run.argtypes = [_vp, C.c_int, P(C.c_longlong), _dp, _dp, _dp,
                                   P(BispecResult)]
    lib.trvg_batch_run.restype = C.c_int
    lib.trvg_batch_run_device.argtypes = [_vp, C.c_int, P(C.c_longlong), _vp, _vp,
                                          _vp, P(BispecResult)]
    lib.trvg_batch_run_device.restype = C.c_int
    lib.trvg_batch_plan_info.argtypes = [_vp, P(C.c_int)]
    lib.trvg_batch_plan_info.restype = C.c_int
    lib.trvg_batch_plan_free.argtypes = [_vp]
    lib.trvg_batch_plan_free.restype = C.c_int
    lib.trvg_diff_assign.argtypes = [P(Params), _vp, _vp, C.c_longlong, _vp, C.c_int]
    lib.trvg_diff_assign.restype = C.c_int
    lib.trvg_diff_assign_backward.argtypes = [P(Params), _vp, _vp, C.c_longlong,
                                              _vp, _vp, _vp]
    lib.trvg_diff_assign_backward.restype = C.c_int
    lib.trvg_diff_bispec_diag.argtypes = [P(Params), P(Binning), C.c_double,
                                          _vp, _dp, P(C.c_int)]
    lib.trvg_diff_bispec_diag.restype = C.c_int
    lib.trvg_diff_bispec_diag_noalloc.argtypes = [P(Params), P(Binning),
                                                  C.c_double, _vp, _dp,
                                                  P(C.c_int)]
    lib.trvg_diff_bispec_diag_noalloc.restype = C.c_int
    lib.trvg_noalloc_scratch_bytes.argtypes = []
    lib.trvg_noalloc_scratch_bytes.restype = C.c_longlong
    lib.trvg_diff_bispec_diag_backward.argtypes = [P(Params), P(Binning),
                                                   C.c_double, _vp, _dp, _vp]
    lib.trvg_diff_bispec_diag_backward.restype = C.c_int
    lib.trvg_copy_calibration.argtypes = [C.c_longlong, _dp]
    lib.trvg_copy_calibration.restype = C.c_int
    lib.trvg_profile_only.argtypes = [C.c_int]
    lib.trvg_profile_only.restype = C.c_int
    lib.trvg_profile_read.argtypes = [C.c_int, _dp, P(C.c_longlong), _dp]
    lib.trvg_profile_read.restype = C.c_int
    lib.trvg_profile_class_name.argtypes = [C.c_int]
    lib.trvg_profile_class_name.restype = C.c_char_p
    lib.trvg_profile_class_count.argtypes = []
    lib.trvg_profile_class_count.restype = C.c_int
    return lib
