"""Three-point statistics: bispectrum and three-point correlation function.

Public surface mirrors ``triumvirate.threept``:

* :func:`compute_bispec` (reference threept.py:743),
* :func:`compute_3pcf` (threept.py:868),
* :func:`compute_bispec_in_gpp_box` (threept.py:1624),
* :func:`compute_3pcf_in_gpp_box` (threept.py:1726),
* :func:`compute_3pcf_window` (threept.py:1832),

with the same arguments and result dictionaries.  The catalogue preparation
follows the reference drivers ``_compute_3pt_stats_survey_like``
(threept.py:384-741) and ``_compute_3pt_stats_sim_like``
(threept.py:1312-1623): lines of sight are computed before the catalogues
are shifted into the box; survey catalogues are centred or padded with the
random catalogue as reference; simulation catalogues are periodised and get
``nz = N/V`` when it is missing; ``alpha = wstotal_data / wstotal_rand``;
the normalisation is chosen by ``norm_convention``.

The mesh work runs in the HIP engine (``libtrvgpu.so``) through its C ABI.
The algorithm backend is passed explicitly to the private drivers (as the
reference passes ``threept_algofunc``) so that one host path serves every
backend.
"""

import contextlib
import ctypes as C

import numpy as np

from . import _abi
from . import io as _io
from .dataobjs import Binning
from .parameters import (InvalidParameterError, ParameterSet,
                         fetch_paramset_template, set_ngrid_from_cutoff)


class InvalidDataError(ValueError):
    """Maps ``trv::sys::InvalidDataError``."""


# ---------------------------------------------------------------------------
# Backend over a C library exporting the trvg_ call shapes
# ---------------------------------------------------------------------------

class CLibBackend:
    """Adapter from the host drivers to a library with the ABI of
    ``include/trv_gpu.h`` (function names ``<prefix>compute_bispec`` ...)."""

    def __init__(self, lib, prefix='trvg_'):
        self.lib = lib
        self.prefix = prefix

    def _fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def _check(self, status):
        if status == _abi.TRVG_OK:
            return
        msg = self._fn('last_error')().decode(errors='replace')
        if status == _abi.TRVG_ERR_PARAM:
            raise InvalidParameterError(msg)
        if status == _abi.TRVG_ERR_DATA:
            raise InvalidDataError(msg)
        from ._lib import EngineError
        raise EngineError(f"[status {status}] {msg}")

    def norm_particles(self, cat, alpha):
        out = C.c_double(0.)
        self._check(self._fn('calc_bispec_normalisation_from_particles')(
            C.byref(cat.struct), C.c_double(alpha), C.byref(out)))
        return out.value

    def norm_mesh(self, cat, params, alpha, handle=None):
        out = C.c_double(0.)
        if handle is not None:
            self._check(self._fn('calc_bispec_normalisation_from_mesh_dev')(
                handle, C.byref(params), C.c_double(alpha), C.byref(out)))
            return out.value
        self._check(self._fn('calc_bispec_normalisation_from_mesh')(
            C.byref(cat.struct), C.byref(params), C.c_double(alpha),
            C.byref(out)))
        return out.value

    def derives_los(self):
        """Whether the library forms lines of sight on the device from an
        observer position (trvg_catalogue_upload_observer)."""
        return hasattr(self.lib, self.prefix + 'catalogue_upload_observer')

    @contextlib.contextmanager
    def resident(self, cat, los=None, observer=None):
        """Device-resident copy of a catalogue for the duration of the block
        (None for a library without device catalogues), so that the mesh
        normalisation and the estimator share one upload.  With `observer`
        and no `los`, the lines of sight are formed on the device."""
        if not hasattr(self.lib, self.prefix + 'catalogue_upload'):
            yield None
            return
        h = C.c_void_p()
        if los is None and observer is not None:
            obs = np.ascontiguousarray(observer, dtype=np.float64)
            self._check(self._fn('catalogue_upload_observer')(
                C.byref(cat.struct), _abi.dptr(obs), -1, C.byref(h)))
        else:
            self._check(self._fn('catalogue_upload')(
                C.byref(cat.struct), _abi.dptr(los), -1, C.byref(h)))
        try:
            yield h
        finally:
            self._fn('catalogue_free')(h)

    def _result(self, params, binning, kind):
        dim = _abi.result_dim(params.shape, binning.struct.num_bins,
                              params.idx_bin)
        return _abi.ResultBuffers(dim, kind)

    def survey(self, kind, data, rand, los_d, los_r, params, binning, norm,
               handles=None):
        res = self._result(params, binning, kind)
        stem = 'compute_bispec' if kind == 'bispec' else 'compute_3pcf'
        if handles is not None:
            self._check(self._fn(stem + '_dev')(
                handles[0], handles[1], C.byref(params),
                C.byref(binning.struct), C.c_double(norm),
                C.byref(res.struct)))
            return res.as_dict()
        self._check(self._fn(stem)(
            C.byref(data.struct), C.byref(rand.struct), _abi.dptr(los_d),
            _abi.dptr(los_r), C.byref(params), C.byref(binning.struct),
            C.c_double(norm), C.byref(res.struct)))
        return res.as_dict()

    def multi(self, data, rand, los_d, los_r, params_list, binning, norms,
              handles=None):
        """Several bispectrum multipoles in one engine call (engine
        extension trvg_compute_bispec_multi)."""
        n = len(params_list)
        arr = (_abi.Params * n)(*params_list)
        res = [self._result(p, binning, 'bispec') for p in params_list]
        outs = (_abi.BispecResult * n)(*[r.struct for r in res])
        nrm = np.ascontiguousarray(norms, dtype=np.float64)
        if handles is not None:
            self._check(self._fn('compute_bispec_multi_dev')(
                handles[0], handles[1], arr, n, C.byref(binning.struct),
                _abi.dptr(nrm), outs))
        else:
            self._check(self._fn('compute_bispec_multi')(
                C.byref(data.struct), C.byref(rand.struct) if rand else None,
                _abi.dptr(los_d), _abi.dptr(los_r), arr, n,
                C.byref(binning.struct), _abi.dptr(nrm), outs))
        for r, o in zip(res, outs):
            r.struct.dim = o.dim
        return [r.as_dict() for r in res]

    def sim(self, kind, data, params, binning, norm, handle=None):
        res = self._result(params, binning, kind)
        if handle is not None:
            fn = self._fn('compute_bispec_dev' if kind == 'bispec'
                          else 'compute_3pcf_dev')
            self._check(fn(handle, None, C.byref(params),
                           C.byref(binning.struct), C.c_double(norm),
                           C.byref(res.struct)))
            return res.as_dict()
        fn = self._fn('compute_bispec_in_gpp_box' if kind == 'bispec'
                      else 'compute_3pcf_in_gpp_box')
        self._check(fn(C.byref(data.struct), C.byref(params),
                       C.byref(binning.struct), C.c_double(norm),
                       C.byref(res.struct)))
        return res.as_dict()


def _engine_backend():
    from . import _lib
    return CLibBackend(_lib.load(), 'trvg_')


# ---------------------------------------------------------------------------
# Parameters
# ---------------------------------------------------------------------------

def _parse_degrees(degrees):
    if isinstance(degrees, str):
        if len(degrees) != 3:
            raise ValueError("`degrees` string must have length 3.")
        return tuple(int(c) for c in degrees)
    ell1, ell2, ELL = degrees
    return int(ell1), int(ell2), int(ELL)


def _amalgamate_parameters(paramset=None, params_sampling=None, degrees=None,
                           binning=None, form=None, idx_bin=None,
                           precision=None, **type_kwargs):
    """Build the full parameter set (reference threept.py:53-194)."""
    if paramset is None:
        if params_sampling is None:
            raise ValueError(
                "Either `paramset` or `sampling_params` must be provided.")
        if degrees is None or binning is None or form is None:
            raise ValueError(
                "`degrees`, `binning` and `form` must be provided when "
                "`paramset` is None.")
        if form in ('row', 'off-diag') and idx_bin is None:
            raise ValueError(
                "`idx_bin` must be provided when `paramset` is None and "
                "`form` is 'row' or 'off-diag'.")
        pdict = fetch_paramset_template('dict')
    else:
        pdict = {k: v for k, v in paramset.items()}
        pdict = {k: (dict(v) if isinstance(v, dict) else v)
                 for k, v in pdict.items()}

    if params_sampling:
        ps = params_sampling
        for key in ('expand', 'alignment', 'assignment', 'interlace',
                    'cutoff_nyq'):
            if key in ps:
                pdict[key] = ps[key]
        if 'boxpad' in ps and 'gridpad' in ps:
            raise ValueError(
                "Cannot set both 'boxpad` and 'gridpad' in `params_sampling`.")
        if 'boxpad' in ps:
            pdict['padscale'], pdict['padfactor'] = 'box', ps['boxpad']
        if 'gridpad' in ps:
            pdict['padscale'], pdict['padfactor'] = 'grid', ps['gridpad']
        if 'boxsize' in ps:
            bs = ps['boxsize']
            bs = [bs] * 3 if np.isscalar(bs) else list(bs)
            pdict['boxsize'] = dict(zip('xyz', bs))
        if 'ngrid' in ps:
            ng = ps['ngrid']
            ng = [ng] * 3 if np.isscalar(ng) else list(ng)
            pdict['ngrid'] = dict(zip('xyz', ng))

    if degrees is not None:
        ell1, ell2, ELL = _parse_degrees(degrees)
        pdict['degrees'] = {'ell1': ell1, 'ell2': ell2, 'ELL': ELL}
    if binning is not None:
        pdict['range'] = [binning.bin_min, binning.bin_max]
        pdict['num_bins'] = binning.num_bins
    if form is not None:
        pdict['form'] = form
    if idx_bin is not None:
        pdict['idx_bin'] = idx_bin
    if precision is not None:
        pdict['precision'] = precision
    pdict.update(type_kwargs)
    return ParameterSet(param_dict=pdict)


def _params_struct(paramset):
    """Pack the validated parameter set into the ABI struct."""
    box = paramset['boxsize']
    ngrid = paramset['ngrid']
    deg = paramset['degrees']
    if any(box[a] is None for a in 'xyz') or any(ngrid[a] is None for a in 'xyz'):
        raise InvalidParameterError("`boxsize` and `ngrid` must be set.")
    if any(deg[k] is None for k in ('ell1', 'ell2', 'ELL')):
        raise InvalidParameterError("`degrees` must be set.")
    shape = paramset['shape']
    if shape not in _abi.SHAPE_CODES:
        raise InvalidParameterError(
            f"Three-point statistic form is not recognised: `form` = "
            f"'{paramset['form']}'.")
    p = _abi.Params()
    for i, a in enumerate('xyz'):
        p.boxsize[i] = float(box[a])
        p.ngrid[i] = int(ngrid[a])
    p.assignment_order = int(paramset['assignment_order'])
    p.interlace = 1 if paramset['interlace'] is True else 0
    p.ell1, p.ell2, p.ELL = int(deg['ell1']), int(deg['ell2']), int(deg['ELL'])
    p.shape = _abi.SHAPE_CODES[shape]
    p.idx_bin = int(paramset['idx_bin'] or 0)
    p.precision = 32 if paramset['precision'] == 'fp32' else 64
    p.verbose = int(paramset['verbose'] or 20)
    p.device = -1
    return p


def _view(catalogue):
    return _abi.CatalogueView(*catalogue.engine_columns())


def _binning_view(binning):
    return _abi.BinningView(binning.bin_edges, binning.bin_centres)


# ---------------------------------------------------------------------------
# Drivers
# ---------------------------------------------------------------------------

def _select_norm(paramset, norm_part, norm_mesh):
    conv = paramset['norm_convention']
    if conv == 'none':
        return 1.
    if conv == 'particle':
        return norm_part
    if conv == 'mesh':
        return norm_mesh
    # 'mesh-mixed' is rejected for three-point statistics by validation.
    raise InvalidParameterError(
        f"Normalisation convention '{conv}' does not apply to three-point "
        "statistics.")


def _multipole_structs(paramset, multipoles):
    """ABI parameter structs for a list of (degrees, form[, idx_bin])."""
    out = []
    for mp in multipoles:
        deg, form = mp[0], mp[1]
        idx = mp[2] if len(mp) > 2 else None
        ell1, ell2, ELL = _parse_degrees(deg)
        pdict = {k: (dict(v) if isinstance(v, dict) else v)
                 for k, v in paramset.items()}
        pdict['degrees'] = {'ell1': ell1, 'ell2': ell2, 'ELL': ELL}
        pdict['form'] = form
        pdict['idx_bin'] = idx
        out.append(_params_struct(ParameterSet(param_dict=pdict)))
    return out


def _compute_3pt_stats_survey_like(backend, kind, catalogue_data,
                                   catalogue_rand, los_data=None,
                                   los_rand=None, paramset=None,
                                   params_sampling=None, degrees=None,
                                   binning=None, form=None, idx_bin=None,
                                   precision=None, logger=None,
                                   multipoles=None, save=False):
    stype = 'bispec' if kind == 'bispec' else '3pcf'
    paramset = _amalgamate_parameters(
        paramset=paramset, params_sampling=params_sampling, degrees=degrees,
        binning=binning, form=form, idx_bin=idx_bin, precision=precision,
        catalogue_type='survey', statistic_type=stype)
    if binning is None:
        binning = Binning.from_parameter_set(paramset)

    # Lines of sight in the observer frame, before alignment
    # (threept.py:525-528 precede :566-582).  A library that forms them on
    # the device from an observer position is given that position instead
    # of two n x 3 arrays (the alignment below only shifts coordinates).
    derive = (los_data is None and los_rand is None and
              getattr(backend, 'derives_los', lambda: False)())
    if not derive:
        if los_data is None:
            los_data = catalogue_data.compute_los()
        if los_rand is None:
            los_rand = catalogue_rand.compute_los()
        los_data = np.ascontiguousarray(los_data, dtype=np.float64)
        los_rand = np.ascontiguousarray(los_rand, dtype=np.float64)
    offset0 = catalogue_data._offset.copy()

    if any(paramset['boxsize'][a] is None for a in 'xyz'):
        for a in 'xyz':
            lo, hi = catalogue_rand.bounds[a]
            paramset['boxsize'][a] = (hi - lo) * paramset['expand']
        paramset._validate()
    # threept.py:554-556: meshes from the Nyquist cutoff when unset.
    if any(paramset['ngrid'][a] is None for a in 'xyz'):
        set_ngrid_from_cutoff(paramset)

    boxsize = [paramset['boxsize'][a] for a in 'xyz']
    if paramset['alignment'] == 'centre':
        catalogue_data.centre(boxsize, catalogue_ref=catalogue_rand)
    elif paramset['alignment'] == 'pad':
        if paramset['padscale'] == 'box':
            kw = {'boxsize_pad': paramset['padfactor']}
        else:
            kw = {'ngrid': [paramset['ngrid'][a] for a in 'xyz'],
                  'ngrid_pad': paramset['padfactor']}
        catalogue_data.pad(boxsize, catalogue_ref=catalogue_rand, **kw)

    data_v, rand_v = _view(catalogue_data), _view(catalogue_rand)
    params = _params_struct(paramset)
    alpha = catalogue_data.wstotal / catalogue_rand.wstotal
    observer = -(catalogue_data._offset - offset0) if derive else None

    norm_part = backend.norm_particles(rand_v, alpha)
    # One upload of each catalogue serves the mesh normalisation and the
    # estimator.
    with backend.resident(rand_v, los_rand, observer) as hr, \
            backend.resident(data_v, los_data, observer) as hd:
        handles = (hd, hr) if hr is not None else None
        norm_mesh = backend.norm_mesh(rand_v, params, alpha, handle=hr)
        norm_factor = _select_norm(paramset, norm_part, norm_mesh)
        if multipoles is not None:
            plist = _multipole_structs(paramset, multipoles)
            return backend.multi(data_v, rand_v, los_data, los_rand, plist,
                                 _binning_view(binning),
                                 [norm_factor] * len(plist), handles=handles)
        results = backend.survey(
            kind, data_v, rand_v, los_data, los_rand, params,
            _binning_view(binning), norm_factor, handles=handles)
    results['_norm_factors'] = (norm_part, norm_mesh)
    _io.record(results, paramset,
               catalogue_data.write_attrs_as_header(
                   catalogue_ref=catalogue_rand),
               results['_norm_factors'], save, logger)
    return results


def _compute_3pt_stats_sim_like(backend, kind, catalogue_data, paramset=None,
                                params_sampling=None, degrees=None,
                                binning=None, form=None, idx_bin=None,
                                precision=None, logger=None,
                                multipoles=None, save=False):
    stype = 'bispec' if kind == 'bispec' else '3pcf'
    paramset = _amalgamate_parameters(
        paramset=paramset, params_sampling=params_sampling, degrees=degrees,
        binning=binning, form=form, idx_bin=idx_bin, precision=precision,
        catalogue_type='sim', statistic_type=stype)
    if binning is None:
        binning = Binning.from_parameter_set(paramset)

    if any(paramset['boxsize'][a] is None for a in 'xyz'):
        for a in 'xyz':
            lo, hi = catalogue_data.bounds[a]
            paramset['boxsize'][a] = (hi - lo) * paramset['expand']
        paramset._validate()
    # threept.py:554-556: meshes from the Nyquist cutoff when unset.
    if any(paramset['ngrid'][a] is None for a in 'xyz'):
        set_ngrid_from_cutoff(paramset)

    boxsize = [paramset['boxsize'][a] for a in 'xyz']
    catalogue_data.periodise(boxsize)
    if not catalogue_data['nz'].any():
        catalogue_data.compute_mean_density(boxsize=boxsize)

    data_v = _view(catalogue_data)
    params = _params_struct(paramset)
    norm_part = backend.norm_particles(data_v, 1.)
    with backend.resident(data_v) as hd:
        norm_mesh = backend.norm_mesh(data_v, params, 1., handle=hd)
        norm_factor = _select_norm(paramset, norm_part, norm_mesh)
        if multipoles is not None:
            plist = _multipole_structs(paramset, multipoles)
            return backend.multi(data_v, None, None, None, plist,
                                 _binning_view(binning),
                                 [norm_factor] * len(plist),
                                 handles=(hd, None) if hd is not None else None)
        results = backend.sim(kind, data_v, params, _binning_view(binning),
                              norm_factor, handle=hd)
    results['_norm_factors'] = (norm_part, norm_mesh)
    _io.record(results, paramset, catalogue_data.write_attrs_as_header(),
               results['_norm_factors'], save, logger)
    return results


def _finish(results):
    results.pop('_norm_factors', None)
    return results


# ---------------------------------------------------------------------------
# Public API
# ---------------------------------------------------------------------------

def compute_bispec(catalogue_data, catalogue_rand, los_data=None,
                   los_rand=None, degrees=None, binning=None, form=None,
                   idx_bin=None, sampling_params=None, paramset=None,
                   save=False, logger=None, precision=None):
    """Bispectrum from survey-like data and random catalogues in the local
    plane-parallel approximation (reference threept.py:743-865).

    `precision` ('fp64' default, or 'fp32') selects the field precision;
    bin sums stay FP64.  Returns a dict with 'k1_bin', 'k1_eff', 'nmodes_1',
    'k2_bin', 'k2_eff', 'nmodes_2', 'bk_raw', 'bk_shot'.
    """
    res = _compute_3pt_stats_survey_like(
        _engine_backend(), 'bispec', catalogue_data, catalogue_rand,
        los_data=los_data, los_rand=los_rand, paramset=paramset,
        params_sampling=sampling_params, degrees=degrees, binning=binning,
        form=form, idx_bin=idx_bin, precision=precision, logger=logger,
        save=save)
    return _finish(res)


def compute_3pcf(catalogue_data, catalogue_rand, los_data=None,
                 los_rand=None, degrees=None, binning=None, form=None,
                 idx_bin=None, sampling_params=None, paramset=None,
                 save=False, logger=None, precision=None):
    """Three-point correlation function from survey-like catalogues
    (reference threept.py:868-990)."""
    res = _compute_3pt_stats_survey_like(
        _engine_backend(), '3pcf', catalogue_data, catalogue_rand,
        los_data=los_data, los_rand=los_rand, paramset=paramset,
        params_sampling=sampling_params, degrees=degrees, binning=binning,
        form=form, idx_bin=idx_bin, precision=precision, logger=logger,
        save=save)
    return _finish(res)


def compute_bispec_in_gpp_box(catalogue_data, degrees=None, binning=None,
                              form=None, idx_bin=None, sampling_params=None,
                              paramset=None, save=False, logger=None,
                              precision=None):
    """Bispectrum from a periodic simulation-box catalogue in the global
    plane-parallel approximation (reference threept.py:1624-1723)."""
    res = _compute_3pt_stats_sim_like(
        _engine_backend(), 'bispec', catalogue_data, paramset=paramset,
        params_sampling=sampling_params, degrees=degrees, binning=binning,
        form=form, idx_bin=idx_bin, precision=precision, logger=logger,
        save=save)
    return _finish(res)


def compute_3pcf_in_gpp_box(catalogue_data, degrees=None, binning=None,
                            form=None, idx_bin=None, sampling_params=None,
                            paramset=None, save=False, logger=None,
                            precision=None):
    """Three-point correlation function from a periodic simulation-box
    catalogue (reference threept.py:1726-1823)."""
    res = _compute_3pt_stats_sim_like(
        _engine_backend(), '3pcf', catalogue_data, paramset=paramset,
        params_sampling=sampling_params, degrees=degrees, binning=binning,
        form=form, idx_bin=idx_bin, precision=precision, logger=logger,
        save=save)
    return _finish(res)


def _compute_3pcf_window(backend, catalogue_rand, los_rand=None, paramset=None,
                         params_sampling=None, degrees=None, wa_orders=None,
                         binning=None, form=None, idx_bin=None, precision=None,
                         logger=None, save=False):
    """Driver of the 3PCF window (reference threept.py:1832-2105): the random
    catalogue centred and periodised in the box, both normalisations with
    alpha = 1, the wide-angle kernel when `wa_orders` is given."""
    paramset = _amalgamate_parameters(
        paramset=paramset, params_sampling=params_sampling, degrees=degrees,
        binning=binning, form=form, idx_bin=idx_bin, precision=precision,
        catalogue_type='random', statistic_type='3pcf-win')
    wide_angle = wa_orders is not None
    if wide_angle:
        i_wa, j_wa = (int(v) for v in wa_orders)
        paramset['wa_orders'] = {'i': i_wa, 'j': j_wa}
    else:
        i_wa = j_wa = 0
    if binning is None:
        binning = Binning.from_parameter_set(paramset)
    if los_rand is None:
        los_rand = catalogue_rand.compute_los()
    los_rand = np.ascontiguousarray(los_rand, dtype=np.float64)
    if any(paramset['ngrid'][a] is None for a in 'xyz'):
        set_ngrid_from_cutoff(paramset)
    boxsize = [paramset['boxsize'][a] for a in 'xyz']
    catalogue_rand.centre(boxsize)
    catalogue_rand.periodise(boxsize)
    rand_v = _view(catalogue_rand)
    params = _params_struct(paramset)
    norm_part = backend.norm_particles(rand_v, 1.)
    norm_mesh = backend.norm_mesh(rand_v, params, 1.)
    norm_factor = _select_norm(paramset, norm_part, norm_mesh)
    bv = _binning_view(binning)
    res = backend._result(params, bv, '3pcf')
    backend._check(backend._fn('compute_3pcf_window')(
        C.byref(rand_v.struct), _abi.dptr(los_rand), C.byref(params),
        C.byref(bv.struct), C.c_double(1.), C.c_double(norm_factor),
        int(wide_angle), i_wa, j_wa, C.byref(res.struct)))
    results = res.as_dict()
    results['_norm_factors'] = (norm_part, norm_mesh)
    _io.record(results, paramset, catalogue_rand.write_attrs_as_header(),
               results['_norm_factors'], save, logger)
    return results


def compute_3pcf_window(catalogue_rand, los_rand=None, degrees=None,
                        wa_orders=None, binning=None, form=None, idx_bin=None,
                        sampling_params=None, paramset=None, save=False,
                        logger=None, precision=None):
    """Three-point correlation function window from a random catalogue, with
    wide-angle corrections of orders `wa_orders` = (i, j) when given
    (reference threept.py:1832-2105)."""
    res = _compute_3pcf_window(
        _engine_backend(), catalogue_rand, los_rand=los_rand,
        paramset=paramset, params_sampling=sampling_params, degrees=degrees,
        wa_orders=wa_orders, binning=binning, form=form, idx_bin=idx_bin,
        precision=precision, logger=logger, save=save)
    return _finish(res)


def compute_bispec_multipoles(catalogue_data, catalogue_rand=None,
                              multipoles=(((0, 0, 0), 'full'),
                                          ((2, 0, 2), 'full')),
                              binning=None, los_data=None, los_rand=None,
                              sampling_params=None, paramset=None,
                              boxsize=None, precision=None, logger=None):
    """Several bispectrum multipoles in one engine call (no reference
    counterpart; each entry equals the corresponding compute_bispec /
    compute_bispec_in_gpp_box result).

    `multipoles` lists (degrees, form) or (degrees, form, idx_bin).  With
    `catalogue_rand` the survey-like (local plane-parallel) estimator is used,
    otherwise the periodic-box one.  Returns a list of result dicts.
    """
    first = multipoles[0]
    kw = dict(paramset=paramset, params_sampling=sampling_params,
              degrees=first[0], binning=binning, form=first[1],
              idx_bin=first[2] if len(first) > 2 else None,
              precision=precision, logger=logger, multipoles=multipoles)
    if catalogue_rand is not None:
        return _compute_3pt_stats_survey_like(
            _engine_backend(), 'bispec', catalogue_data, catalogue_rand,
            los_data=los_data, los_rand=los_rand, **kw)
    return _compute_3pt_stats_sim_like(_engine_backend(), 'bispec',
                                       catalogue_data, **kw)
