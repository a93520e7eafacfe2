"""Two-point statistics: power spectrum, correlation function and the
correlation-function window.

Public surface mirrors ``triumvirate.twopt``:

* :func:`compute_powspec` (reference twopt.py:668),
* :func:`compute_corrfunc` (twopt.py:786),
* :func:`compute_powspec_in_gpp_box` (twopt.py:1179),
* :func:`compute_corrfunc_in_gpp_box` (twopt.py:1262),
* :func:`compute_corrfunc_window` (twopt.py:1348),

with the same arguments and result dictionaries ('kbin', 'keff', 'nmodes',
'pk_raw', 'pk_shot' for power spectra; 'rbin', 'reff', 'npairs', 'xi' for
correlation functions).  Catalogue preparation follows the reference
drivers ``_compute_2pt_stats_survey_like`` (twopt.py:322-665),
``_compute_2pt_stats_sim_like`` (twopt.py:883-1176) and
``compute_corrfunc_window`` (twopt.py:1348-1587): lines of sight before
alignment, centring or padding with the random catalogue as reference,
periodisation of boxes, ``alpha = wstotal_data / wstotal_rand``, and all
normalisations computed with the one selected by ``norm_convention``.
Interlacing follows the parameter set.  The mesh work runs in the HIP
engine through its C ABI.
"""

import ctypes as C
import math

import numpy as np

from . import _abi
from . import io as _io
from .catalogue import ParticleCatalogue
from .dataobjs import Binning
from .parameters import (InvalidParameterError, ParameterSet,
                         fetch_paramset_template, set_ngrid_from_cutoff)
from .threept import CLibBackend, _engine_backend, _finish, _view

# Mixed-mesh normalisation geometry (reference twopt.py:52-54).
PADDING = .10
CELLSIZE = 10.
ASSIGNMENT = 'cic'


# ---------------------------------------------------------------------------
# Backend calls
# ---------------------------------------------------------------------------

def _norm_call(backend, name, *args):
    out = C.c_double(0.)
    backend._check(backend._fn(name)(*args, C.byref(out)))
    return out.value


def _norms_part(backend, cat, alpha):
    return _norm_call(backend, 'calc_powspec_normalisation_from_particles',
                      C.byref(cat.struct), C.c_double(alpha))


def _norms_mesh(backend, cat, params, alpha):
    return _norm_call(backend, 'calc_powspec_normalisation_from_mesh',
                      C.byref(cat.struct), C.byref(params), C.c_double(alpha))


def _run(backend, name, kind, nbins, *args):
    buf = _abi.TwoPtBuffers(nbins, kind)
    backend._check(backend._fn(name)(*args, C.byref(buf.struct)))
    return buf.as_dict()


# ---------------------------------------------------------------------------
# Parameters
# ---------------------------------------------------------------------------

def _amalgamate_parameters(paramset=None, params_sampling=None, degree=None,
                           binning=None, **type_kwargs):
    """Full parameter set for a two-point measurement (reference
    twopt.py:59-155)."""
    if paramset is None:
        if params_sampling is None or degree is None or binning is None:
            raise ValueError(
                "`degree`, `binning` and `params_sampling` must be provided "
                "when `paramset` is None.")
        pdict = fetch_paramset_template('dict')
    else:
        pdict = {k: (dict(v) if isinstance(v, dict) else v)
                 for k, v in paramset.items()}
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
        for key in ('boxsize', 'ngrid'):
            if key in ps:
                v = ps[key]
                v = [v] * 3 if np.isscalar(v) else list(v)
                pdict[key] = dict(zip('xyz', v))
    if degree is not None:
        deg = dict(pdict.get('degrees') or {})
        deg.update({'ell1': None, 'ell2': None, 'ELL': int(degree)})
        pdict['degrees'] = deg
    if binning is not None:
        pdict['range'] = [binning.bin_min, binning.bin_max]
        pdict['num_bins'] = binning.num_bins
        pdict['binning'] = binning.scheme
    pdict.update(type_kwargs)
    return ParameterSet(param_dict=pdict)


def _params_struct(paramset):
    """ABI parameters of a two-point call (the degree is ELL)."""
    box, ngrid = paramset['boxsize'], paramset['ngrid']
    if any(box[a] is None for a in 'xyz') or any(ngrid[a] is None for a in 'xyz'):
        raise InvalidParameterError("`boxsize` and `ngrid` must be set.")
    ELL = (paramset['degrees'] or {}).get('ELL')
    if ELL is None:
        raise InvalidParameterError("`degrees['ELL']` must be set.")
    p = _abi.Params()
    for i, a in enumerate('xyz'):
        p.boxsize[i] = float(box[a])
        p.ngrid[i] = int(ngrid[a])
    p.assignment_order = int(paramset['assignment_order'])
    p.interlace = 1 if paramset['interlace'] is True else 0
    p.ell1, p.ell2, p.ELL = 0, 0, int(ELL)
    p.shape = 0
    p.idx_bin = 0
    p.precision = 64
    p.verbose = int(paramset['verbose'] or 20)
    p.device = -1
    return p


def _binning_view(binning):
    return _abi.BinningView(binning.bin_edges, binning.bin_centres)


def _select_norm(paramset, part, mesh, meshes):
    conv = paramset['norm_convention']
    if conv == 'none':
        return 1.
    if conv == 'particle':
        return part
    if conv == 'mesh':
        return mesh
    if conv == 'mesh-mixed':
        return meshes
    raise InvalidParameterError(
        f"Unknown normalisation convention: '{conv}'.")


def _mesh_mixed_norm(backend, catalogue_data, catalogue_rand, paramset, alpha):
    """Mixed-mesh normalisation (twopt.cpp:111-268 with the geometry of
    twopt.py:52-54): both catalogues re-aligned, as the measurement, in a
    box (1 + PADDING) times the larger span, of cells CELLSIZE wide (an even
    number of them), assigned with CIC; 1 / (alpha V_cell sum n_d n_r)."""
    span = max(max(hi - lo for lo, hi in c.bounds.values())
               for c in (catalogue_data, catalogue_rand))
    L = (1. + PADDING) * span
    n = int(math.ceil(L / CELLSIZE))
    n += n % 2
    L = n * CELLSIZE
    d = ParticleCatalogue(*catalogue_data.columns()[:3],
                          nz=catalogue_data['nz'], ws=catalogue_data['ws'],
                          wc=catalogue_data['wc'])
    r = ParticleCatalogue(*catalogue_rand.columns()[:3],
                          nz=catalogue_rand['nz'], ws=catalogue_rand['ws'],
                          wc=catalogue_rand['wc'])
    if paramset['alignment'] == 'pad':
        if paramset['padscale'] == 'grid':
            d.pad([L] * 3, ngrid=[n] * 3, ngrid_pad=paramset['padfactor'],
                  catalogue_ref=r)
        else:
            d.pad([L] * 3, boxsize_pad=paramset['padfactor'], catalogue_ref=r)
    else:
        d.centre([L] * 3, catalogue_ref=r)
    p = _params_struct(paramset)
    for i in range(3):
        p.boxsize[i] = L
        p.ngrid[i] = n
    p.assignment_order = 2
    dv, rv = _view(d), _view(r)
    return _norm_call(backend, 'calc_powspec_normalisation_from_meshes',
                      C.byref(dv.struct), C.byref(rv.struct), C.byref(p),
                      C.c_double(alpha))


def _prepare_box(paramset, catalogue):
    if any(paramset['boxsize'][a] is None for a in 'xyz'):
        for a in 'xyz':
            lo, hi = catalogue.bounds[a]
            paramset['boxsize'][a] = (hi - lo) * paramset['expand']
        paramset._validate()
    if any(paramset['ngrid'][a] is None for a in 'xyz'):
        set_ngrid_from_cutoff(paramset)


# ---------------------------------------------------------------------------
# Drivers
# ---------------------------------------------------------------------------

def _compute_2pt_stats_survey_like(backend, kind, catalogue_data,
                                   catalogue_rand, los_data=None,
                                   los_rand=None, paramset=None,
                                   params_sampling=None, degree=None,
                                   binning=None, logger=None, save=False):
    stype = 'powspec' if kind == 'powspec' else '2pcf'
    paramset = _amalgamate_parameters(
        paramset=paramset, params_sampling=params_sampling, degree=degree,
        binning=binning, catalogue_type='survey', statistic_type=stype)
    if binning is None:
        binning = Binning.from_parameter_set(paramset)
    if los_data is None:
        los_data = catalogue_data.compute_los()
    if los_rand is None:
        los_rand = catalogue_rand.compute_los()
    los_data = np.ascontiguousarray(los_data, dtype=np.float64)
    los_rand = np.ascontiguousarray(los_rand, dtype=np.float64)
    _prepare_box(paramset, catalogue_rand)
    boxsize = [paramset['boxsize'][a] for a in 'xyz']
    if paramset['alignment'] == 'centre':
        catalogue_data.centre(boxsize, catalogue_ref=catalogue_rand)
    else:
        if paramset['padscale'] == 'box':
            kw = {'boxsize_pad': paramset['padfactor']}
        else:
            kw = {'ngrid': [paramset['ngrid'][a] for a in 'xyz'],
                  'ngrid_pad': paramset['padfactor']}
        catalogue_data.pad(boxsize, catalogue_ref=catalogue_rand, **kw)

    data_v, rand_v = _view(catalogue_data), _view(catalogue_rand)
    params = _params_struct(paramset)
    alpha = catalogue_data.wstotal / catalogue_rand.wstotal
    norm_part = _norms_part(backend, rand_v, alpha)
    norm_mesh = _norms_mesh(backend, rand_v, params, alpha)
    norm_meshes = _mesh_mixed_norm(backend, catalogue_data, catalogue_rand,
                                   paramset, alpha)
    norm = _select_norm(paramset, norm_part, norm_mesh, norm_meshes)
    name = 'compute_powspec' if kind == 'powspec' else 'compute_corrfunc'
    bv = _binning_view(binning)
    res = _run(backend, name, kind, binning.num_bins, C.byref(data_v.struct),
               C.byref(rand_v.struct), _abi.dptr(los_data),
               _abi.dptr(los_rand), C.byref(params), C.byref(bv.struct),
               C.c_double(norm))
    res['_norm_factors'] = (norm_part, norm_mesh, norm_meshes)
    _io.record(res, paramset,
               catalogue_data.write_attrs_as_header(
                   catalogue_ref=catalogue_rand),
               res['_norm_factors'], save, logger)
    return res


def _compute_2pt_stats_sim_like(backend, kind, catalogue_data, paramset=None,
                                params_sampling=None, degree=None,
                                binning=None, logger=None, save=False):
    stype = 'powspec' if kind == 'powspec' else '2pcf'
    paramset = _amalgamate_parameters(
        paramset=paramset, params_sampling=params_sampling, degree=degree,
        binning=binning, catalogue_type='sim', statistic_type=stype)
    if binning is None:
        binning = Binning.from_parameter_set(paramset)
    _prepare_box(paramset, catalogue_data)
    boxsize = [paramset['boxsize'][a] for a in 'xyz']
    catalogue_data.periodise(boxsize)
    if not catalogue_data['nz'].any():
        catalogue_data.compute_mean_density(boxsize=boxsize)
    data_v = _view(catalogue_data)
    params = _params_struct(paramset)
    norm_part = _norms_part(backend, data_v, 1.)
    norm_mesh = _norms_mesh(backend, data_v, params, 1.)
    if paramset['norm_convention'] == 'mesh-mixed':
        raise InvalidParameterError(
            "Normalisation convention 'mesh-mixed' needs paired survey-type "
            "catalogues.")
    norm = _select_norm(paramset, norm_part, norm_mesh, 0.)
    name = ('compute_powspec_in_gpp_box' if kind == 'powspec'
            else 'compute_corrfunc_in_gpp_box')
    bv = _binning_view(binning)
    res = _run(backend, name, kind, binning.num_bins, C.byref(data_v.struct),
               C.byref(params), C.byref(bv.struct), C.c_double(norm))
    res['_norm_factors'] = (norm_part, norm_mesh, 0.)
    _io.record(res, paramset, catalogue_data.write_attrs_as_header(),
               res['_norm_factors'], save, logger)
    return res


def _compute_corrfunc_window(backend, catalogue_rand, los_rand=None,
                             paramset=None, params_sampling=None, degree=None,
                             binning=None, logger=None, save=False):
    paramset = _amalgamate_parameters(
        paramset=paramset, params_sampling=params_sampling, degree=degree,
        binning=binning, catalogue_type='random', statistic_type='2pcf-win')
    if binning is None:
        binning = Binning.from_parameter_set(paramset)
    if los_rand is None:
        los_rand = catalogue_rand.compute_los()
    los_rand = np.ascontiguousarray(los_rand, dtype=np.float64)
    _prepare_box(paramset, catalogue_rand)
    boxsize = [paramset['boxsize'][a] for a in 'xyz']
    catalogue_rand.centre(boxsize)
    catalogue_rand.periodise(boxsize)
    rand_v = _view(catalogue_rand)
    params = _params_struct(paramset)
    norm_part = _norms_part(backend, rand_v, 1.)
    norm_mesh = _norms_mesh(backend, rand_v, params, 1.)
    if paramset['norm_convention'] == 'mesh-mixed':
        raise InvalidParameterError(
            "Normalisation convention 'mesh-mixed' does not apply to the "
            "correlation-function window.")
    norm = _select_norm(paramset, norm_part, norm_mesh, 0.)
    bv = _binning_view(binning)
    res = _run(backend, 'compute_corrfunc_window', '2pcf', binning.num_bins,
               C.byref(rand_v.struct), _abi.dptr(los_rand), C.byref(params),
               C.byref(bv.struct), C.c_double(1.), C.c_double(norm))
    res['_norm_factors'] = (norm_part, norm_mesh, 0.)
    _io.record(res, paramset, catalogue_rand.write_attrs_as_header(),
               res['_norm_factors'], save, logger)
    return res


# ---------------------------------------------------------------------------
# Public API
# ---------------------------------------------------------------------------

def compute_powspec(catalogue_data, catalogue_rand, los_data=None,
                    los_rand=None, degree=None, binning=None,
                    sampling_params=None, paramset=None, save=False,
                    logger=None):
    """Power spectrum multipole from survey-like data and random catalogues
    in the local plane-parallel approximation (reference twopt.py:668-783).
    """
    res = _compute_2pt_stats_survey_like(
        _engine_backend(), 'powspec', catalogue_data, catalogue_rand,
        los_data=los_data, los_rand=los_rand, paramset=paramset,
        params_sampling=sampling_params, degree=degree, binning=binning,
        logger=logger, save=save)
    return _finish(res)


def compute_corrfunc(catalogue_data, catalogue_rand, los_data=None,
                     los_rand=None, degree=None, binning=None,
                     sampling_params=None, paramset=None, save=False,
                     logger=None):
    """Two-point correlation function multipole from survey-like catalogues
    (reference twopt.py:786-880)."""
    res = _compute_2pt_stats_survey_like(
        _engine_backend(), '2pcf', catalogue_data, catalogue_rand,
        los_data=los_data, los_rand=los_rand, paramset=paramset,
        params_sampling=sampling_params, degree=degree, binning=binning,
        logger=logger, save=save)
    return _finish(res)


def compute_powspec_in_gpp_box(catalogue_data, degree=None, binning=None,
                               sampling_params=None, paramset=None,
                               save=False, logger=None):
    """Power spectrum multipole from a periodic simulation-box catalogue in
    the global plane-parallel approximation (reference twopt.py:1179-1259).
    """
    res = _compute_2pt_stats_sim_like(
        _engine_backend(), 'powspec', catalogue_data, paramset=paramset,
        params_sampling=sampling_params, degree=degree, binning=binning,
        logger=logger, save=save)
    return _finish(res)


def compute_corrfunc_in_gpp_box(catalogue_data, degree=None, binning=None,
                                sampling_params=None, paramset=None,
                                save=False, logger=None):
    """Two-point correlation function multipole from a periodic
    simulation-box catalogue (reference twopt.py:1262-1345)."""
    res = _compute_2pt_stats_sim_like(
        _engine_backend(), '2pcf', catalogue_data, paramset=paramset,
        params_sampling=sampling_params, degree=degree, binning=binning,
        logger=logger, save=save)
    return _finish(res)


def compute_corrfunc_window(catalogue_rand, los_rand=None, degree=None,
                            binning=None, sampling_params=None, paramset=None,
                            save=False, logger=None):
    """Correlation-function window multipole from a random catalogue
    (reference twopt.py:1348-1587)."""
    res = _compute_corrfunc_window(
        _engine_backend(), catalogue_rand, los_rand=los_rand,
        paramset=paramset, params_sampling=sampling_params, degree=degree,
        binning=binning, logger=logger, save=save)
    return _finish(res)


__all__ = ['compute_powspec', 'compute_corrfunc', 'compute_powspec_in_gpp_box',
           'compute_corrfunc_in_gpp_box', 'compute_corrfunc_window', 'CLibBackend']
