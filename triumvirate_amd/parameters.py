"""Parameter set for three-point measurements.

Mirrors the dict-like ``triumvirate.parameters.ParameterSet`` for the entries
the three-point path reads, and restates the validation/derivation rules of
``ParameterSet::validate`` (src/triumvirate/src/parameters.cpp:466-1270):

* ``assignment`` -> ``assignment_order`` (parameters.cpp:659-670);
* ``interlace`` accepts True/False/'on'/'off'/'true'/'false' and is
  transmuted to a bool (parameters.cpp:672-690, parameters.pyx:576-585);
* ``statistic_type`` -> ``npoint``/``space`` (parameters.cpp:700-815);
* ``form`` 'full' with ``ell1 == ell2`` -> ``shape`` 'triu', otherwise
  ``shape = form`` (parameters.cpp:830-850);
* interlacing is switched off for every three-point statistic
  (parameters.cpp:1240-1249);
* ``volume``/``nmesh`` derived (parameters.cpp:1028-1031).

Defaults follow the reference template
(src/triumvirate/resources/params_template.yml).
"""

import copy
import logging
import math
from pathlib import Path
from pprint import pformat, pprint

import yaml

_ASSIGNMENT_ORDERS = {'ngp': 1, 'cic': 2, 'tsc': 3, 'pcs': 4}
_FORMS = ('diag', 'off-diag', 'row', 'full')
_NORMS = ('none', 'particle', 'mesh', 'mesh-mixed')

_TEMPLATE = {
    'directories': {'catalogues': None, 'measurements': None},
    'files': {'data_catalogue': None, 'rand_catalogue': None},
    'catalogue_columns': [],
    'catalogue_dataset': None,
    'tags': {'output': None},
    'boxsize': {'x': None, 'y': None, 'z': None},
    'ngrid': {'x': None, 'y': None, 'z': None},
    'expand': 1.,
    'cutoff_nyq': None,
    'alignment': 'centre',
    'padscale': 'box',
    'padfactor': None,
    'assignment': 'tsc',
    'interlace': 'off',
    'catalogue_type': None,
    'statistic_type': None,
    'degrees': {'ell1': None, 'ell2': None, 'ELL': None},
    'wa_orders': {'i': None, 'j': None},
    'form': 'diag',
    'norm_convention': 'particle',
    'binning': 'lin',
    'range': [None, None],
    'num_bins': None,
    'idx_bin': None,
    'fftw_scheme': 'measure',
    'use_fftw_wisdom': False,
    'save_binned_vectors': False,
    'verbose': 20,
    'progbar': 'off',
    # Engine-specific (no reference counterpart): field precision.
    'precision': 'fp64',
}


def fetch_paramset_template(format='dict', ret_defaults=False):
    """Parameter-set template as a dict (reference:
    ``triumvirate.parameters.fetch_paramset_template``)."""
    if format != 'dict':
        raise ValueError("Only the 'dict' template format is provided.")
    tmpl = copy.deepcopy(_TEMPLATE)
    if ret_defaults:
        defaults = {k: v for k, v in tmpl.items()
                    if v is not None and not isinstance(v, (dict, list))}
        return tmpl, defaults
    return tmpl


class InvalidParameterError(ValueError):
    """Maps ``trv::sys::InvalidParameterError``."""


class ParameterSet:
    """Dict-like parameter set with the reference's validation rules.

    Parameters
    ----------
    param_filepath : str or path, optional
        YAML parameter file (loaded with ``yaml.safe_load``).
    param_dict : dict, optional
        Parameter dictionary (used if `param_filepath` is None).
    logger : logging.Logger, optional
    """

    def __init__(self, param_filepath=None, param_dict=None, logger=None):
        self._logger = logger or logging.getLogger(__name__)
        params = copy.deepcopy(_TEMPLATE)
        if param_filepath is not None:
            with open(param_filepath) as f:
                loaded = yaml.safe_load(f) or {}
            params.update(loaded)
            self._source = f'extfile:{param_filepath}'
        elif param_dict is not None:
            params.update(copy.deepcopy(param_dict))
            self._source = 'extdict'
        else:
            self._source = 'template'
        self._original = True
        self._params = params
        self._validate()

    # dict interface ---------------------------------------------------------
    def __getitem__(self, key):
        return self._params[key]

    def __setitem__(self, key, value):
        self._params[key] = value
        self._original = False

    def __contains__(self, key):
        return key in self._params

    def get(self, key, default=None):
        return self._params.get(key, default)

    def items(self):
        return self._params.items()

    def keys(self):
        return self._params.keys()

    def names(self):
        return self._params.keys()

    def update(self, *args, **kwargs):
        self._params.update(*args, **kwargs)
        self._validate()
        self._original = False

    def __repr__(self):
        return f"ParameterSet({self._params!r})"

    def __str__(self):
        return "\n".join([
            f"ParameterSet(source={self._source}, original={self._original})",
            pformat(self._params, sort_dicts=False)])

    def print(self):
        """Print the parameters as a dictionary (parameters.pyx:265-270)."""
        pprint(self._params, sort_dicts=False)

    def save(self, filepath=None):
        """Write the validated parameters to a YAML file; by default
        ``parameters_used{output tag}.yml`` in the measurement directory
        (parameters.pyx:272-296)."""
        if filepath is None:
            filepath = Path(
                self._params['directories']['measurements'] or '',
                "parameters_used{}.yml".format(
                    self._params['tags']['output'] or ''))
        with open(filepath, 'w') as f:
            yaml.safe_dump(self._params, f, sort_keys=False,
                           default_flow_style=False)
        if self._logger:
            self._logger.info("Saved parameters to %s.", filepath)
        return filepath

    # validation -------------------------------------------------------------
    def _validate(self):
        p = self._params

        if p.get('alignment') not in ('centre', 'pad'):
            raise InvalidParameterError(
                "Box alignment must be 'centre' or 'pad': "
                f"`alignment` = '{p.get('alignment')}'.")
        if p.get('padscale') not in ('box', 'grid'):
            raise InvalidParameterError(
                "Pad scale must be 'box' or 'grid': "
                f"`padscale` = '{p.get('padscale')}'.")

        assignment = p.get('assignment')
        if assignment not in _ASSIGNMENT_ORDERS:
            raise InvalidParameterError(
                "Mesh assignment scheme must be 'ngp', 'cic', 'tsc' or "
                f"'pcs': `assignment` = '{assignment}'.")
        p['assignment_order'] = _ASSIGNMENT_ORDERS[assignment]

        interlace = p.get('interlace')
        if interlace in (True, 'true', 'on'):
            p['interlace'] = True
        elif interlace in (False, 'false', 'off', None):
            p['interlace'] = False
        else:
            raise InvalidParameterError(
                "Interlacing must be 'true'/'on' or 'false'/'off': "
                f"`interlace` = '{interlace}'.")

        # parameters.cpp:704-828: statistic type -> npoint and space, with
        # the catalogue type each statistic requires.
        stype = p.get('statistic_type')
        ctype = p.get('catalogue_type')
        needs = {'powspec': ('2pt', 'fourier', 'obs'),
                 '2pcf': ('2pt', 'config', 'obs'),
                 '2pcf-win': ('2pt', 'config', 'random'),
                 'bispec': ('3pt', 'fourier', 'obs'),
                 '3pcf': ('3pt', 'config', 'obs'),
                 '3pcf-win': ('3pt', 'config', 'random'),
                 '3pcf-win-wa': ('3pt', 'config', 'random')}
        if stype in needs:
            p['npoint'], p['space'], need = needs[stype]
            names = {'powspec': 'Power spectrum',
                     '2pcf': 'Two-point correlation function',
                     'bispec': 'Bispectrum',
                     '3pcf': 'Three-point correlation function'}
            if need == 'obs' and ctype in ('random', 'none'):
                raise InvalidParameterError(
                    f"{names[stype]} requires 'sim' or 'survey' "
                    f"catalogue(s): `catalogue_type` = '{ctype}'.")
            if need == 'random' and ctype is not None and ctype != 'random':
                what = ('Two-point' if stype == '2pcf-win' else 'Three-point')
                raise InvalidParameterError(
                    f"{what} correlation function window requires 'random' "
                    f"catalogue: `catalogue_type` = '{ctype}'.")
        elif stype == 'modes':
            p['npoint'], p['space'] = 'none', 'fourier'
        elif stype == 'pairs':
            p['npoint'], p['space'] = 'none', 'config'
        elif stype is None:
            p['npoint'], p['space'] = None, None
        else:
            raise InvalidParameterError(
                f"Unrecognised statistic type: `statistic_type` = '{stype}'.")

        form = p.get('form')
        if form not in _FORMS:
            raise InvalidParameterError(
                "Three-point statistic form is not recognised: "
                f"`form` = '{form}'.")
        deg = p.get('degrees') or {}
        ell1, ell2 = deg.get('ell1'), deg.get('ell2')
        if form == 'full' and ell1 is not None and ell1 == ell2:
            p['shape'] = 'triu'
        else:
            p['shape'] = form

        if p.get('norm_convention') not in _NORMS:
            raise InvalidParameterError(
                "Normalisation convention must be 'none', 'particle', "
                "'mesh' or 'mesh-mixed': "
                f"`norm_convention` = '{p.get('norm_convention')}'.")
        # parameters.cpp:870-882
        if p.get('norm_convention') == 'mesh-mixed' and \
                p.get('npoint') not in (None, '2pt'):
            raise InvalidParameterError(
                "Normalisation convention 'mesh-mixed' only applies to "
                f"two-point statistics: `npoint` = '{p.get('npoint')}'.")

        prec = str(p.get('precision', 'fp64')).lower()
        if prec not in ('fp64', 'fp32'):
            raise InvalidParameterError(
                f"Precision must be 'fp64' or 'fp32': `precision` = '{prec}'.")
        p['precision'] = prec

        box = p.get('boxsize') or {}
        ngrid = p.get('ngrid') or {}
        if all(box.get(a) is not None for a in 'xyz'):
            p['volume'] = float(box['x']) * float(box['y']) * float(box['z'])
        if all(ngrid.get(a) is not None for a in 'xyz'):
            for a in 'xyz':
                ngrid[a] = int(ngrid[a])
            p['nmesh'] = ngrid['x'] * ngrid['y'] * ngrid['z']

        rng = p.get('range') or [None, None]
        p['bin_min'], p['bin_max'] = rng[0], rng[1]

        idx_bin = p.get('idx_bin')
        nb = p.get('num_bins')
        if form in ('row', 'off-diag') and idx_bin is not None and nb:
            if abs(int(idx_bin)) >= int(nb):
                raise InvalidParameterError(
                    "Bin index `idx_bin` must be < `num_bins` in absolute "
                    "value.")

        if p.get('npoint') == '3pt' and p['interlace'] is True:
            p['interlace'] = False
            self._logger.info(
                "Interlacing is unsupported for three-point measurements. "
                "`interlace` is set to 'false'.")


def set_ngrid_from_cutoff(paramset):
    """Mesh numbers from the Nyquist cutoff (parameters.cpp:1434-1463):
    ceil(L k_cut / pi) in Fourier space, ceil(2 L / r_cut) in configuration
    space, rounded up to even."""
    cut = paramset['cutoff_nyq']
    if cut is None:
        raise InvalidParameterError(
            "`ngrid` is unset and `cutoff_nyq` is not given.")
    ngrid = {}
    for a in 'xyz':
        L = float(paramset['boxsize'][a])
        if paramset['space'] == 'fourier':
            n = int(math.ceil(L * cut / math.pi))
        elif paramset['space'] == 'config':
            n = int(math.ceil(2. * L / cut))
        else:
            raise InvalidParameterError(
                "Space must be 'fourier' or 'config': `space` = "
                f"'{paramset['space']}'.")
        ngrid[a] = n + n % 2
    paramset['ngrid'] = ngrid
    paramset._validate()
