"""Shared fixtures.

The reference's own end-to-end fixtures (``tests/golden/reference_stats``)
were produced with a 3-particle equilateral-triangle data catalogue and
30 000 uniform randoms drawn with ``numpy.random.default_rng(seed=42)``
(reference tests/conftest.py:149-219), on a 64^3 TSC mesh in a 1000 Mpc/h
box with centre alignment and particle normalisation
(reference tests/test_input/params/test_params.yml).
"""

import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / 'tests' / 'golden'
REF_STATS = GOLDEN / 'reference_stats'


def pytest_configure(config):
    config.addinivalue_line(
        'markers', 'gpu: test needs an MI355X (HIP device) to run')


@pytest.fixture(scope='session')
def ref_stats_dir():
    return REF_STATS


def make_test_data_catalogue():
    from triumvirate_amd.catalogue import ParticleCatalogue
    d = np.loadtxt(REF_STATS / 'test_data_catalogue.txt')
    return ParticleCatalogue(d[:, 0], d[:, 1], d[:, 2], nz=d[:, 3])


def make_test_rand_catalogue():
    from triumvirate_amd.catalogue import ParticleCatalogue
    gen = np.random.default_rng(seed=42)
    x, y, z = gen.uniform(-500., 500., size=(3, 30000))
    return ParticleCatalogue(x, y, z, nz=3. / 1.e9)


def make_test_paramset(**overrides):
    from triumvirate_amd.parameters import ParameterSet
    p = {
        'boxsize': {'x': 1000., 'y': 1000., 'z': 1000.},
        'ngrid': {'x': 64, 'y': 64, 'z': 64},
        'alignment': 'centre', 'padscale': 'box', 'assignment': 'tsc',
        'interlace': 'off', 'form': 'diag', 'norm_convention': 'particle',
        'binning': 'lin', 'range': [0.005, 0.105], 'num_bins': 4,
        'degrees': {'ell1': None, 'ell2': None, 'ELL': 0},
        'verbose': 20,
    }
    p.update(overrides)
    return ParameterSet(param_dict=p)


def binning_fourier():
    from triumvirate_amd.dataobjs import Binning
    return Binning('fourier', 'lin', bin_min=0.005, bin_max=0.105, num_bins=4)


def binning_config():
    from triumvirate_amd.dataobjs import Binning
    return Binning('config', 'lin', bin_min=50., bin_max=150., num_bins=4)


def load_ref_stats(name):
    """Columns of a reference fixture file (unpacked)."""
    return np.loadtxt(REF_STATS / name, unpack=True)


@pytest.fixture(scope='session')
def oracle_lib():
    import oracle
    return oracle.load()
