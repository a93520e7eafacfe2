"""Reduced, seeded versions of the benchmark configurations, shared by the
golden-vector generator (tests/golden/make_golden.py) and the tests.

  cfg1_gpp_b000_diag : periodic box, uniform positions, TSC, 10 diag bins
  cfg1_full_gpp_b000_diag : the same at the stated size (64^3, 2e5
                       particles in L = 1000 Mpc/h, bins [0.01, 0.11])
  cfg2_lpp_b000_full : survey quadrant shell, lognormal data + randoms, PCS,
                       B_000 full (-> upper triangle)
  cfg2_lpp_b202_full : same inputs, B_202 full
  cfg3_gpp_b202_full : periodic lognormal mock, TSC, B_202 full
  cfg5_lpp_z110_full : survey inputs, TSC, zeta_110 full (-> upper triangle)
"""

import numpy as np

CASES = ['cfg1_gpp_b000_diag', 'cfg1_full_gpp_b000_diag',
         'cfg2_lpp_b000_full', 'cfg2_lpp_b202_full', 'cfg3_gpp_b202_full',
         'cfg5_lpp_z110_full']


def _paramset(ngrid, L, scheme):
    from triumvirate_amd.parameters import ParameterSet
    return ParameterSet(param_dict={
        'boxsize': {'x': L, 'y': L, 'z': L},
        'ngrid': {'x': ngrid, 'y': ngrid, 'z': ngrid},
        'assignment': scheme, 'norm_convention': 'particle',
        'degrees': {'ell1': None, 'ell2': None, 'ELL': 0},
    })


def _survey_inputs(nd, nr, seed):
    from triumvirate_amd import mocks
    from triumvirate_amd.catalogue import ParticleCatalogue
    xd, yd, zd, nbar = mocks.lognormal_survey(nd, ngrid=32, seed=seed)
    xr, yr, zr = mocks.uniform_shell_randoms(nr, seed=seed + 1)
    return (ParticleCatalogue(xd, yd, zd, nz=nbar),
            ParticleCatalogue(xr, yr, zr, nz=nbar))


def run_case(backend, name, precision=None):
    from triumvirate_amd import mocks
    from triumvirate_amd import threept as T
    from triumvirate_amd.catalogue import ParticleCatalogue
    from triumvirate_amd.dataobjs import Binning
    if name == 'cfg1_gpp_b000_diag':
        g = np.random.default_rng(1)
        x, y, z = g.uniform(0., 1000., (3, 20000))
        d = ParticleCatalogue(x, y, z, nz=20000 / 1e9)
        b = Binning('fourier', 'lin', bin_min=0.01, bin_max=0.11, num_bins=10)
        return T._compute_3pt_stats_sim_like(
            backend, 'bispec', d, paramset=_paramset(32, 1000., 'tsc'),
            degrees=(0, 0, 0), binning=b, form='diag', precision=precision)
    if name == 'cfg1_full_gpp_b000_diag':
        g = np.random.default_rng(11)
        x, y, z = g.uniform(0., 1000., (3, 200000))
        d = ParticleCatalogue(x, y, z, nz=200000 / 1e9)
        b = Binning('fourier', 'lin', bin_min=0.01, bin_max=0.11, num_bins=10)
        return T._compute_3pt_stats_sim_like(
            backend, 'bispec', d, paramset=_paramset(64, 1000., 'tsc'),
            degrees=(0, 0, 0), binning=b, form='diag', precision=precision)
    if name in ('cfg2_lpp_b000_full', 'cfg2_lpp_b202_full'):
        d, r = _survey_inputs(20000, 100000, seed=2)
        b = Binning('fourier', 'lin', bin_min=0.005, bin_max=0.065, num_bins=6)
        deg = (0, 0, 0) if name.endswith('b000_full') else (2, 0, 2)
        return T._compute_3pt_stats_survey_like(
            backend, 'bispec', d, r, paramset=_paramset(32, 3200., 'pcs'),
            degrees=deg, binning=b, form='full', precision=precision)
    if name == 'cfg3_gpp_b202_full':
        x, y, z = mocks.lognormal_box(30000, 1000., 32, seed=0)
        d = ParticleCatalogue(x, y, z, nz=len(x) / 1e9)
        b = Binning('fourier', 'lin', bin_min=0.01, bin_max=0.16, num_bins=5)
        return T._compute_3pt_stats_sim_like(
            backend, 'bispec', d, paramset=_paramset(32, 1000., 'tsc'),
            degrees=(2, 0, 2), binning=b, form='full', precision=precision)
    if name == 'cfg5_lpp_z110_full':
        d, r = _survey_inputs(20000, 100000, seed=5)
        b = Binning('config', 'lin', bin_min=100., bin_max=700., num_bins=5)
        return T._compute_3pt_stats_survey_like(
            backend, '3pcf', d, r, paramset=_paramset(32, 3200., 'tsc'),
            degrees=(1, 1, 0), binning=b, form='full', precision=precision)
    raise KeyError(name)


def kind_of(name):
    return '3pcf' if '_z' in name else 'bispec'


def load_golden(path, name):
    data = np.load(path)
    pre = name + '/'
    return {k[len(pre):]: data[k] for k in data.files if k.startswith(pre)}
