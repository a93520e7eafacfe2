"""Two-point path of the engine (triumvirate_amd.twopt over trvg_compute_
powspec / corrfunc / *_in_gpp_box / corrfunc_window):

* the reference's own fixtures (pk0/pk2, xi0/xi2 for survey pairs and
  boxes, the windows xiw0/xiw2) and their normalisation headers, through
  the public drivers;
* the oracle on configurations no fixture covers: every assignment scheme
  with and without interlacing, degrees 0, 1 and 2, weighted survey
  catalogues, non-cubic meshes; counts bit-exact, values within 1e-10 of
  the largest |value| of the multipole.
"""

import numpy as np
import pytest

import conftest as cf
from test_oracle_twopt import (TWOPT_FIXTURES, check_twopt_fixture,
                               run_twopt_fixture)

pytestmark = pytest.mark.gpu


def _engine():
    from triumvirate_amd import _lib
    from triumvirate_amd.threept import CLibBackend
    return CLibBackend(_lib.load(), 'trvg_')


def _oracle():
    import oracle
    return oracle.backend(schedule=0)


@pytest.mark.parametrize('kind,geo,degree,name', TWOPT_FIXTURES)
def test_engine_twopt_matches_reference_fixture(kind, geo, degree, name):
    print(f"twopt fixture {name}", flush=True)
    res = run_twopt_fixture(_engine(), kind, geo, degree)
    check_twopt_fixture(res, kind, name)


def _survey(seed, weighted):
    from triumvirate_amd import mocks
    from triumvirate_amd.catalogue import ParticleCatalogue
    xd, yd, zd, nbar = mocks.lognormal_survey(8000, ngrid=32, seed=seed)
    xr, yr, zr = mocks.uniform_shell_randoms(40000, seed=seed + 1)
    g = np.random.default_rng(seed + 7)
    kw_d, kw_r = {}, {}
    if weighted:
        kw_d = dict(ws=g.uniform(0.5, 1.5, len(xd)), wc=g.uniform(0.8, 1.2, len(xd)))
        kw_r = dict(ws=g.uniform(0.5, 1.5, len(xr)), wc=g.uniform(0.8, 1.2, len(xr)))
    return (ParticleCatalogue(xd, yd, zd, nz=nbar, **kw_d),
            ParticleCatalogue(xr, yr, zr, nz=nbar, **kw_r))


def _paramset(ng, L, scheme, interlace, degree):
    from triumvirate_amd.parameters import ParameterSet
    return ParameterSet(param_dict={
        'boxsize': dict(zip('xyz', L)), 'ngrid': dict(zip('xyz', ng)),
        'assignment': scheme, 'interlace': interlace,
        'norm_convention': 'particle',
        'degrees': {'ell1': None, 'ell2': None, 'ELL': degree}})


def _compare(a, b, kind, tol=1e-10):
    cnt, eff, vals = (('nmodes', 'keff', ('pk_raw', 'pk_shot')) if kind == 'powspec'
                      else ('npairs', 'reff', ('xi',)))
    assert np.array_equal(a[cnt], b[cnt])
    np.testing.assert_allclose(a[eff], b[eff], rtol=1e-12)
    out = {}
    # Raw and shot parts share units; the shot noise of a box at l > 0 is
    # round-off (the shell average of y_l0 vanishes), so both are measured
    # against the larger of the two.
    scale = max(np.max(np.abs(b[key])) for key in vals)
    for key in vals:
        err = np.max(np.abs(a[key] - b[key])) / scale if scale > 0 else 0.
        out[key] = err
        assert err < tol, (key, err)
    return out


CASES = [
    # (kind, geometry, scheme, interlace, degree, ngrid, weighted)
    ('powspec', 'lpp', 'ngp', 'off', 0, (32, 32, 32), False),
    ('powspec', 'lpp', 'cic', 'on', 2, (32, 32, 32), True),
    ('powspec', 'lpp', 'tsc', 'on', 0, (32, 16, 64), False),
    ('powspec', 'lpp', 'pcs', 'on', 2, (32, 32, 32), False),
    ('powspec', 'lpp', 'pcs', 'off', 1, (32, 32, 32), True),
    ('2pcf', 'lpp', 'tsc', 'on', 2, (32, 32, 32), True),
    ('2pcf', 'lpp', 'cic', 'off', 0, (32, 32, 32), False),
    ('powspec', 'gpp', 'tsc', 'on', 2, (32, 32, 32), False),
    ('powspec', 'gpp', 'ngp', 'on', 0, (16, 32, 32), False),
    ('2pcf', 'gpp', 'pcs', 'on', 0, (32, 32, 32), False),
    ('2pcf', 'win', 'tsc', 'on', 2, (32, 32, 32), True),
    ('2pcf', 'win', 'cic', 'off', 0, (32, 32, 32), False),
]


@pytest.mark.parametrize('case', CASES, ids=lambda c: '-'.join(map(str, c[:5])))
def test_engine_twopt_matches_oracle(case):
    from triumvirate_amd import twopt as T
    from triumvirate_amd.dataobjs import Binning
    kind, geo, scheme, il, degree, ng, weighted = case
    print(f"twopt vs oracle {case}", flush=True)
    if kind == 'powspec':
        b = Binning('fourier', 'lin', bin_min=0.005, bin_max=0.065, num_bins=6)
    else:
        b = Binning('config', 'lin', bin_min=100., bin_max=700., num_bins=6)
    out = []
    for backend in (_engine(), _oracle()):
        d, r = _survey(3, weighted)
        if geo == 'gpp':
            from triumvirate_amd import mocks
            from triumvirate_amd.catalogue import ParticleCatalogue
            x, y, z = mocks.lognormal_box(20000, 1000., 32, seed=4)
            cat = ParticleCatalogue(x, y, z, nz=len(x) / 1e9)
            ps = _paramset(ng, (1000.,) * 3, scheme, il, degree)
            out.append(T._compute_2pt_stats_sim_like(backend, kind, cat, paramset=ps,
                                                     degree=degree, binning=b))
        elif geo == 'lpp':
            ps = _paramset(ng, (3200.,) * 3, scheme, il, degree)
            out.append(T._compute_2pt_stats_survey_like(backend, kind, d, r, paramset=ps,
                                                        degree=degree, binning=b))
        else:
            ps = _paramset(ng, (3200.,) * 3, scheme, il, degree)
            out.append(T._compute_corrfunc_window(backend, r, paramset=ps, degree=degree,
                                                  binning=b))
    errs = _compare(out[0], out[1], kind)
    print("  ", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    for i in range(3):
        np.testing.assert_allclose(out[0]['_norm_factors'][i],
                                   out[1]['_norm_factors'][i], rtol=1e-12)


def test_public_twopt_api():
    """The public functions run end to end with the reference's argument
    names and return the reference's keys."""
    from triumvirate_amd import twopt as T
    print("twopt public api", flush=True)
    ps = cf.make_test_paramset()
    r = T.compute_powspec(cf.make_test_data_catalogue(), cf.make_test_rand_catalogue(),
                          degree=0, binning=cf.binning_fourier(), paramset=ps)
    assert set(r) == {'kbin', 'keff', 'nmodes', 'pk_raw', 'pk_shot'}
    r = T.compute_corrfunc_in_gpp_box(cf.make_test_data_catalogue(), degree=2,
                                      binning=cf.binning_config(), paramset=ps)
    assert set(r) == {'rbin', 'reff', 'npairs', 'xi'}
    r = T.compute_corrfunc_window(cf.make_test_rand_catalogue(), degree=0,
                                  binning=cf.binning_config(), paramset=ps)
    assert set(r) == {'rbin', 'reff', 'npairs', 'xi'}
