"""Meshes whose sides are not powers of two (even, up to 2048): the
reference plans FFTW for any size (field.cpp:246-339) and its drivers derive
even meshes from the Nyquist cutoff (parameters.cpp:1434-1463).

* 3-D transforms against numpy.fft (FFTW convention): 1e-12 (FP64),
  2e-6 (FP32) of max |X|, for radices 2, 3, 5, 13, 53 and a 2048-point
  axis (one line per workgroup).
* Estimators against the oracle (whose transform takes a direct DFT for
  such sizes) at 1e-10: box and survey bispectra, 3PCFs, power spectrum
  with interlacing, correlation function and window; FP32 within 1e-5 of
  FP64.
* The drivers with ngrid unset take the mesh from cutoff_nyq, and the
  result equals an explicit call on that mesh (to 1e-12: atomic sums).
"""

import ctypes as C

import numpy as np
import pytest

import conftest as cf
import parity as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def eng():
    from triumvirate_amd import _lib
    return _lib.load()


@pytest.fixture(scope='module')
def gpu_backend(eng):
    from triumvirate_amd.threept import CLibBackend
    return CLibBackend(eng)


@pytest.fixture(scope='module')
def oracle_backend():
    import oracle
    return oracle.backend(schedule=1)


@pytest.mark.parametrize('shape', [(12, 18, 30), (96, 8, 10), (8, 106, 8),
                                   (24, 40, 26), (10, 8, 2048), (2048, 8, 10)])
@pytest.mark.parametrize('prec', [64, 32])
def test_fft3d_general_sizes(eng, shape, prec):
    from triumvirate_amd import _abi
    print(f"fft3d general shape={shape} prec={prec}", flush=True)
    g = np.random.default_rng(11)
    a = g.normal(size=shape) + 1j * g.normal(size=shape)
    p = _abi.Params()
    for i in range(3):
        p.boxsize[i] = 1.
        p.ngrid[i] = shape[i]
    p.assignment_order = 3
    p.precision = prec
    p.device = -1
    tol = 1e-12 if prec == 64 else 2e-6
    for sign, ref in ((-1, np.fft.fftn(a)), (+1, np.fft.ifftn(a) * a.size)):
        buf = np.ascontiguousarray(a).view(np.float64).copy()
        assert eng.trvg_fft3d(C.byref(p), _abi.dptr(buf), sign) == 0, \
            eng.trvg_last_error()
        out = buf.view(np.complex128).reshape(shape)
        err = np.max(np.abs(out - ref)) / np.max(np.abs(ref))
        assert err < tol, (sign, err)


CASES = [
    # (kind, degrees, form, geometry, scheme, ngrid)
    ('bispec', (0, 0, 0), 'full', 'gpp', 'tsc', (24, 24, 24)),
    ('bispec', (2, 0, 2), 'diag', 'lpp', 'pcs', (30, 40, 48)),
    ('3pcf', (1, 1, 0), 'full', 'lpp', 'cic', (30, 30, 30)),
    ('3pcf', (0, 0, 0), 'diag', 'gpp', 'tsc', (26, 26, 26)),
]


@pytest.mark.parametrize('case', CASES, ids=lambda c: '-'.join(map(str, c[:5])))
def test_three_point_general_meshes_match_oracle(gpu_backend, oracle_backend, case):
    from triumvirate_amd import threept as T
    from triumvirate_amd.dataobjs import Binning
    kind, deg, form, geo, scheme, ng = case
    print(f"general-mesh 3pt {case}", flush=True)
    ps = cf.make_test_paramset(assignment=scheme, ngrid=dict(zip('xyz', ng)),
                               boxsize={'x': 800., 'y': 800., 'z': 800.})
    if kind == 'bispec':
        b = Binning('fourier', 'lin', bin_min=0.01, bin_max=0.09, num_bins=4)
    else:
        b = Binning('config', 'lin', bin_min=40., bin_max=160., num_bins=4)
    out = []
    for be in (gpu_backend, oracle_backend):
        if geo == 'gpp':
            d = P.uniform_box_catalogue(3000, 800., seed=5, weights=False)
            out.append(T._compute_3pt_stats_sim_like(
                be, kind, d, paramset=ps, degrees=deg, binning=b, form=form))
        else:
            d, r = P.survey_pair(2000, 8000, 700., seed=9, weights=True)
            out.append(T._compute_3pt_stats_survey_like(
                be, kind, d, r, paramset=ps, degrees=deg, binning=b, form=form))
    P.compare_results(out[0], out[1], kind, rtol=1e-10)


TWOPT = [
    # (kind, geometry, scheme, interlace, degree, ngrid)
    ('powspec', 'lpp', 'tsc', 'on', 2, (48, 48, 48)),
    ('powspec', 'gpp', 'pcs', 'on', 0, (30, 36, 40)),
    ('2pcf', 'gpp', 'cic', 'off', 0, (40, 40, 40)),
    ('2pcf', 'win', 'tsc', 'on', 2, (36, 36, 36)),
]


@pytest.mark.parametrize('case', TWOPT, ids=lambda c: '-'.join(map(str, c[:5])))
def test_two_point_general_meshes_match_oracle(case):
    import oracle
    from test_gpu_twopt import _compare, _paramset, _survey
    from triumvirate_amd import _lib, mocks
    from triumvirate_amd import twopt as T
    from triumvirate_amd.catalogue import ParticleCatalogue
    from triumvirate_amd.dataobjs import Binning
    from triumvirate_amd.threept import CLibBackend
    kind, geo, scheme, interlace, degree, ng = case
    print(f"general-mesh 2pt {case}", flush=True)
    b = (Binning('fourier', 'lin', bin_min=0.005, bin_max=0.065, num_bins=6)
         if kind == 'powspec' else
         Binning('config', 'lin', bin_min=100., bin_max=700., num_bins=6))
    out = []
    for be in (CLibBackend(_lib.load()), oracle.backend(schedule=0)):
        if geo == 'gpp':
            x, y, z = mocks.lognormal_box(20000, 1000., 32, seed=4)
            cat = ParticleCatalogue(x, y, z, nz=len(x) / 1e9)
            ps = _paramset(ng, (1000.,) * 3, scheme, interlace, degree)
            out.append(T._compute_2pt_stats_sim_like(
                be, kind, cat, paramset=ps, degree=degree, binning=b))
            continue
        d, r = _survey(3, geo == 'win')
        ps = _paramset(ng, (3200.,) * 3, scheme, interlace, degree)
        if geo == 'lpp':
            out.append(T._compute_2pt_stats_survey_like(
                be, kind, d, r, paramset=ps, degree=degree, binning=b))
        else:
            out.append(T._compute_corrfunc_window(
                be, r, paramset=ps, degree=degree, binning=b))
    _compare(out[0], out[1], kind)


def test_fp32_general_mesh_bound(gpu_backend):
    from triumvirate_amd import threept as T
    from triumvirate_amd.dataobjs import Binning
    print("general-mesh fp32", flush=True)
    ps = cf.make_test_paramset(assignment='pcs', ngrid={'x': 48, 'y': 48, 'z': 48},
                               boxsize={'x': 800., 'y': 800., 'z': 800.})
    b = Binning('fourier', 'lin', bin_min=0.01, bin_max=0.13, num_bins=6)
    res = {}
    for prec in ('fp32', 'fp64'):
        d, r = P.survey_pair(4000, 16000, 700., seed=4, weights=False)
        res[prec] = T._compute_3pt_stats_survey_like(
            gpu_backend, 'bispec', d, r, paramset=ps, degrees=(2, 0, 2), binning=b,
            form='full', precision=prec)
    for key in ('bk_raw', 'bk_shot'):
        err = np.max(np.abs(res['fp32'][key] - res['fp64'][key])) / \
            np.max(np.abs(res['fp64'][key]))
        assert err < 1e-5, (key, err)


def test_mesh_from_cutoff_equals_explicit_mesh(gpu_backend):
    """The box bispectrum with ngrid unset and cutoff_nyq = 0.13 in a
    1000 Mpc/h box: ceil(1000 * 0.13 / pi) = 42 cells per side (even), and
    the result equals the call on that mesh."""
    from triumvirate_amd import threept as T
    from triumvirate_amd.dataobjs import Binning
    b = Binning('fourier', 'lin', bin_min=0.01, bin_max=0.09, num_bins=4)
    out = []
    for ngrid in (None, 'derived'):
        d = P.uniform_box_catalogue(3000, 1000., seed=5, weights=False)
        if ngrid is None:
            ps = cf.make_test_paramset(ngrid={'x': None, 'y': None, 'z': None},
                                       boxsize={'x': 1000., 'y': 1000., 'z': 1000.},
                                       cutoff_nyq=0.13)
        else:
            n = int(np.ceil(1000. * 0.13 / np.pi))
            n += n % 2
            assert n == 42
            ps = cf.make_test_paramset(ngrid={'x': n, 'y': n, 'z': n},
                                       boxsize={'x': 1000., 'y': 1000., 'z': 1000.})
        out.append(T._compute_3pt_stats_sim_like(
            gpu_backend, 'bispec', d, paramset=ps, degrees=(0, 0, 0), binning=b,
            form='diag'))
    assert np.array_equal(out[0]['nmodes_1'], out[1]['nmodes_1'])
    # Equal up to the order of the assignment's atomic sums.
    for key in ('bk_raw', 'bk_shot'):
        np.testing.assert_allclose(out[0][key], out[1][key], rtol=1e-12)
