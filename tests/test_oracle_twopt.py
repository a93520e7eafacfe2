"""Pins of the oracle's two-point path (src/twopt.cpp).

* Every reference fixture of the two-point path (pk0/pk2 and xi0/xi2 for
  survey pairs and periodic boxes, the correlation-function windows
  xiw0/xiw2): counts equal, bins and effective coordinates to 1e-9, real
  parts within 2e-9 of the largest |Re| of the column (the files print 10
  digits), imaginary parts round-off small; and the three normalisations
  in the fixture headers (particle, mesh, mixed-mesh).
* Interlacing (field.cpp:929-1112 shadow passes, 1609-1650 average), which
  no fixture exercises: the interlaced transform of one particle equals
  the alias series of the continuous kernel transform with every term of
  odd total alias order removed, 1/2 [prod_a S_a + prod_a T_a] with
  S_a = sum_n K(q_a + 2 pi n) e^{-i (q_a + 2 pi n) t_a} and T_a the same
  with (-1)^n (K = sinc(w / 2)^p, the transform pinned in
  test_oracle_pins.py).
"""

import ctypes as C
import re

import numpy as np
import pytest

import conftest as cf
from triumvirate_amd import _abi


def _oracle():
    import oracle
    return oracle.backend(schedule=0)


def _fixture(name):
    return cf.load_ref_stats(name)


def _header_norms(name):
    txt = (cf.REF_STATS / name).read_text()
    m = re.search(r"Normalisation factor alternatives: ([0-9.e+-]+) \(particle\), "
                  r"([0-9.e+-]+) \(mesh\), ([0-9.e+-]+) \(mesh-mixed\)", txt)
    return tuple(float(v) for v in m.groups())


def _check_powspec(res, ext):
    np.testing.assert_allclose(res['kbin'], ext[0], rtol=1e-9)
    np.testing.assert_allclose(res['keff'], ext[1], rtol=1e-9)
    assert np.array_equal(res['nmodes'], ext[2].astype(int))
    for col, key in ((3, 'pk_raw'),):
        scale = np.max(np.abs(ext[col]))
        assert np.max(np.abs(res[key].real - ext[col])) <= 2e-9 * scale
        assert np.max(np.abs(res[key].imag)) <= 1e-6 * scale
    # Shot noise: the reference test compares it with atol 1e-6 (its
    # imaginary and, for l > 0, real parts are round-off).
    sn = ext[5] + 1j * ext[6]
    scale = max(np.max(np.abs(sn)), 1.)
    assert np.max(np.abs(res['pk_shot'] - sn)) <= max(1e-6, 2e-9 * scale)


def _check_corrfunc(res, ext):
    np.testing.assert_allclose(res['rbin'], ext[0], rtol=1e-9)
    np.testing.assert_allclose(res['reff'], ext[1], rtol=1e-9)
    assert np.array_equal(res['npairs'], ext[2].astype(int))
    scale = np.max(np.abs(ext[3]))
    assert np.max(np.abs(res['xi'].real - ext[3])) <= 2e-9 * scale
    assert np.max(np.abs(res['xi'].imag)) <= 1e-9 * scale


def run_twopt_fixture(backend, kind, geo, degree):
    """One reference two-point fixture configuration through the drivers."""
    from triumvirate_amd import twopt as T
    ps = cf.make_test_paramset(degrees={'ell1': None, 'ell2': None,
                                        'ELL': degree})
    if kind == 'powspec':
        b = cf.binning_fourier()
    else:
        b = cf.binning_config()
    if geo == 'gpp':
        return T._compute_2pt_stats_sim_like(
            backend, kind, cf.make_test_data_catalogue(), paramset=ps,
            degree=degree, binning=b)
    if geo == 'lpp':
        return T._compute_2pt_stats_survey_like(
            backend, kind, cf.make_test_data_catalogue(),
            cf.make_test_rand_catalogue(), paramset=ps, degree=degree,
            binning=b)
    return T._compute_corrfunc_window(
        backend, cf.make_test_rand_catalogue(), paramset=ps, degree=degree,
        binning=b)


TWOPT_FIXTURES = [
    ('powspec', 'gpp', 0, 'pk0_gpp.txt'), ('powspec', 'gpp', 2, 'pk2_gpp.txt'),
    ('powspec', 'lpp', 0, 'pk0_lpp.txt'), ('powspec', 'lpp', 2, 'pk2_lpp.txt'),
    ('2pcf', 'gpp', 0, 'xi0_gpp.txt'), ('2pcf', 'gpp', 2, 'xi2_gpp.txt'),
    ('2pcf', 'lpp', 0, 'xi0_lpp.txt'), ('2pcf', 'lpp', 2, 'xi2_lpp.txt'),
    ('2pcf', 'win', 0, 'xiw0.txt'), ('2pcf', 'win', 2, 'xiw2.txt'),
]


def check_twopt_fixture(res, kind, name):
    ext = _fixture(name)
    if kind == 'powspec':
        _check_powspec(res, ext)
    else:
        _check_corrfunc(res, ext)
    part, mesh, mixed = _header_norms(name)
    got = res['_norm_factors']
    np.testing.assert_allclose(got[0], part, rtol=2e-9)
    np.testing.assert_allclose(got[1], mesh, rtol=2e-9)
    if mixed != 0.:
        np.testing.assert_allclose(got[2], mixed, rtol=2e-9)


@pytest.mark.parametrize('kind,geo,degree,name', TWOPT_FIXTURES)
def test_oracle_twopt_matches_reference_fixture(kind, geo, degree, name):
    res = run_twopt_fixture(_oracle(), kind, geo, degree)
    check_twopt_fixture(res, kind, name)


def _sinc_p(w, p):
    w = np.asarray(w, dtype=float)
    out = np.ones_like(w)
    nz = w != 0.
    out[nz] = (np.sin(w[nz] / 2.) / (w[nz] / 2.)) ** p
    return out


def _alias_pair(q, t, p, M):
    """S = sum_n K(q + 2 pi n) e^{-i (q + 2 pi n) t}, T = the same with
    (-1)^n, |n| <= M (terms decay as n^-p)."""
    n = np.arange(-M, M + 1, dtype=float)
    w = q + 2. * np.pi * n
    terms = _sinc_p(w, p) * np.exp(-1j * w * t)
    sign = np.where(n.astype(np.int64) % 2 == 0, 1., -1.)
    return terms.sum(), (sign * terms).sum()


@pytest.mark.parametrize('order,tol', [(2, 2e-5), (3, 1e-9), (4, 1e-11)])
def test_interlaced_transform_cancels_odd_aliases(oracle_lib, order, tol):
    N = 16
    p = _abi.Params()
    for i in range(3):
        p.ngrid[i] = N
        p.boxsize[i] = float(N)            # unit cells: V_cell = 1
    p.assignment_order = order
    p.precision = 64
    t = np.array([5.3, 7.71, 9.137])       # away from the shadow-pass edges
    one = _abi.CatalogueView(t[:1], t[1:2], t[2:], np.zeros(1), np.ones(1),
                             np.ones(1))
    wb = np.array([1., 0.])
    main = np.zeros(2 * N ** 3)
    shadow = np.zeros(2 * N ** 3)
    for out, sh in ((main, 0), (shadow, 1)):
        assert oracle_lib.trvo_assign(C.byref(one.struct), _abi.dptr(wb),
                                      C.byref(p), sh, _abi.dptr(out)) == 0
    p.interlace = 1
    assert oracle_lib.trvo_fourier_transform(C.byref(p), _abi.dptr(main),
                                             _abi.dptr(shadow)) == 0
    F = main.view(np.complex128).reshape(N, N, N)
    M = {2: 400000, 3: 100000, 4: 20000}[order]
    rng = np.random.default_rng(order)
    modes = rng.integers(0, N, size=(12, 3))
    for j in modes:
        js = np.where(j < N // 2, j, j - N)
        q = 2. * np.pi * js / N
        S, T = zip(*[_alias_pair(q[a], t[a], order, M) for a in range(3)])
        expect = 0.5 * (np.prod(S) + np.prod(T))
        got = F[tuple(j)]
        assert abs(got - expect) <= tol, (order, j, got, expect)
