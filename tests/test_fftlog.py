"""FFTLog Hankel and spherical Bessel transforms, extrapolation and
resampling (CPU).

Pinned by
* the reference's known-answer test of the biased transform
  (tests/test_fftlog.cpp:15-72: mu = 1/2, q = 0.1, eight samples at
  0.1 * 2^j, a_j = j + 1, low-ringing pivot from kr_c = 1;
  b[0] = 2.492884552, b[7] = 2.080611583 to 1e-9);
* the analytic transform pairs and tolerances of the reference's tests
  (tests/test_transforms.py:17-165, rtol 1e-3);
* the extrapolation fixtures of tests/test__arrayops.py:120-183;
* closed forms the reference does not test: the Hankel transform with
  q = 0 is its own inverse, the low-ringing pivot makes the kernel real
  at the Nyquist mode, and the transforms are linear over batches.
"""

import numpy as np
import pytest
from scipy.special import gamma, loggamma

from triumvirate_amd.fftlog import (HankelTransform, extrap_lin,
                                    extrap_loglin, extrap_pad)
from triumvirate_amd.transforms import (DoubleSphericalBesselTransform,
                                        SphericalBesselTransform,
                                        resample_lglin, resample_lin)


def test_biased_transform_known_answer():
    ht = HankelTransform(.5, .1, [.1, .2, .4, .8, 1.6, 3.2, 6.4, 12.8],
                         kr_c=1., lowring=True)
    _, b = ht.transform(np.arange(1., 9.).astype(complex))
    assert abs(b[0].real - 2.492884552) < 1e-9
    assert abs(b[7].real - 2.080611583) < 1e-9


HANKEL_PAIRS = {
    'sym': (lambda r, mu: r ** (mu + 1.) * np.exp(-r * r / 2.),
            lambda k, mu: k ** (mu + 1.) * np.exp(-k * k / 2.)),
    'asym': (lambda r, mu: r * (1. + r ** 2) ** -1.5,
             lambda k, mu: k * np.exp(-k)),
}


def _sj_pair(kind, ell, s=1):
    if kind == 'sym':
        return (lambda r: r ** ell * np.exp(-r * r / 2.),
                lambda k: k ** ell * np.exp(-k * k / 2.) * (2 * np.pi) ** 1.5)
    return (lambda r: r ** (s - 3.),
            lambda k: np.pi ** 1.5 * gamma(s / 2. + ell / 2.)
            / gamma(1.5 - s / 2. + ell / 2.) * (2. / k) ** s)


def _in_range(pts, lgrange):
    lg = np.log10(pts)
    return (lgrange[0] < lg) & (lg < lgrange[1])


@pytest.mark.parametrize('order, nsamp, extrap, pair, lgrange', [
    (2, 2 ** 10, 0, 'sym', (-3., .5)),
    (0, 2 ** 11, 4, 'asym', (-3., .5)),
])
def test_hankel_analytic_pairs(order, nsamp, extrap, pair, lgrange):
    f, g = HANKEL_PAIRS[pair]
    x = np.logspace(-5., 5., nsamp, base=10, endpoint=False)
    ht = HankelTransform(order, 0, x, kr_c=1., lowring=True, extrap=extrap)
    k, b = ht.transform(f(x, order))
    np.testing.assert_array_equal(k, ht._post_sampts)
    m = _in_range(k, lgrange)
    assert m.sum() > 300
    np.testing.assert_allclose(b[m].real, g(k[m], order), rtol=1e-3)


@pytest.mark.parametrize('degree, nsamp, extrap, pair, lgrange', [
    (0, 2 ** 10, 0, 'sym', (-3.5, .5)),
    (2, 2 ** 11, 4, 'asym', (-1., 3.)),
])
def test_spherical_bessel_analytic_pairs(degree, nsamp, extrap, pair, lgrange):
    f, g = _sj_pair(pair, degree)
    x = np.logspace(-5., 5., nsamp, base=10, endpoint=False)
    sbt = SphericalBesselTransform(degree, 0, x, lowring=True, extrap=extrap)
    k, b = sbt.transform(f(x))
    m = _in_range(k, lgrange)
    assert m.sum() > 300
    np.testing.assert_allclose(b[m].real, g(k[m]), rtol=1e-3)


@pytest.mark.parametrize('degrees, nsamp, extrap, pair, lgrange', [
    ((0, 0), 2 ** 10, 0, 'sym', (-3., 0.)),
    ((2, 2), 2 ** 11, 4, 'asym', (-.5, 3.)),
])
def test_double_spherical_bessel_analytic_pairs(degrees, nsamp, extrap, pair,
                                                lgrange):
    x = np.logspace(-5., 5., nsamp, base=10, endpoint=False)
    (f1, g1), (f2, g2) = (_sj_pair(pair, ell) for ell in degrees)
    dsbt = DoubleSphericalBesselTransform(degrees, (0, 0), x, lowring=True,
                                          extrap=extrap)
    (k1, k2), b = dsbt.transform(f1(x)[:, None] * f2(x)[None, :])
    m = _in_range(k1, lgrange) & _in_range(k2, lgrange)
    assert m.sum() > 1000
    np.testing.assert_allclose(b[m].real, (g1(k1) * g2(k2))[m], rtol=1e-3)


def test_double_transform_equals_one_dimensional_transforms():
    """The 2-d transform of a separable product is the outer product of the
    1-d transforms, with 2-d extrapolation by end values as well."""
    x = np.logspace(-3., 3., 256, endpoint=False)
    f1 = x ** 2 * np.exp(-x)
    f2 = np.exp(-x * x / 3.)
    for kw in ({}, {'extrap': 2, 'extrap2d': True}):
        dsbt = DoubleSphericalBesselTransform((1, 3), (0, 0), x, **kw)
        _, b = dsbt.transform(f1[:, None] * f2[None, :])
        n_ext = dsbt._n_ext
        ext = {'extrap': 2} if kw else {}
        g1 = SphericalBesselTransform(1, 0, x, **ext).transform(f1)[1]
        g2 = SphericalBesselTransform(3, 0, x, **ext).transform(f2)[1]
        scale = np.max(np.abs(g1)) * np.max(np.abs(g2))
        assert n_ext == (128 if kw else 0)
        np.testing.assert_allclose(b, g1[:, None] * g2[None, :],
                                   rtol=0, atol=1e-9 * scale)


def test_hankel_is_its_own_inverse():
    """With q = 0, int_0^inf J_mu(k r) J_mu(k r') k dk = delta(r - r') / r,
    so transforming twice (at the pivot 1, where the output points are the
    inputs reversed) returns the samples of a smooth, well-contained
    function."""
    x = np.logspace(-4., 4., 2048, endpoint=False)
    f = x ** 1.5 * np.exp(-(np.log(x) - .3) ** 2)
    ht = HankelTransform(1.5, 0., x, kr_c=1., lowring=False)
    k, b = ht.transform(f)
    ht2 = HankelTransform(1.5, 0., k, kr_c=1., lowring=False)
    r, c = ht2.transform(b)
    np.testing.assert_allclose(r, x, rtol=1e-12)
    np.testing.assert_allclose(c.real, f, atol=1e-9 * np.max(f))


@pytest.mark.parametrize('mu, q, nsamp', [(.5, .1, 8), (2.5, 0., 1024),
                                          (0., .7, 256)])
def test_lowring_pivot_makes_nyquist_kernel_real(mu, q, nsamp):
    """The low-ringing pivot puts the phase of
    2^q Gamma(x_+ + i pi / (2 dL)) / Gamma(x_- - i pi / (2 dL)) (k_c r_c / 2)^(-i pi / dL)
    on a multiple of pi, and moves the given pivot by less than one sample
    spacing in ln."""
    x = np.logspace(-2., 2., nsamp, endpoint=False)
    dL = np.log(x[1] / x[0])
    ht = HankelTransform(mu, q, x, kr_c=1., lowring=True)
    y = np.pi / (2. * dL)
    phase = (loggamma((mu + 1. + q) / 2. + 1j * y).imag
             + loggamma((mu + 1. - q) / 2. + 1j * y).imag
             - 2. * y * np.log(ht.pivot / 2.))
    frac = phase / np.pi - np.round(phase / np.pi)
    assert abs(frac) < 1e-9
    assert abs(np.log(ht.pivot)) <= dL / 2. + 1e-12


def test_transforms_are_batched_and_linear():
    x = np.logspace(-3., 3., 512, endpoint=False)
    g = np.random.default_rng(3)
    a = g.normal(size=(3, 512))
    sbt = SphericalBesselTransform(2, 0, x)
    _, b = sbt.transform(a)
    for i in range(3):
        np.testing.assert_array_equal(b[i], sbt.transform(a[i])[1])
    _, lin = sbt.transform(2. * a[0] - a[1])
    np.testing.assert_allclose(lin, 2. * b[0] - b[1],
                               atol=1e-12 * np.max(np.abs(b)))


def test_cosmo_multipole_factors():
    x = np.logspace(-3., 3., 256, endpoint=False)
    f = np.exp(-x)
    sbt = SphericalBesselTransform(2, 0, x)
    _, b = sbt.transform(f)
    np.testing.assert_allclose(sbt.transform_cosmo_multipoles(1, f)[1], -b)
    np.testing.assert_allclose(sbt.transform_cosmo_multipoles('backward', f)[1],
                               -b / (2 * np.pi) ** 3)
    with pytest.raises(ValueError):
        sbt.transform_cosmo_multipoles(0, f)


def test_invalid_inputs():
    x = np.logspace(-3., 3., 257, endpoint=False)
    with pytest.raises(ValueError):
        HankelTransform(0., 0., x, kr_c=1., extrap=1)
    with pytest.raises(ValueError):
        HankelTransform(0., 0., x, kr_c=-1., lowring=False)
    with pytest.raises(ValueError):
        HankelTransform(0., 0., np.linspace(1., 2., 16), kr_c=1.)
    with pytest.raises(ValueError):
        SphericalBesselTransform(0, 0, x[:-1], extrap=6)
    with pytest.raises(ValueError):
        SphericalBesselTransform(0, 0, x[:-1], extrap=1, extrap_layer='inner')


@pytest.mark.parametrize('a, n_ext, lo, hi, expected', [
    ([1, 2, 3], 2, -1, 1, [-1, -1, 1, 2, 3, 1, 1]),
    ([1, 2, 3], 0, 0, 0, [1, 2, 3]),
])
def test_extrap_pad(a, n_ext, lo, hi, expected):
    np.testing.assert_allclose(extrap_pad(np.array(a), n_ext, lo, hi), expected)


@pytest.mark.parametrize('a, n_ext, expected', [
    ([1, 2, 2], 2, [-1, 0, 1, 2, 2, 2, 2]),
    ([1, 2, 3], 0, [1, 2, 3]),
])
def test_extrap_lin(a, n_ext, expected):
    np.testing.assert_allclose(extrap_lin(np.array(a), n_ext), expected)


@pytest.mark.parametrize('a, n_ext, expected', [
    ([1., 2., 3.], 1, [.5, 1., 2., 3., 4.5]),
    ([-1., -2., -3.], 1, [-.5, -1., -2., -3., -4.5]),
    ([0, 1, 2], 2, [0, 0, 0, 1, 2, 4, 8]),
])
def test_extrap_loglin(a, n_ext, expected):
    np.testing.assert_allclose(extrap_loglin(np.array(a), n_ext), expected)


@pytest.mark.parametrize('a', [[-1., 2., 3.], [1., 0., 3.], [1., 2., -3.]])
def test_extrap_loglin_rejects_mixed_end_signs(a):
    with pytest.raises(ValueError):
        extrap_loglin(np.array(a), 1)


@pytest.mark.parametrize('sampts, size, func, ndim', [
    (np.arange(1., 11.), None, lambda x: x ** 2, 1),
    (np.arange(1., 51.) * 1e-3, 10, lambda x, y: np.exp(-x * np.sqrt(y)), 2),
])
def test_resample_lglin(sampts, size, func, ndim):
    samples = func(sampts) if ndim == 1 else \
        func(*np.meshgrid(sampts, sampts, indexing='ij'))
    pts, res = resample_lglin(sampts, samples, size=size)
    exact = func(pts) if ndim == 1 else \
        func(*np.meshgrid(pts[0], pts[-1], indexing='ij'))
    np.testing.assert_allclose(res, exact, rtol=1e-5, atol=1e-8)
    lg = np.log(pts if ndim == 1 else pts[0])
    np.testing.assert_allclose(np.diff(lg), np.diff(lg)[0], rtol=1e-12)


@pytest.mark.parametrize('sampts, size, func, ndim', [
    (np.arange(1., 11.), None, lambda x: x ** 2, 1),
    (np.arange(1., 51.) * 1e-3, 10, lambda x, y: x + y, 2),
])
def test_resample_lin(sampts, size, func, ndim):
    samples = func(sampts) if ndim == 1 else \
        func(*np.meshgrid(sampts, sampts, indexing='ij'))
    pts, res = resample_lin(sampts, samples, size=size)
    exact = func(pts) if ndim == 1 else \
        func(*np.meshgrid(pts[0], pts[-1], indexing='ij'))
    np.testing.assert_allclose(res, exact, rtol=1e-5, atol=1e-8)
