"""Spherical Bessel transforms of clustering multipoles and resampling
(reference: transforms.py:35-704, over the FFTLog transform of
:mod:`triumvirate_amd.fftlog`).

Extrapolation options of the transforms: 0 none, 1 zeros, 2 end values,
3 linear, 4 log-linear, 5 log-linear with oscillation-aware tails (2-d
samples: rows first, then columns).
"""

from collections.abc import Sequence

import numpy as np
from scipy.interpolate import InterpolatedUnivariateSpline, RectBivariateSpline

from .fftlog import HankelTransform, _extrapolate, extrap_loglin

__all__ = ['SphericalBesselTransform', 'DoubleSphericalBesselTransform',
           'resample_lglin', 'resample_lin']


def _check_extrap(extrap):
    if extrap not in {0, 1, 2, 3, 4, 5}:
        raise ValueError(f"Extrapolation option must be in 0 to 5: {extrap=}.")


class SphericalBesselTransform:
    """Spherical Bessel transform g(k) = 4 pi int r^2 dr j_l(k r) f(r)
    (transforms.py:35-272): the Hankel transform of order l + 1/2 of
    r^{3/2} f, times (2 pi / k)^{3/2}.

    Parameters
    ----------
    degree : int
        Degree l.
    bias : int
        Power-law bias index.
    sample_pts : array of float
        Log-spaced pre-transform points (even in length with
        extrapolation).
    pivot : float, optional
        Pivot value (default 1), adjusted when `lowring`.
    lowring : bool, optional
        Low-ringing pivot (default True).
    extrap : int, optional
        Extrapolation option (default 0).
    extrap_exp : float, optional
        Sample-size expansion factor (default 2): the extended size is the
        smallest power of two >= extrap_exp times the sample size.
    extrap_layer : {'outer', 'native'}, optional
        'outer' (default) extends the samples before the r^{3/2} factor,
        'native' inside the Hankel transform after it.
    threaded : bool, optional
        Accepted for interface compatibility.
    """

    def __init__(self, degree, bias, sample_pts, pivot=1., lowring=True,
                 extrap=0, extrap_exp=2., extrap_layer='outer',
                 threaded=False):
        _check_extrap(extrap)
        self.degree = degree
        self.bias = bias
        self._lowring = lowring
        self._extrap = extrap
        self._extrap_layer = extrap_layer
        native = 0
        self._n_ext = 0
        if extrap:
            nsamp = len(sample_pts)
            if nsamp % 2:
                raise ValueError(
                    "Input sample size must be even when extrapolation is "
                    f"used: {nsamp=}.")
            if extrap_layer == 'native':
                native = extrap
            elif extrap_layer == 'outer':
                nsamp_trans = int(2 ** np.ceil(np.log2(extrap_exp * nsamp)))
                self._n_ext = (nsamp_trans - nsamp) // 2
                sample_pts = extrap_loglin(sample_pts, self._n_ext)
            else:
                raise ValueError(
                    "Extrapolation layer must be 'native' or 'outer': "
                    f"extrap_layer={extrap_layer}.")
        self._fbht = HankelTransform(degree + 0.5, bias, sample_pts, kr_c=pivot,
                                     lowring=lowring, extrap=native,
                                     extrap_exp=extrap_exp)
        self._post_slice = (slice(self._n_ext, -self._n_ext) if self._n_ext
                            else slice(None))
        self._logres = self._fbht._logres
        self._pre_sampts = np.asarray(self._fbht._pre_sampts)
        self._post_sampts = np.asarray(self._fbht._post_sampts)[self._post_slice]

    @property
    def size(self):
        """Sample size of the transform."""
        return self._fbht.size

    @property
    def pivot(self):
        """Pivot value used."""
        return self._fbht._pivot

    def _perform_extrap(self, arr):
        if self._extrap and self._extrap_layer == 'outer':
            return _extrapolate(arr, self._n_ext, self._extrap)
        return np.asarray(arr)

    def transform(self, pre_samples):
        """Transform samples at the initialised points; returns the
        post-transform points and samples.  Leading axes of the samples,
        if any, are independent sample sets transformed together."""
        a = self._perform_extrap(pre_samples) * self._pre_sampts ** 1.5
        post_sampts, b = self._fbht.transform(a)
        post_sampts = post_sampts[self._post_slice]
        b = b[..., self._post_slice] * (2 * np.pi / self._post_sampts) ** 1.5
        return post_sampts, b

    def transform_cosmo_multipoles(self, direction, pre_samples):
        """Clustering multipoles between configuration and Fourier space
        (transforms.py:211-246): forward (1, 'forward') multiplies by
        (-i)^l, backward (-1, 'backward') by i^l / (2 pi)^3."""
        if direction in {'forward', 1}:
            factor = (-1j) ** self.degree
        elif direction in {'backward', -1}:
            factor = (1j) ** self.degree / (2 * np.pi) ** 3
        else:
            raise ValueError(
                "Transform direction must be 'forward'/1 or 'backward'/-1.")
        post_sampts, post_samples = self.transform(pre_samples)
        return post_sampts, post_samples * factor


class DoubleSphericalBesselTransform:
    """Two-dimensional spherical Bessel transform (transforms.py:275-523):
    rows transformed with (degrees[1], biases[1]), then columns with
    (degrees[0], biases[0]), on one log-spaced grid for both dimensions.

    With `extrap2d` the samples are extended in both dimensions before
    the transforms; otherwise each one-dimensional transform extends its
    own samples.
    """

    def __init__(self, degrees, biases, sample_pts, pivot=1., lowring=True,
                 extrap=0, extrap_exp=2., extrap2d=False, threaded=False):
        _check_extrap(extrap)
        self.degrees = tuple(degrees)
        self.biases = tuple(biases)
        self._extrap = extrap
        self._extrap2d = extrap2d
        self._n_ext = 0
        inner = 0
        if extrap:
            nsamp = len(sample_pts)
            if nsamp % 2:
                raise ValueError(
                    "Input sample size must be even when extrapolation is "
                    f"used: {nsamp=}.")
            if extrap2d:
                nsamp_trans = int(2 ** np.ceil(np.log2(extrap_exp * nsamp)))
                self._n_ext = (nsamp_trans - nsamp) // 2
                sample_pts = extrap_loglin(sample_pts, self._n_ext)
            else:
                inner = extrap
        self._fbsjt = tuple(
            SphericalBesselTransform(d, b, sample_pts, pivot=pivot,
                                     lowring=lowring, extrap=inner,
                                     extrap_exp=extrap_exp)
            for d, b in zip(self.degrees, self.biases))
        self._post_slice = (slice(self._n_ext, -self._n_ext) if self._n_ext
                            else slice(None))
        self._logres = self._fbsjt[0]._logres
        self._pre_sampts = self._fbsjt[0]._pre_sampts
        self._post_sampts = self._fbsjt[0]._post_sampts[self._post_slice]

    @property
    def size(self):
        """Sample size of the transform in either dimension."""
        return self._fbsjt[0].size

    @property
    def pivot(self):
        """Pivot values of both dimensions."""
        return (self._fbsjt[0].pivot, self._fbsjt[1].pivot)

    def _perform_extrap(self, arr):
        a = np.asarray(arr)
        if not (self._extrap and self._extrap2d):
            return a
        rows = _extrapolate(a, self._n_ext, self._extrap)
        return _extrapolate(rows.T, self._n_ext, self._extrap).T

    def transform(self, pre_samples):
        """Transform 2-d samples; returns the post-transform points (in
        numpy.meshgrid 'ij' form) and samples."""
        a = self._perform_extrap(pre_samples)
        if a.ndim != 2 or a.shape[0] != a.shape[1]:
            raise ValueError("Sample array must be square and 2-d.")
        inter = self._fbsjt[1].transform(a)[1]
        post = self._fbsjt[0].transform(inter.T)[1].T
        post = post[self._post_slice, self._post_slice]
        return np.meshgrid(self._post_sampts, self._post_sampts, indexing='ij'), post

    def transform_cosmo_multipoles(self, direction, pre_samples):
        """Three-point multipoles between configuration and Fourier space
        (transforms.py:456-492): forward multiplies by (-i)^(l1 + l2),
        backward by i^(l1 + l2) / (2 pi)^6."""
        l1, l2 = self.degrees
        if direction in {'forward', 1}:
            factor = (-1j) ** (l1 + l2)
        elif direction in {'backward', -1}:
            factor = (1j) ** (l1 + l2) / (2 * np.pi) ** 6
        else:
            raise ValueError(
                "Transform direction must be 'forward'/1 or 'backward'/-1.")
        post_sampts, post_samples = self.transform(pre_samples)
        return post_sampts, post_samples * factor


def _resample(sampts, samples, size, spline, space):
    sampts = np.squeeze(sampts)
    samples = np.squeeze(samples)

    def grid(lo, hi, n):
        if space == 'log':
            return np.logspace(np.log10(lo), np.log10(hi), n, base=10)
        return np.linspace(lo, hi, n)

    if samples.ndim == 1:
        if sampts.ndim != 1:
            raise ValueError("Sample points must be 1-d.")
        res = grid(sampts[0], sampts[-1], size or len(sampts))
        return res, InterpolatedUnivariateSpline(sampts, samples, k=spline,
                                                 ext='raise')(res)
    if samples.ndim == 2:
        if sampts.ndim == 1:
            sampts = np.array([sampts, sampts])
        if len(sampts) != 2 or sampts[0].ndim != 1 or sampts[-1].ndim != 1:
            raise ValueError("Must provide (two) 1-d array(s) as sample points "
                             "for 2-d resampling.")
        size = size if isinstance(size, Sequence) else (size, size)
        res = (grid(sampts[0][0], sampts[0][-1], size[0] or len(sampts[0])),
               grid(sampts[-1][0], sampts[-1][-1], size[-1] or len(sampts[-1])))
        return res, RectBivariateSpline(sampts[0], sampts[-1], samples,
                                        kx=spline, ky=spline)(*res)
    raise ValueError(f"Sample points must be 1-d or 2-d: {samples.ndim=}.")


def resample_lglin(sampts, samples, size=None, spline=3):
    """Resample 1-d or 2-d samples at log-spaced points over the original
    range(s) by spline interpolation (transforms.py:526-614)."""
    return _resample(sampts, samples, size, spline, 'log')


def resample_lin(sampts, samples, size=None, spline=3):
    """Resample 1-d or 2-d samples at linearly spaced points over the
    original range(s) by spline interpolation (transforms.py:617-704)."""
    return _resample(sampts, samples, size, spline, 'lin')
