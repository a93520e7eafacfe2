"""FFTLog transforms (reference: src/fftlog.cpp, include/fftlog.hpp:60-388,
_fftlog.pyx, transforms.py, _arrayops.py:249-330).

Downstream tooling of the estimators (window convolution of model
multipoles): one-dimensional transforms of 10^3-10^4 log-spaced samples,
which run on the host with numpy's FFT.

The biased Hankel transform

    b(k) = int_0^inf k dr (k r)^q J_mu(k r) a(r)

of samples a(r_j) at log-spaced r_j (spacing dL in ln r) is evaluated by
FFTLog (A. J. S. Hamilton, MNRAS 312, 257 (2000)): the samples are
transformed by an N-point FFT, multiplied by the kernel

    u_m = 2^q Gamma((mu + 1 + q)/2 + i m y) / Gamma((mu + 1 - q)/2 - i m y)
          * exp(-2 i m y ln(k_c r_c / 2)),   y = pi / (N dL),

for m <= N/2 (u_{N-m} = conj(u_m), u_{N/2} real), divided by N and
transformed by a second forward FFT, which returns b at
k_j = k_c r_c exp(-dL) / r_{N-1-j}.  The pivot k_c r_c, given or chosen
by the low-ringing condition (the phase of the kernel at the Nyquist mode
a multiple of pi), follows fftlog.cpp's conventions so that the output
sample points coincide with the reference's.  Gamma of complex arguments
is scipy's loggamma; only its exponential enters the kernel, and the
low-ringing adjustment is invariant under 2 pi shifts of its phase.

The spherical Bessel transform of degree l,

    g(k) = 4 pi int_0^inf r^2 dr j_l(k r) f(r),

is the Hankel transform of order l + 1/2 of r^{3/2} f, times
(2 pi / k)^{3/2}.
"""

import warnings

import numpy as np
from scipy.special import loggamma

__all__ = ['HankelTransform', 'extrap_pad', 'extrap_lin', 'extrap_loglin',
           'extrap_loglin_oscil', 'DivergenceWarning']


# ---------------------------------------------------------------------------
# Extrapolation (_arrayops.py:249-330), along the last axis
# ---------------------------------------------------------------------------

def extrap_pad(a, n_ext, c_lower, c_upper):
    """Pad an array along its last axis with constants on either side."""
    a = np.asarray(a)
    side = a.shape[:-1] + (n_ext,)
    dtype = np.result_type(a, c_lower, c_upper)
    return np.concatenate([np.broadcast_to(np.asarray(c_lower, dtype), side),
                           a.astype(dtype, copy=False),
                           np.broadcast_to(np.asarray(c_upper, dtype), side)],
                          axis=-1)


def _as_float(a):
    a = np.asarray(a)
    return a if np.iscomplexobj(a) else a.astype(float, copy=False)


def extrap_lin(a, n_ext):
    """Extend an array along its last axis linearly with the spacing of
    the two end entries on either side."""
    a = _as_float(a)
    left = a[..., :1] + (a[..., 1:2] - a[..., :1]) * np.arange(-n_ext, 0)
    right = a[..., -1:] + (a[..., -1:] - a[..., -2:-1]) * np.arange(1, n_ext + 1)
    return np.concatenate([left, a, right], axis=-1)


def extrap_loglin(a, n_ext):
    """Extend an array along its last axis log-linearly with the ratio of
    the two end entries on either side.  Each end pair must share its
    sign, or the end value be zero, in which case that side is extended
    with zeros (_arrayops.py:112-123, 304-329).  Complex arrays are
    extended by the complex ratio, with the sign test on the real part of
    the end products."""
    a = _as_float(a)
    if a.ndim < 1 or a.shape[-1] < 2:
        raise ValueError("Log-linear extrapolation needs at least two "
                         "entries.")
    lo, lo1, hi, hi1 = a[..., :1], a[..., 1:2], a[..., -1:], a[..., -2:-1]
    same_lo = (np.real(lo * lo1) > 0.) | (lo == 0.)
    same_hi = (np.real(hi * hi1) > 0.) | (hi == 0.)
    if not np.all(same_lo & same_hi):
        raise ValueError("Log-linear extrapolation needs end pairs of one "
                         "sign (or zero end values).")
    with np.errstate(divide='ignore', invalid='ignore'):
        left = lo * (lo / lo1) ** np.arange(n_ext, 0, -1)
        right = hi * (hi / hi1) ** np.arange(1, n_ext + 1)
    return np.concatenate([left, a, right], axis=-1)


class DivergenceWarning(UserWarning):
    """A tail of the samples grows outwards; it is padded with its end
    value instead."""


def _lngrad(a):
    with np.errstate(divide='ignore', invalid='ignore'):
        return np.nan_to_num(np.gradient(np.log(np.abs(a))))


def _sign_changes(v):
    return np.nonzero(np.diff(np.sign(v)))[0]


def _tail_rate(a, inner, quarter, side):
    """Log-slope per sample that extends the samples beyond `end` on
    `side` ('l' or 'r'), so that the extension decays away from the data:
    the mean log-gradient of the tail when it oscillates, else the
    end-point log-gradient, else the mean over the quarter of the samples
    on that side; zero (constant padding) when each of these grows
    outwards."""
    tail = a[:inner] if side == 'l' else a[inner:]
    g = _lngrad(tail)
    out = 1. if side == 'l' else -1.     # sign of an outward-decaying slope
    if _sign_changes(g).size:
        m = np.mean(g)
        if out * m >= 0.:
            return -m if side == 'l' else m
        warnings.warn(f"Divergent behaviour detected at the "
                      f"{'left' if side == 'l' else 'right'} tail. "
                      "Constant padding is used instead.", DivergenceWarning,
                      stacklevel=4)
        return 0.
    with np.errstate(divide='ignore', invalid='ignore'):
        if side == 'l':
            gend = np.nan_to_num(np.log(np.abs(a[1])) - np.log(np.abs(a[0])))
        else:
            gend = np.nan_to_num(np.log(np.abs(a[-1])) - np.log(np.abs(a[-2])))
    if out * gend >= 0.:
        return -gend if side == 'l' else gend
    m = np.nan_to_num(np.mean(_lngrad(a[:quarter] if side == 'l' else a[quarter:])))
    if out * m >= 0.:
        return -m if side == 'l' else m
    warnings.warn(f"Divergent behaviour detected at the "
                  f"{'left' if side == 'l' else 'right'} endpoints. "
                  "Constant padding is used instead.", DivergenceWarning,
                  stacklevel=4)
    return 0.


def extrap_loglin_oscil(a, n_ext):
    """Extend a 1-d array of possibly oscillating samples log-linearly on
    either side (_arrayops.py:332-487).

    The slope of each side is taken from a tail chosen to avoid sign
    changes and turning points: at most up to the first (last) sign change,
    the last (first) turning point of the log-magnitude before (after) it,
    and the quarter of the samples on that side of the peak, with at least
    three samples kept; see `_tail_rate` for the slope itself.  Complex
    samples take their sign changes from the real part and are extended
    with the complex end values."""
    a = _as_float(a)
    if a.ndim != 1 or a.size < 2:
        raise ValueError("Extrapolation needs a 1-d array of at least two "
                         "entries.")
    n = a.size
    if n < 8:
        warnings.warn("Array length is less than 8. Extrapolation may not "
                      "be accurate.", UserWarning, stacklevel=2)
    # Complex samples (a transform's output whose imaginary part is
    # round-off): sign changes from the real part, slopes from |a|.
    sc = _sign_changes(np.real(a))
    lbank, rbank = ((max(sc[0] + 1, 3), min(sc[-1] + 1, n - 4)) if sc.size
                    else (n - 1, 0))
    tl = _sign_changes(_lngrad(a[:lbank]))
    ldbank = max(tl[-1] + 1, 3) if tl.size else n - 1
    tr = _sign_changes(_lngrad(a[rbank:]))
    rdbank = min(rbank + tr[0], n - 4) if tr.size else 0
    peak = int(np.argmax(np.abs(a)))
    if peak in (0, n - 1):
        peak = n // 2
    lquarter = max((peak + 1) // 2 - 1, 3)
    rquarter = min(peak + (n - 1 - peak) // 2, n - 4)
    ltail = min(lbank, ldbank, lquarter)
    rtail = max(rbank, rdbank, rquarter)
    sl = _tail_rate(a, ltail, lquarter, 'l')
    sr = _tail_rate(a, rtail, rquarter, 'r')
    left = a[0] * np.exp(sl * np.arange(n_ext, 0, -1))
    right = a[-1] * np.exp(sr * np.arange(1, n_ext + 1))
    return np.concatenate([left, a, right])


def _extrapolate(a, n_ext, option):
    if option == 1:
        return extrap_pad(a, n_ext, 0., 0.)
    if option == 2:
        a = np.asarray(a)
        side = a.shape[:-1] + (n_ext,)
        return np.concatenate([np.broadcast_to(a[..., :1], side), a,
                               np.broadcast_to(a[..., -1:], side)], axis=-1)
    if option == 3:
        return extrap_lin(a, n_ext)
    if option == 4:
        return extrap_loglin(a, n_ext)
    if option == 5:
        a = np.asarray(a)
        if a.ndim == 1:
            return extrap_loglin_oscil(a, n_ext)
        rows = a.reshape(-1, a.shape[-1])
        return np.array([extrap_loglin_oscil(r, n_ext) for r in rows]).reshape(
            a.shape[:-1] + (a.shape[-1] + 2 * n_ext,))
    raise ValueError(f"Unsupported extrapolation option: {option}.")


def _check_loglin(x):
    x = np.asarray(x, dtype=float)
    if x.ndim != 1 or x.size < 2:
        raise ValueError("The number of sample points must be at least 2.")
    if not np.all(x > 0):
        raise ValueError("The sample points are not logarithmically spaced.")
    d = np.diff(np.log(x))
    if not np.allclose(d, d[0], rtol=1e-5, atol=0.):
        raise ValueError("The sample points are not logarithmically spaced.")
    return x


# ---------------------------------------------------------------------------
# Hankel transform
# ---------------------------------------------------------------------------

class HankelTransform:
    """Biased Hankel transform by FFTLog (fftlog.hpp:60-240,
    _fftlog.pyx:20-145).

    Parameters
    ----------
    mu : float
        Order of the transform.
    q : float
        Power-law bias index.
    x : array of float
        Logarithmically spaced pre-transform sample points (even in length
        if extrapolation is used).
    kr_c : float
        Pivot value; adjusted by the low-ringing condition when `lowring`
        (computed from the kernel when non-positive).
    lowring : bool, optional
        Low-ringing pivot (default True).
    extrap : int, optional
        Extrapolation of the samples inside the transform: 0 none
        (default), 1 zeros, 2 end values, 3 linear, 4 log-linear, 5
        log-linear with oscillation-aware tails; the
        transform then takes the smallest power of two >= `extrap_exp`
        times the sample size, and the samples are taken as real.
    extrap_exp : float, optional
        Sample-size expansion factor for extrapolation (default 2).
    threaded : bool, optional
        Accepted for interface compatibility (numpy's FFT is used).
    """

    def __init__(self, mu, q, x, kr_c, lowring=True, extrap=0, extrap_exp=2.,
                 threaded=False):
        del threaded
        x = _check_loglin(x)
        self.order = float(mu)
        self.bias = float(q)
        self._nsamp = x.size
        self._logres = np.log(x[-1] / x[0]) / (x.size - 1)
        self._extrap = int(extrap or 0)
        if self._extrap:
            if x.size % 2:
                raise ValueError(
                    "The number of sample points must be even for "
                    "extrapolation.")
            n_trans = int(2 ** np.ceil(np.log2(extrap_exp * x.size)))
            if n_trans < x.size:
                raise ValueError("The sample size expansion factor results "
                                 "in a shrunken sample size.")
        else:
            n_trans = x.size
        self._nsamp_trans = n_trans
        self._n_ext = (n_trans - x.size) // 2
        if lowring:
            self._pivot = self._lowring_pivot(self._logres, kr_c)
        else:
            if not kr_c > 0.:
                raise ValueError("Pivot value must be positive.")
            self._pivot = float(kr_c)
        self._kernel = self._kernel_coeff()
        self._pre_sampts = x
        self._post_sampts = self._pivot * np.exp(-self._logres) / x[::-1]

    @property
    def size(self):
        """Sample size of the transform."""
        return self._nsamp_trans

    @property
    def pivot(self):
        """Pivot value used."""
        return self._pivot

    def _lowring_pivot(self, dL, kr_c):
        mu, q = self.order, self.bias
        y = np.pi / (2. * dL)
        if kr_c > 0.:
            phase = (loggamma((mu + 1. + q) / 2. + 1j * y).imag
                     + loggamma((mu + 1. - q) / 2. + 1j * y).imag)
            arg = np.log(2. / kr_c) / dL + phase / np.pi
            # Nearest integer with halves away from zero.
            nearest = np.copysign(np.floor(abs(arg) + 0.5), arg)
            return kr_c * np.exp(dL * (arg - nearest))
        nu = q + 2j * y
        lnratio = (loggamma((mu + 1. + nu) / 2.) - loggamma((mu + 1. - nu) / 2.))
        lnkr = np.log(2.) + lnratio.imag / (2. * y)
        return np.exp(lnkr - np.floor(lnkr / dL) * dL)

    def _kernel_coeff(self):
        mu, q, n = self.order, self.bias, self._nsamp_trans
        y = np.pi / (n * self._logres)
        t = -2. * y * np.log(self._pivot / 2.)
        m = np.arange(n // 2 + 1)
        lg_p = loggamma((mu + 1. + q) / 2. + 1j * m * y)
        lg_m = loggamma((mu + 1. - q) / 2. + 1j * m * y)
        # Gamma(x_m - i m y) = conj(Gamma(x_m + i m y)).
        u_half = np.exp(q * np.log(2.) + lg_p.real - lg_m.real
                        + 1j * (m * t + lg_p.imag + lg_m.imag))
        u = np.empty(n, dtype=complex)
        u[:n // 2 + 1] = u_half
        u[n // 2 + 1:] = np.conj(u_half[1:(n + 1) // 2][::-1])
        if n % 2 == 0:
            u[n // 2] = u[n // 2].real
        return u

    def transform(self, fx):
        """Transform samples at the pre-transform points; returns the
        post-transform points and samples.  Leading axes of `fx`, if any,
        are independent sample sets transformed together."""
        a = np.asarray(fx)
        if a.ndim < 1 or a.shape[-1] != self._nsamp:
            raise ValueError("Sample array does not match the sample points.")
        if self._extrap:
            a = _extrapolate(np.real(a).astype(float), self._n_ext, self._extrap)
        a = a.astype(complex)
        b = np.fft.fft(np.fft.fft(a, axis=-1) * (self._kernel / self._nsamp_trans),
                       axis=-1)
        b = b[..., self._n_ext:self._n_ext + self._nsamp]
        return self._post_sampts.copy(), b
