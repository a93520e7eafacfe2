"""Seeded synthetic catalogues for benchmarks and tests.

* :func:`lognormal_box` -- periodic-box lognormal mock: Gaussian field with
  power spectrum P(k) on a grid, delta = exp(delta_G - sigma^2/2) - 1,
  Poisson counts per cell with mean nbar V_cell (1 + delta), positions
  uniform inside each cell.
* :func:`lognormal_survey` -- the same field generated in a cubic box around
  an observer at the origin, sampled only inside a spherical-shell footprint
  (radius range, one quadrant of the sky x >= 0, y >= 0).
* :func:`uniform_shell_randoms` -- uniform randoms in that footprint.

The default spectrum is P(k) = 1e4 (k/0.02) / (1 + k/0.02)^3 (Mpc/h)^3.
Everything is deterministic given the seed (numpy PCG64).
"""

import numpy as np


def default_pk(k):
    x = k / 0.02
    return 1.e4 * x / (1. + x) ** 3


def gaussian_field(ngrid, boxsize, pk=default_pk, seed=0):
    """Real Gaussian random field on an ngrid^3 mesh with power spectrum pk
    (variance sum_k P(k)/V)."""
    rng = np.random.default_rng(seed)
    n, L = int(ngrid), float(boxsize)
    vcell = (L / n) ** 3
    white = rng.standard_normal((n, n, n))
    wk = np.fft.rfftn(white)
    del white
    kf = 2. * np.pi / L
    kx = np.fft.fftfreq(n, d=1. / n) * kf
    kz = np.fft.rfftfreq(n, d=1. / n) * kf
    k = np.sqrt(kx[:, None, None] ** 2 + kx[None, :, None] ** 2
                + kz[None, None, :] ** 2)
    amp = np.sqrt(pk(np.where(k > 0, k, 1.)) / vcell)
    amp[0, 0, 0] = 0.
    wk *= amp
    del amp, k
    return np.fft.irfftn(wk, s=(n, n, n), axes=(0, 1, 2))


def _lognormal(delta_g):
    s2 = delta_g.var()
    return np.exp(delta_g - 0.5 * s2) - 1.


def _poisson_positions(lam, origin, cell, rng):
    """Poisson counts per cell, uniform positions inside each cell."""
    counts = rng.poisson(lam.ravel())
    idx = np.repeat(np.arange(counts.size), counts)
    n = lam.shape[0]
    i, j, k = np.unravel_index(idx, (n, n, n))
    u = rng.random((len(idx), 3))
    x = origin + (i + u[:, 0]) * cell
    y = origin + (j + u[:, 1]) * cell
    z = origin + (k + u[:, 2]) * cell
    return x, y, z


def lognormal_box(nmean, boxsize, ngrid, seed=0, pk=default_pk):
    """Periodic-box lognormal mock with Poisson(nmean) particles in
    [0, boxsize)^3 (counts per cell Poisson, so the total is uneven across
    seeds)."""
    L = float(boxsize)
    delta = _lognormal(gaussian_field(ngrid, L, pk, seed))
    cell = L / ngrid
    lam = (nmean / ngrid ** 3) * (1. + delta)
    rng = np.random.default_rng(seed + 1_000_003)
    x, y, z = _poisson_positions(lam, 0., cell, rng)
    # Keep positions strictly inside [0, L).
    for a in (x, y, z):
        np.clip(a, 0., np.nextafter(L, 0.), out=a)
    return x, y, z


def shell_volume(rmin, rmax, sky_fraction=0.25):
    return sky_fraction * 4. / 3. * np.pi * (rmax ** 3 - rmin ** 3)


def in_footprint(x, y, z, rmin, rmax):
    r = np.sqrt(x * x + y * y + z * z)
    return (r >= rmin) & (r <= rmax) & (x >= 0.) & (y >= 0.)


def lognormal_survey(ndata, rmin=500., rmax=1500., gen_box=3200.,
                     ngrid=256, seed=2, pk=default_pk):
    """Lognormal data in the quadrant shell footprint, observer at the
    origin; expected count `ndata`."""
    L = float(gen_box)
    delta = _lognormal(gaussian_field(ngrid, L, pk, seed))
    cell = L / ngrid
    nbar = ndata / shell_volume(rmin, rmax)
    c = -0.5 * L + (np.arange(ngrid) + 0.5) * cell
    r2 = c[:, None, None] ** 2 + c[None, :, None] ** 2 + c[None, None, :] ** 2
    near = ((c[:, None, None] >= -cell) & (c[None, :, None] >= -cell)
            & (r2 >= (rmin - cell) ** 2) & (r2 <= (rmax + cell) ** 2))
    lam = np.where(near, nbar * cell ** 3 * (1. + delta), 0.)
    del delta, r2
    rng = np.random.default_rng(seed + 1_000_003)
    x, y, z = _poisson_positions(lam, -0.5 * L, cell, rng)
    keep = in_footprint(x, y, z, rmin, rmax)
    return x[keep], y[keep], z[keep], nbar


def uniform_shell_randoms(nrand, rmin=500., rmax=1500., seed=3):
    """Uniform points in the quadrant shell footprint."""
    rng = np.random.default_rng(seed)
    u = rng.random((3, int(nrand)))
    r = np.cbrt(rmin ** 3 + u[0] * (rmax ** 3 - rmin ** 3))
    mu = 2. * u[1] - 1.
    phi = 0.5 * np.pi * u[2]
    s = np.sqrt(1. - mu * mu)
    return r * s * np.cos(phi), r * s * np.sin(phi), r * mu
