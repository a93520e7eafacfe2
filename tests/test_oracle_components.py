"""Component pins of the CPU oracle: each restated routine is checked
against an independent implementation or a closed form, chosen so that a
plausible slip (dropped term, wrong sign or index, transposed operand) fails.

* reduced spherical harmonics vs scipy.special.sph_harm_y;
* Wigner 3-j symbols vs sympy's exact values;
* spherical Bessel j_l vs scipy.special.spherical_jn;
* the natural cubic spline vs scipy.interpolate.CubicSpline('natural');
* the 3-D FFT vs numpy.fft (FFTW sign convention);
* mesh assignment: total mass, single-particle window values, wrap-around;
* mode counts vs brute-force enumeration of the k-grid.
"""

import ctypes as C

import numpy as np
import pytest

from triumvirate_amd import _abi


def _ylm(lib, ell, m, v):
    out = np.zeros(2)
    lib.trvo_ylm(ell, m, _abi.dptr(np.ascontiguousarray(v, dtype=float)),
                 _abi.dptr(out))
    return out[0] + 1j * out[1]


@pytest.mark.parametrize('ell', [0, 1, 2, 3, 4])
def test_ylm_matches_scipy(oracle_lib, ell):
    from scipy.special import sph_harm_y
    g = np.random.default_rng(ell)
    for _ in range(40):
        v = g.normal(size=3) * g.uniform(0.1, 10.)
        r = np.linalg.norm(v)
        theta = np.arccos(v[2] / r)
        phi = np.arctan2(v[1], v[0]) % (2 * np.pi)
        for m in range(-ell, ell + 1):
            ref = np.sqrt(4 * np.pi / (2 * ell + 1)) * np.conj(
                sph_harm_y(ell, m, theta, phi))
            got = _ylm(oracle_lib, ell, m, v)
            assert abs(got - ref) < 1e-13, (ell, m, v)


def test_ylm_special_cases(oracle_lib):
    # (0,0) is exactly 1 everywhere, including the origin.
    assert _ylm(oracle_lib, 0, 0, [0., 0., 0.]) == 1.
    # The origin maps to 0 for l > 0.
    assert _ylm(oracle_lib, 2, 1, [0., 0., 0.]) == 0.
    # On the z axis the azimuth is 0: Y_l0 = 1 (mu = 1), Y_lm = 0 (m != 0).
    assert abs(_ylm(oracle_lib, 2, 0, [0., 0., 3.]) - 1.) < 1e-15
    assert abs(_ylm(oracle_lib, 2, 0, [0., 0., -3.]) - 1.) < 1e-15
    assert _ylm(oracle_lib, 2, 2, [0., 0., 1.]) == 0.
    # y < 0 branch: phi = 2 pi - acos(x / r_xy); Y_11 reduced = -sqrt(1/2)
    # sin(theta) e^{-i phi} for the reduced conjugated form.
    v = np.array([1., -1., 0.]) / np.sqrt(2.)
    ref = -np.sqrt(0.5) * np.exp(-1j * (2 * np.pi - np.pi / 4))
    assert abs(_ylm(oracle_lib, 1, 1, v) - ref) < 1e-15


def test_wigner_3j_matches_sympy(oracle_lib):
    from sympy.physics.wigner import wigner_3j
    for j1 in range(4):
        for j2 in range(4):
            for j3 in range(abs(j1 - j2), j1 + j2 + 1):
                for m1 in range(-j1, j1 + 1):
                    for m2 in range(-j2, j2 + 1):
                        m3 = -m1 - m2
                        if abs(m3) > j3:
                            continue
                        ref = float(wigner_3j(j1, j2, j3, m1, m2, m3))
                        got = oracle_lib.trvo_wigner_3j(j1, j2, j3, m1, m2, m3)
                        assert abs(got - ref) < 1e-14, (j1, j2, j3, m1, m2, m3)
    # Non-conserving m vanishes; triangle violation vanishes.
    assert oracle_lib.trvo_wigner_3j(1, 1, 0, 1, 0, 0) == 0.
    assert oracle_lib.trvo_wigner_3j(1, 0, 2, 0, 0, 0) == 0.


@pytest.mark.parametrize('ell', [0, 1, 2, 3, 6])
def test_spherical_bessel_matches_scipy(oracle_lib, ell):
    from scipy.special import spherical_jn
    xs = np.concatenate([[0., 1e-6, 1e-3, 0.05, 0.5], np.linspace(0.7, 60, 400),
                         [999.95, 1000., 1234.5, 4000.]])
    for x in xs:
        got = oracle_lib.trvo_sph_bessel(ell, float(x))
        ref = spherical_jn(ell, x)
        assert abs(got - ref) < 2e-14, (ell, x, got, ref)


@pytest.mark.parametrize('ell', [0, 1, 2])
def test_bessel_spline_matches_scipy_natural_spline(oracle_lib, ell):
    """Same nodes (x_i = 0.05 i, i <= 20000), natural boundary conditions:
    the oracle's spline and scipy's CubicSpline agree to round-off.  Against
    exact j_l the spline is within 2e-7 for x >= 0.5; below that the natural
    end condition (zero curvature at x = 0, while j_0''(0) = -1/3) leaves up
    to ~4e-5 -- a property of the reference's interpolator
    (maths.cpp:309-341) that is reproduced, not corrected."""
    from scipy.interpolate import CubicSpline
    from scipy.special import spherical_jn
    xn = 0.05 * np.arange(20001)
    cs = CubicSpline(xn, spherical_jn(ell, xn), bc_type='natural')
    g = np.random.default_rng(3)
    x = np.concatenate([g.uniform(0, 999.99, 3000), xn[:50] + 0.013,
                        [0., 0.05, 999.9]])
    out = np.zeros_like(x)
    oracle_lib.trvo_bessel_spline(ell, _abi.dptr(x), _abi.dptr(out), len(x))
    assert np.max(np.abs(out - cs(x))) < 1e-13
    far = x >= 0.5
    assert np.max(np.abs(out[far] - spherical_jn(ell, x[far]))) < 2e-7
    assert np.max(np.abs(out - spherical_jn(ell, x))) < 5e-5
    # Beyond the split the exact function is used.
    xb = np.array([1000., 1500.25])
    ob = np.zeros(2)
    oracle_lib.trvo_bessel_spline(ell, _abi.dptr(xb), _abi.dptr(ob), 2)
    assert np.max(np.abs(ob - spherical_jn(ell, xb))) < 1e-15


@pytest.mark.parametrize('shape', [(8, 8, 8), (16, 4, 32), (6, 10, 3)])
def test_fft_matches_numpy(oracle_lib, shape):
    g = np.random.default_rng(1)
    a = g.normal(size=shape) + 1j * g.normal(size=shape)
    ng = (C.c_int * 3)(*shape)
    for sign, ref in ((-1, np.fft.fftn(a)), (1, np.fft.ifftn(a) * a.size)):
        buf = a.copy().view(np.float64).ravel().copy()
        oracle_lib.trvo_fft3d(_abi.dptr(buf), ng, sign)
        got = buf.view(np.complex128).reshape(shape)
        assert np.max(np.abs(got - ref)) < 1e-11 * np.max(np.abs(ref))


def _params(ng, L, order):
    p = _abi.Params()
    for i in range(3):
        p.ngrid[i] = ng[i]
        p.boxsize[i] = L[i]
    p.assignment_order = order
    p.precision = 64
    return p


@pytest.mark.parametrize('order', [1, 2, 3, 4])
def test_assignment_mass_and_windows(oracle_lib, order):
    """Sum(field) * V_cell = sum(w); a single particle's per-axis weights are
    the scheme's window values at its fractional offset."""
    ng, L = (8, 8, 8), (80., 80., 80.)
    g = np.random.default_rng(order)
    n = 500
    x, y, z = g.uniform(0, 80., (3, n))
    w = g.uniform(0.5, 2., n)
    cat = _abi.CatalogueView(x, y, z, np.zeros(n), w, np.ones(n))
    wbuf = np.zeros(2 * n)
    wbuf[0::2] = w
    p = _params(ng, L, order)
    out = np.zeros(2 * 512)
    assert oracle_lib.trvo_assign(C.byref(cat.struct), _abi.dptr(wbuf),
                                  C.byref(p), 0, _abi.dptr(out)) == 0
    vcell = 1000.
    assert abs(out[0::2].sum() * vcell - w.sum()) < 1e-10 * w.sum()

    # One particle at grid coordinate (3.3, 0.7, 7.8) (wraps on z).
    loc = np.array([3.3, 0.7, 7.8])
    one = _abi.CatalogueView(*(loc * 10.)[:, None], np.zeros(1), np.ones(1),
                             np.ones(1))
    wb1 = np.array([1., 0.])
    out1 = np.zeros(2 * 512)
    oracle_lib.trvo_assign(C.byref(one.struct), _abi.dptr(wb1), C.byref(p),
                           0, _abi.dptr(out1))
    field = out1[0::2].reshape(8, 8, 8) * vcell

    def axis_weights(t):
        i0 = int(np.floor(t))
        s = t - i0
        if order == 1:
            c = i0 if s < 0.5 else i0 + 1
            return {c % 8: 1.}
        if order == 2:
            return {i0 % 8: 1 - s, (i0 + 1) % 8: s}
        if order == 3:
            c = i0 if s < 0.5 else i0 + 1
            d = t - c
            return {(c - 1) % 8: 0.5 * (0.5 - d) ** 2, c % 8: 0.75 - d * d,
                    (c + 1) % 8: 0.5 * (0.5 + d) ** 2}
        return {(i0 - 1) % 8: (1 - s) ** 3 / 6,
                i0 % 8: (4 - 6 * s * s + 3 * s ** 3) / 6,
                (i0 + 1) % 8: (4 - 6 * (1 - s) ** 2 + 3 * (1 - s) ** 3) / 6,
                (i0 + 2) % 8: s ** 3 / 6}

    expect = np.zeros((8, 8, 8))
    wx, wy, wz = (axis_weights(t) for t in loc)
    for i, a in wx.items():
        for j, b in wy.items():
            for k, c in wz.items():
                expect[i, j, k] += a * b * c
    assert np.max(np.abs(field - expect)) < 1e-12
    for ax in (wx, wy, wz):
        assert abs(sum(ax.values()) - 1.) < 1e-14


def test_mode_counts_brute_force():
    """n_modes per shell on 64^3, L = 1000 by direct enumeration of the
    k-grid equals the fixture's integer columns."""
    import conftest as cf
    n, L = 64, 1000.
    dk = 2 * np.pi / L
    idx = np.arange(n)
    s = np.where(idx < n // 2, idx, idx - n) * dk
    kx, ky, kz = np.meshgrid(s, s, s, indexing='ij')
    km = np.sqrt(kx * kx + ky * ky + kz * kz)
    edges = [0.005 + 0.025 * i for i in range(4)] + [0.105]
    counts = [int(((km >= edges[b]) & (km < edges[b + 1])).sum())
              for b in range(4)]
    ext = cf.load_ref_stats('bk000_diag_gpp.txt')
    assert counts == [int(v) for v in ext[2]]
