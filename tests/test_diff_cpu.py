"""CPU checks of the differentiable B_000 chain (tests/diff_reference.py):
the numpy forward equals the oracle's periodic-box B_000 diag (CIC) on the
same catalogue, and the adjoints equal central finite differences."""

import numpy as np

import diff_reference as DR
import parity as P

L, NG = 800., 16


def _catalogue(n=3000, seed=11):
    d = P.uniform_box_catalogue(n, L, seed=seed)
    x, y, z, nz, ws, wc = d.columns()
    return d, np.stack([x, y, z], axis=1)


def _oracle_diag(d):
    import oracle
    from triumvirate_amd import threept as T
    from triumvirate_amd.dataobjs import Binning
    import conftest as cf
    ps = cf.make_test_paramset(assignment='cic',
                               ngrid={'x': NG, 'y': NG, 'z': NG},
                               boxsize={'x': L, 'y': L, 'z': L})
    b = Binning('fourier', 'lin', bin_min=0.01, bin_max=0.05, num_bins=4)
    res = T._compute_3pt_stats_sim_like(
        oracle.backend(schedule=1), 'bispec', d, paramset=ps,
        degrees=(0, 0, 0), binning=b, form='diag')
    return res, b


def test_forward_matches_oracle():
    d, _ = _catalogue()
    res, b = _oracle_diag(d)          # re-centres d in the box, as reference
    x, y, z, _, _, _ = d.columns()
    pos = np.stack([x, y, z], axis=1)
    norm = res['_norm_factors'][0]
    delta = DR.assign(pos, np.ones(len(pos)), NG, L) - len(pos) / L ** 3
    bk, nmodes = DR.forward(delta, L, b.bin_edges, norm)
    assert np.array_equal(nmodes, res['nmodes_1'])
    np.testing.assert_allclose(bk, res['bk_raw'].real, rtol=1e-10)


def test_mesh_adjoint_matches_finite_differences():
    rng = np.random.default_rng(3)
    d, pos = _catalogue()
    delta = DR.assign(pos, np.ones(len(pos)), NG, L) - len(pos) / L ** 3
    edges = np.linspace(0.01, 0.05, 5)
    gB = rng.standard_normal(4)
    grad = DR.adjoint(delta, L, edges, 1., gB)
    for trial in range(3):
        v = rng.standard_normal(delta.shape) * np.std(delta)
        h = 1e-4
        lp = gB @ DR.forward(delta + h * v, L, edges, 1.)[0]
        lm = gB @ DR.forward(delta - h * v, L, edges, 1.)[0]
        fd = (lp - lm) / (2 * h)
        an = np.sum(grad * v)
        assert abs(fd - an) < 1e-7 * max(abs(an), abs(fd)), (trial, fd, an)


def test_assign_adjoint_matches_finite_differences():
    rng = np.random.default_rng(4)
    _, pos = _catalogue(200, seed=5)
    w = rng.uniform(0.5, 1.5, len(pos))
    gm = rng.standard_normal((NG, NG, NG))

    def f(p, ww):
        return np.sum(gm * DR.assign(p, ww, NG, L))

    gpos, gw = DR.assign_adjoint(gm, pos, w, NG, L)
    eps = 1e-3
    for p in range(5):
        for a in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[p, a] += eps
            pm[p, a] -= eps
            fd = (f(pp, w) - f(pm, w)) / (2 * eps)
            assert abs(fd - gpos[p, a]) <= 1e-6 * abs(gpos).max(), (p, a)
        wp, wm = w.copy(), w.copy()
        wp[p] += 1e-3
        wm[p] -= 1e-3
        fd = (f(pos, wp) - f(pos, wm)) / 2e-3
        assert abs(fd - gw[p]) <= 1e-8 * abs(gw).max()
