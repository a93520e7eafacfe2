"""Helpers shared by the parity tests: synthetic catalogues, result
comparison with the tolerances the tests state, and reference-fixture cases.
"""

import numpy as np

import conftest as cf

# Reference fixture cases: (statistic, degrees, form, idx_bin, geometry).
FIXTURE_CASES = [
    ('bispec', (0, 0, 0), 'diag', None, 'gpp'),
    ('bispec', (0, 0, 0), 'diag', None, 'lpp'),
    ('bispec', (0, 0, 0), 'row', 0, 'gpp'),
    ('bispec', (0, 0, 0), 'row', 0, 'lpp'),
    ('bispec', (2, 0, 2), 'diag', None, 'gpp'),
    ('bispec', (2, 0, 2), 'diag', None, 'lpp'),
    ('3pcf', (0, 0, 0), 'diag', None, 'gpp'),
    ('3pcf', (0, 0, 0), 'diag', None, 'lpp'),
    ('3pcf', (0, 0, 0), 'row', 0, 'gpp'),
    ('3pcf', (0, 0, 0), 'row', 0, 'lpp'),
    ('3pcf', (2, 0, 2), 'diag', None, 'gpp'),
    ('3pcf', (2, 0, 2), 'diag', None, 'lpp'),
]


def fixture_name(kind, deg, form, idx, geo):
    stem = 'bk' if kind == 'bispec' else 'zeta'
    ftag = form if form == 'diag' else f'row{idx}'
    return f"{stem}{''.join(map(str, deg))}_{ftag}_{geo}.txt"


def keys(kind):
    if kind == 'bispec':
        return (['k1_bin', 'k1_eff', 'nmodes_1', 'k2_bin', 'k2_eff',
                 'nmodes_2'], 'bk_raw', 'bk_shot')
    return (['r1_bin', 'r1_eff', 'npairs_1', 'r2_bin', 'r2_eff', 'npairs_2'],
            'zeta_raw', 'zeta_shot')


def run_fixture_case(backend, kind, deg, form, idx, geo, precision=None):
    """Run one reference-fixture configuration through the host drivers."""
    from triumvirate_amd import threept as T
    ps = cf.make_test_paramset()
    b = cf.binning_fourier() if kind == 'bispec' else cf.binning_config()
    if geo == 'gpp':
        return T._compute_3pt_stats_sim_like(
            backend, kind, cf.make_test_data_catalogue(), paramset=ps,
            degrees=deg, binning=b, form=form, idx_bin=idx,
            precision=precision)
    return T._compute_3pt_stats_survey_like(
        backend, kind, cf.make_test_data_catalogue(),
        cf.make_test_rand_catalogue(), paramset=ps, degrees=deg, binning=b,
        form=form, idx_bin=idx, precision=precision)


def check_against_fixture(res, kind, name, rtol=2e-9):
    """Compare with a reference fixture file.

    The fixture prints 10 significant digits, so real parts are pinned to
    ~5e-10 relative; `rtol` is taken relative to max |Re| of the column.
    Integer counts must be equal.  Imaginary parts in the fixtures are FFT
    round-off (|Im| ~ 1e-16 |Re|), so they are only required to be small.
    """
    ext = cf.load_ref_stats(name)
    cols, raw, shot = keys(kind)
    for i, k in enumerate(cols):
        if k.startswith(('nmodes', 'npairs')):
            assert np.array_equal(res[k], ext[i].astype(np.int64)), (name, k)
        else:
            np.testing.assert_allclose(res[k], ext[i], rtol=1e-9, atol=0,
                                       err_msg=f"{name} {k}")
    for key, col in ((raw, 6), (shot, 8)):
        scale = np.max(np.abs(ext[col]))
        err = np.max(np.abs(res[key].real - ext[col])) / scale
        assert err < rtol, (name, key, err)
        assert np.max(np.abs(res[key].imag)) < 1e-8 * scale, (name, key)


def compare_results(a, b, kind, rtol=1e-10):
    """GPU vs oracle: counts bit-exact; Re and Im of raw/shot within `rtol`
    relative to max |value| of the multipole; bins and effective
    coordinates within 1e-12."""
    cols, raw, shot = keys(kind)
    for k in cols:
        if k.startswith(('nmodes', 'npairs')):
            assert np.array_equal(a[k], b[k]), k
        else:
            np.testing.assert_allclose(a[k], b[k], rtol=1e-12, atol=0,
                                       err_msg=k)
    errs = {}
    for key in (raw, shot):
        scale = np.max(np.abs(b[key]))
        err = np.max(np.abs(a[key] - b[key])) / scale if scale > 0 else 0.
        errs[key] = err
        assert err < rtol, (key, err)
    return errs


def uniform_box_catalogue(n, L, seed, weights=False):
    from triumvirate_amd.catalogue import ParticleCatalogue
    g = np.random.default_rng(seed)
    x, y, z = g.uniform(0., L, size=(3, n))
    ws = g.uniform(0.5, 1.5, n) if weights else 1.
    wc = g.uniform(0.8, 1.2, n) if weights else 1.
    return ParticleCatalogue(x, y, z, nz=n / L**3, ws=ws, wc=wc)


def survey_pair(nd, nr, L, seed, weights=False):
    """A clustered-ish data catalogue and uniform randoms in a sphere-cut box
    (observer at the origin, positions offset so lines of sight vary)."""
    from triumvirate_amd.catalogue import ParticleCatalogue
    g = np.random.default_rng(seed)

    def draw(n):
        pts = []
        while sum(len(p) for p in pts) < n:
            p = g.uniform(-0.5 * L, 0.5 * L, size=(3 * n, 3))
            p = p[np.linalg.norm(p, axis=1) < 0.45 * L]
            pts.append(p)
        p = np.concatenate(pts)[:n]
        return p + np.array([0.3 * L, 0.2 * L, 0.1 * L])

    pd = draw(nd)
    # Clustering: pull a third of the data towards a few centres.
    centres = draw(5)
    k = nd // 3
    pd[:k] = centres[g.integers(0, 5, k)] + g.normal(0, 0.02 * L, (k, 3))
    pr = draw(nr)
    nbar = nd / (4. / 3. * np.pi * (0.45 * L) ** 3)
    kw_d = {}
    kw_r = {}
    if weights == 'uniform':
        # One weight per catalogue (not 1): the engine's uniform-weight path.
        kw_d = dict(ws=np.full(nd, 1.5), wc=np.full(nd, 0.8))
        kw_r = dict(ws=np.full(nr, 0.7), wc=np.full(nr, 1.1))
    elif weights == 'one-off':
        # Uniform but for the last particle of each catalogue: the upload's
        # weight column is filled lazily for ranges that stay uniform.
        kw_d = dict(ws=np.ones(nd), wc=np.ones(nd))
        kw_r = dict(ws=np.ones(nr), wc=np.ones(nr))
        kw_d['ws'][-1] = 2.
        kw_r['wc'][-1] = 0.5
    elif weights == 'random-one-off':
        # Uniform data weights; randoms uniform but for their last particle,
        # so only the data side of the engine's uniform-weight path applies.
        kw_d = dict(ws=np.full(nd, 1.5), wc=np.full(nd, 0.8))
        kw_r = dict(ws=np.ones(nr), wc=np.ones(nr))
        kw_r['wc'][-1] = 0.5
    elif weights:
        kw_d = dict(ws=g.uniform(0.5, 1.5, nd), wc=g.uniform(0.8, 1.2, nd))
        kw_r = dict(ws=g.uniform(0.5, 1.5, nr), wc=g.uniform(0.8, 1.2, nr))
    d = ParticleCatalogue(pd[:, 0], pd[:, 1], pd[:, 2], nz=nbar, **kw_d)
    r = ParticleCatalogue(pr[:, 0], pr[:, 1], pr[:, 2], nz=nbar, **kw_r)
    return d, r
