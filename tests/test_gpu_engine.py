"""GPU parity tests of the HIP engine (run on an MI355X).

Every estimator call here goes through the engine's C ABI
(``libtrvgpu.so``); results are compared with the CPU FP64 oracle and the
reference's own fixture files.

Tolerances (stated per test):
  * mode / pair counts: bit-exact;
  * FP64 estimates vs the oracle: 1e-10 relative to max |B| of the multipole;
  * FP64 estimates vs reference fixtures: 2e-9 (the files print 10 digits);
  * FP32 estimates vs FP64: 1e-4 relative to max |B| (see test docstring).
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
    lib = _lib.load()
    n = C.c_int(0)
    assert lib.trvg_device_count(C.byref(n)) == 0
    if n.value == 0:
        pytest.fail("No HIP device visible to the engine.")
    return lib


@pytest.fixture(scope='module')
def gpu_backend(eng):
    from triumvirate_amd.threept import CLibBackend
    return CLibBackend(eng, 'trvg_')


@pytest.fixture(scope='module')
def oracle_backend():
    import oracle
    return oracle.backend(schedule=1)


def _params(ngrid, L, order=3, prec=64):
    from triumvirate_amd import _abi
    p = _abi.Params()
    for i in range(3):
        p.ngrid[i] = ngrid[i]
        p.boxsize[i] = L[i]
    p.assignment_order = order
    p.precision = prec
    p.device = -1
    return p


# ---------------------------------------------------------------------------
# Components
# ---------------------------------------------------------------------------

@pytest.mark.parametrize('shape', [(32, 32, 32), (16, 32, 64), (64, 16, 8),
                                   (8, 8, 256), (8, 8, 512), (8, 256, 8),
                                   (256, 8, 8)])
@pytest.mark.parametrize('prec', [64, 32])
def test_fft3d_matches_numpy(eng, shape, prec):
    """Forward/backward unnormalised FFT vs numpy.fft (FFTW convention).
    Tolerance: 1e-12 (FP64) / 2e-6 (FP32) relative to max |X|."""
    from triumvirate_amd import _abi
    print(f"fft3d shape={shape} prec={prec}", flush=True)
    g = np.random.default_rng(7)
    a = g.normal(size=shape) + 1j * g.normal(size=shape)
    p = _params(shape, (1., 1., 1.), prec=prec)
    tol = 1e-12 if prec == 64 else 2e-6
    for sign, ref in ((-1, np.fft.fftn(a)), (+1, np.fft.ifftn(a) * a.size)):
        buf = np.ascontiguousarray(a.astype(np.complex128)).view(np.float64).copy()
        assert eng.trvg_fft3d(C.byref(p), _abi.dptr(buf), sign) == 0, \
            eng.trvg_last_error()
        out = buf.view(np.complex128).reshape(shape)
        err = np.max(np.abs(out - ref)) / np.max(np.abs(ref))
        assert err < tol, (sign, err)


@pytest.mark.parametrize('order', [1, 2, 3, 4])
@pytest.mark.parametrize('mode,ell,m', [(0, 0, 0), (1, 0, 0), (2, 2, 1),
                                        (3, 2, -2)])
def test_assignment_matches_oracle(eng, oracle_lib, order, mode, ell, m):
    """Mesh assignment vs the oracle's restatement of field.cpp:569-1112,
    including wrap-around at both box faces.  Tolerance 1e-12 relative to
    the largest cell value (FP64 atomics reorder the sums)."""
    from triumvirate_amd import _abi
    print(f"assign order={order} mode={mode} l={ell} m={m}", flush=True)
    ng, L = (16, 32, 16), (100., 200., 160.)
    g = np.random.default_rng(11 + order)
    n = 9000
    x = g.uniform(0, L[0], n)
    y = g.uniform(0, L[1], n)
    z = g.uniform(0, L[2], n)
    # A dense clump inside one 8^3-cell tile: several work chunks per tile.
    x[4000:] = g.uniform(10., 40., 5000)
    y[4000:] = g.uniform(10., 40., 5000)
    z[4000:] = g.uniform(10., 40., 5000)
    # Put some particles right at the faces and cell edges.
    x[:20] = 0.
    y[20:40] = np.nextafter(L[1], 0)
    z[40:60] = L[2] / ng[2] * 3
    ws = g.uniform(0.5, 1.5, n)
    wc = g.uniform(0.5, 1.5, n)
    los = g.normal(size=(n, 3))
    los /= np.linalg.norm(los, axis=1)[:, None]
    los = np.ascontiguousarray(los)
    cat = _abi.CatalogueView(x, y, z, np.zeros(n), ws, wc)
    p = _params(ng, L, order=order)
    nm = int(np.prod(ng))
    out = np.zeros(2 * nm)
    assert eng.trvg_mesh_assign(C.byref(cat.struct), _abi.dptr(los),
                                C.byref(p), mode, ell, m, _abi.dptr(out)) == 0, \
        eng.trvg_last_error()
    # Oracle: explicit complex weights.
    w = ws * wc
    if mode == 0:
        wt = np.ones(n, complex)
    elif mode == 1:
        wt = w.astype(complex)
    else:
        yl = np.zeros((n, 2))
        for i in range(n):
            oracle_lib.trvo_ylm(ell, m, _abi.dptr(los[i].copy()),
                                _abi.dptr(yl[i]))
        yv = yl[:, 0] + 1j * yl[:, 1]
        wt = yv * w if mode == 2 else np.conj(yv) * w**2
    wbuf = np.ascontiguousarray(wt).view(np.float64).copy()
    ref = np.zeros(2 * nm)
    assert oracle_lib.trvo_assign(C.byref(cat.struct), _abi.dptr(wbuf),
                                  C.byref(p), 0, _abi.dptr(ref)) == 0
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(out - ref)) / scale < 1e-12


# ---------------------------------------------------------------------------
# End to end
# ---------------------------------------------------------------------------

@pytest.mark.parametrize('case', P.FIXTURE_CASES,
                         ids=[P.fixture_name(*c) for c in P.FIXTURE_CASES])
def test_reference_fixtures_public_api(eng, case):
    """The public API (compute_bispec, compute_3pcf, ..._in_gpp_box) on the
    reference's test catalogues reproduces the reference fixture files."""
    import triumvirate_amd as trv
    kind, deg, form, idx, geo = case
    name = P.fixture_name(*case)
    print(f"fixture {name}", flush=True)
    ps = cf.make_test_paramset()
    b = cf.binning_fourier() if kind == 'bispec' else cf.binning_config()
    data = cf.make_test_data_catalogue()
    if geo == 'gpp':
        fn = (trv.compute_bispec_in_gpp_box if kind == 'bispec'
              else trv.compute_3pcf_in_gpp_box)
        res = fn(data, degrees=deg, binning=b, form=form, idx_bin=idx,
                 paramset=ps)
    else:
        fn = trv.compute_bispec if kind == 'bispec' else trv.compute_3pcf
        res = fn(data, cf.make_test_rand_catalogue(), degrees=deg, binning=b,
                 form=form, idx_bin=idx, paramset=ps)
    P.check_against_fixture(res, kind, name)


@pytest.mark.parametrize('case', P.FIXTURE_CASES,
                         ids=[P.fixture_name(*c) for c in P.FIXTURE_CASES])
def test_fixture_configs_match_oracle(gpu_backend, oracle_backend, case):
    """Same configurations, GPU vs oracle at FP64 tolerance 1e-10."""
    name = P.fixture_name(*case)
    print(f"oracle-compare {name}", flush=True)
    kind = case[0]
    a = P.run_fixture_case(gpu_backend, *case)
    b = P.run_fixture_case(oracle_backend, *case)
    P.compare_results(a, b, kind, rtol=1e-10)


FORM_CASES = [
    ('bispec', (0, 0, 0), 'full', None, 'gpp', 4),
    ('bispec', (2, 0, 2), 'full', None, 'gpp', 2),
    ('bispec', (0, 0, 0), 'off-diag', 1, 'lpp', 3),
    ('bispec', (0, 0, 0), 'off-diag', -2, 'gpp', 1),
    ('bispec', (2, 0, 2), 'row', 1, 'lpp', 2),
    ('bispec', (1, 1, 0), 'full', None, 'lpp', 3),
    ('bispec', (0, 2, 2), 'full', None, 'lpp', 4),
    ('3pcf', (1, 1, 0), 'full', None, 'lpp', 3),
    ('3pcf', (0, 0, 0), 'full', None, 'gpp', 2),
    ('3pcf', (2, 0, 2), 'off-diag', 0, 'lpp', 4),
]


@pytest.mark.parametrize('case', FORM_CASES,
                         ids=['-'.join(map(str, c)) for c in FORM_CASES])
def test_forms_and_schemes_match_oracle(gpu_backend, oracle_backend, case):
    """Forms beyond the fixtures (full/triu, off-diag, row), other
    assignment schemes, non-unit weights, 32^3 mesh: GPU vs oracle, 1e-10."""
    from triumvirate_amd import threept as T
    from triumvirate_amd.dataobjs import Binning
    kind, deg, form, idx, geo, order = case
    print(f"form-case {case}", flush=True)
    scheme = {1: 'ngp', 2: 'cic', 3: 'tsc', 4: 'pcs'}[order]
    ps = cf.make_test_paramset(
        assignment=scheme, ngrid={'x': 32, 'y': 32, 'z': 32},
        boxsize={'x': 800., 'y': 800., 'z': 800.})
    if kind == 'bispec':
        b = Binning('fourier', 'lin', bin_min=0.01, bin_max=0.11, num_bins=5)
    else:
        b = Binning('config', 'lin', bin_min=30., bin_max=130., num_bins=5)
    out = []
    for be in (gpu_backend, oracle_backend):
        if geo == 'gpp':
            d = P.uniform_box_catalogue(3000, 800., seed=5, weights=False)
            out.append(T._compute_3pt_stats_sim_like(
                be, kind, d, paramset=ps, degrees=deg, binning=b, form=form,
                idx_bin=idx))
        else:
            d, r = P.survey_pair(2000, 8000, 700., seed=9, weights=True)
            out.append(T._compute_3pt_stats_survey_like(
                be, kind, d, r, paramset=ps, degrees=deg, binning=b,
                form=form, idx_bin=idx))
    P.compare_results(out[0], out[1], kind, rtol=1e-10)


def test_fp32_error_bound(gpu_backend):
    """FP32 fields (assignment, FFT, Y_lm products) with FP64 sums vs the
    FP64 engine path: max |B32 - B64| / max |B64| per multipole < 1e-4 on a
    64^3 periodic box of 2e4 particles (measured values are reported by the
    benchmark; this bound guards regressions)."""
    from triumvirate_amd import threept as T
    from triumvirate_amd.dataobjs import Binning
    b = Binning('fourier', 'lin', bin_min=0.01, bin_max=0.21, num_bins=8)
    res = {}
    for prec in ('fp64', 'fp32'):
        ps = cf.make_test_paramset(ngrid={'x': 64, 'y': 64, 'z': 64},
                                   precision=prec)
        d = P.uniform_box_catalogue(20000, 1000., seed=3)
        res[prec] = T._compute_3pt_stats_sim_like(
            gpu_backend, 'bispec', d, paramset=ps, degrees=(0, 0, 0),
            binning=b, form='full')
    for key in ('bk_raw', 'bk_shot'):
        a, r = res['fp32'][key], res['fp64'][key]
        err = np.max(np.abs(a - r)) / np.max(np.abs(r))
        print(f"fp32 {key} rel err {err:.3e}", flush=True)
        assert err < 1e-4
    assert np.array_equal(res['fp32']['nmodes_1'], res['fp64']['nmodes_1'])


def test_errors_map_to_reference_exceptions(gpu_backend):
    """Vanishing 3-j couplings raise InvalidParameterError; all-zero nz
    raises InvalidDataError (threept.cpp:73-89, 120-131)."""
    from triumvirate_amd import threept as T
    from triumvirate_amd.parameters import InvalidParameterError
    ps = cf.make_test_paramset()
    with pytest.raises(InvalidParameterError):
        T._compute_3pt_stats_sim_like(
            gpu_backend, 'bispec', cf.make_test_data_catalogue(), paramset=ps,
            degrees=(1, 0, 0), binning=cf.binning_fourier(), form='diag')
    d = cf.make_test_data_catalogue()
    d['nz'] = 0.
    d2 = P.uniform_box_catalogue(100, 1000., seed=1)
    x, y, z, nz, ws, wc = d2.columns()
    from triumvirate_amd import _abi
    view = _abi.CatalogueView(x, y, z, np.zeros_like(nz), ws, wc)
    with pytest.raises(T.InvalidDataError):
        gpu_backend.norm_particles(view, 1.)


@pytest.mark.parametrize('geo', ['lpp', 'gpp'])
def test_multipole_call_equals_single_calls(gpu_backend, oracle_backend, geo):
    """compute_bispec_multipoles (shared fields and shells) gives the same
    numbers as separate single-multipole calls and as the oracle."""
    from triumvirate_amd import threept as T
    from triumvirate_amd.dataobjs import Binning
    print(f"multi {geo}", flush=True)
    ps = cf.make_test_paramset(assignment='pcs',
                               ngrid={'x': 32, 'y': 32, 'z': 32},
                               boxsize={'x': 800., 'y': 800., 'z': 800.})
    b = Binning('fourier', 'lin', bin_min=0.01, bin_max=0.11, num_bins=5)
    mps = [((0, 0, 0), 'full'), ((2, 0, 2), 'full'), ((0, 0, 0), 'row', 2)]

    def run(be, deg=None, form=None, idx=None, multipoles=None):
        if geo == 'gpp':
            d = P.uniform_box_catalogue(3000, 800., seed=5)
            return T._compute_3pt_stats_sim_like(
                be, 'bispec', d, paramset=ps, degrees=deg or (0, 0, 0),
                binning=b, form=form or 'full', idx_bin=idx,
                multipoles=multipoles)
        d, r = P.survey_pair(2000, 8000, 700., seed=9, weights=True)
        return T._compute_3pt_stats_survey_like(
            be, 'bispec', d, r, paramset=ps, degrees=deg or (0, 0, 0),
            binning=b, form=form or 'full', idx_bin=idx, multipoles=multipoles)

    multi = run(gpu_backend, multipoles=mps)
    for mp, got in zip(mps, multi):
        single = run(gpu_backend, mp[0], mp[1], mp[2] if len(mp) > 2 else None)
        single.pop('_norm_factors', None)
        P.compare_results(got, single, 'bispec', rtol=1e-12)
        ref = run(oracle_backend, mp[0], mp[1], mp[2] if len(mp) > 2 else None)
        P.compare_results(got, ref, 'bispec', rtol=1e-10)


@pytest.mark.parametrize('weights', ['uniform', 'random-one-off'])
def test_data_fields_assigned_ahead_of_randoms(gpu_backend, oracle_backend, weights):
    """With uniform data weights, survey calls assign the data side of every
    (L, M) field while the random catalogue is still being uploaded; the
    estimator uses those fields when the random weights are uniform too
    ('uniform') and assigns afresh when they are not ('random-one-off').
    Multipole call (B_000 full, B_202 full, B_000 row) and single calls,
    GPU vs oracle, 1e-10."""
    from triumvirate_amd import threept as T
    from triumvirate_amd.dataobjs import Binning
    print(f"data fields ahead of randoms, weights={weights}", flush=True)
    ps = cf.make_test_paramset(assignment='pcs',
                               ngrid={'x': 32, 'y': 32, 'z': 32},
                               boxsize={'x': 800., 'y': 800., 'z': 800.})
    b = Binning('fourier', 'lin', bin_min=0.01, bin_max=0.11, num_bins=5)
    mps = [((0, 0, 0), 'full'), ((2, 0, 2), 'full'), ((0, 0, 0), 'row', 2)]

    def run(be, deg=None, form=None, idx=None, multipoles=None):
        d, r = P.survey_pair(2000, 8000, 700., seed=9, weights=weights)
        return T._compute_3pt_stats_survey_like(
            be, 'bispec', d, r, paramset=ps, degrees=deg or (0, 0, 0),
            binning=b, form=form or 'full', idx_bin=idx, multipoles=multipoles)

    multi = run(gpu_backend, multipoles=mps)
    for mp, got in zip(mps, multi):
        args = (mp[0], mp[1], mp[2] if len(mp) > 2 else None)
        ref = run(oracle_backend, *args)
        P.compare_results(got, ref, 'bispec', rtol=1e-10)
        single = run(gpu_backend, *args)
        single.pop('_norm_factors', None)
        P.compare_results(single, ref, 'bispec', rtol=1e-10)


@pytest.mark.parametrize('kind,deg', [('bispec', (0, 0, 0)), ('bispec', (2, 0, 2)),
                                      ('3pcf', (1, 1, 0))])
def test_device_lines_of_sight_equal_host_arrays(gpu_backend, kind, deg):
    """Survey drivers without explicit lines of sight form them on the
    device from the observer position (the offset the alignment applied);
    with the reference's host n x 3 arrays the estimates agree to round-off
    (the device takes positions minus the offset, one rounding apart)."""
    from triumvirate_amd import mocks
    from triumvirate_amd import threept as T
    from triumvirate_amd.catalogue import ParticleCatalogue
    from triumvirate_amd.dataobjs import Binning
    xd, yd, zd, nbar = mocks.lognormal_survey(20000, ngrid=32, seed=7)
    xr, yr, zr = mocks.uniform_shell_randoms(100000, seed=8)
    ps = cf.make_test_paramset(boxsize={'x': 3200., 'y': 3200., 'z': 3200.},
                               ngrid={'x': 64, 'y': 64, 'z': 64},
                               assignment='pcs')
    b = (cf.binning_fourier() if kind == 'bispec' else
         Binning('config', 'lin', bin_min=100., bin_max=700., num_bins=4))
    out = []
    for explicit in (False, True):
        d = ParticleCatalogue(xd, yd, zd, nz=nbar)
        r = ParticleCatalogue(xr, yr, zr, nz=nbar)
        kw = {}
        if explicit:
            kw = dict(los_data=d.compute_los(), los_rand=r.compute_los())
        print(f"device LOS {kind} {deg} explicit={explicit}", flush=True)
        out.append(T._compute_3pt_stats_survey_like(
            gpu_backend, kind, d, r, paramset=ps, degrees=deg, binning=b,
            form='diag', **kw))
    raw = 'bk_raw' if kind == 'bispec' else 'zeta_raw'
    shot = 'bk_shot' if kind == 'bispec' else 'zeta_shot'
    for key in (raw, shot):
        scale = np.max(np.abs(out[1][key]))
        assert np.max(np.abs(out[0][key] - out[1][key])) <= 1e-12 * scale, key


def _multi_call(eng, dview, rview, observer, plist, kbv, handles=None):
    from triumvirate_amd import _abi
    res = [_abi.ResultBuffers(_abi.result_dim(p.shape, kbv.struct.num_bins, 0), 'bispec')
           for p in plist]
    arr = (_abi.Params * len(plist))(*plist)
    outs = (_abi.BispecResult * len(plist))(*[r.struct for r in res])
    norms = np.ones(len(plist))
    if handles is None:
        st = eng.trvg_compute_bispec_multi_observer(
            C.byref(dview.struct), C.byref(rview.struct), _abi.dptr(observer), arr,
            len(plist), C.byref(kbv.struct), _abi.dptr(norms), outs)
    else:
        st = eng.trvg_compute_bispec_multi_dev(handles[0], handles[1], arr, len(plist),
                                               C.byref(kbv.struct), _abi.dptr(norms), outs)
    assert st == 0, eng.trvg_last_error()
    for r, o in zip(res, outs):
        r.struct.dim = o.dim
    return [r.as_dict() for r in res]


@pytest.mark.parametrize('outlier', [False, True])
def test_fp32_fixed_point_upload(eng, outlier):
    """FP32 survey calls on host catalogues send positions as 32-bit fixed
    point over [-L, 2L) (step 3L / 2^32): the estimates agree with the same
    FP32 call on device-resident FP64 positions within 2e-6 of the column
    maximum.  A chunk holding a position outside that range (here one far
    particle in each catalogue) crosses in FP64, and the call then equals
    the device-resident one up to the order of atomic sums (1e-12)."""
    from triumvirate_amd import _abi, mocks
    from triumvirate_amd.dataobjs import Binning
    print(f"fixed-point upload outlier={outlier}", flush=True)
    xd, yd, zd, _ = mocks.lognormal_survey(20000, ngrid=32, seed=7)
    xr, yr, zr = mocks.uniform_shell_randoms(100000, seed=8)
    L = 3200.
    # Box frame: shift every catalogue by the same offset.
    shift = L / 2.
    cols_d = [xd + shift, yd + shift, zd + shift]
    cols_r = [xr + shift, yr + shift, zr + shift]
    if outlier:
        for c in cols_d + cols_r:
            c[5] = 3.5 * L
    observer = np.full(3, shift)
    dview = _abi.CatalogueView(*cols_d)
    rview = _abi.CatalogueView(*cols_r)
    kb = Binning('fourier', 'lin', bin_min=0.005, bin_max=0.105, num_bins=10)
    kbv = _abi.BinningView(kb.bin_edges, kb.bin_centres)
    plist = []
    for deg, shape in (((0, 0, 0), 'diag'), ((2, 0, 2), 'full')):
        p = _abi.Params()
        for i in range(3):
            p.ngrid[i] = 64
            p.boxsize[i] = L
        p.assignment_order = 4
        p.ell1, p.ell2, p.ELL = deg
        p.shape = _abi.SHAPE_CODES[shape]
        p.precision = 32
        p.device = 0
        plist.append(p)
    host = _multi_call(eng, dview, rview, observer, plist, kbv)
    hs = []
    try:
        for view in (dview, rview):
            h = C.c_void_p()
            assert eng.trvg_catalogue_upload_observer(
                C.byref(view.struct), _abi.dptr(observer), 0, C.byref(h)) == 0, \
                eng.trvg_last_error()
            hs.append(h)
        dev = _multi_call(eng, dview, rview, observer, plist, kbv, handles=hs)
    finally:
        for h in hs:
            eng.trvg_catalogue_free(h)
    tol = 1e-12 if outlier else 2e-6
    for a, b in zip(host, dev):
        assert np.array_equal(a['nmodes_1'], b['nmodes_1'])
        for key in ('bk_raw', 'bk_shot'):
            scale = np.max(np.abs(b[key]))
            err = np.max(np.abs(a[key] - b[key])) / scale
            assert err < tol, (key, err)


@pytest.mark.parametrize('deg', [(0, 0, 0), (2, 0, 2)])
def test_one_off_weights_match_oracle(gpu_backend, oracle_backend, deg):
    """Catalogues whose weights are all 1 but for their last particle: the
    upload writes the weight column only from the first differing weight of
    each thread's range and fills the uniform ranges afterwards (GPU vs
    oracle, 1e-10)."""
    from triumvirate_amd import threept as T
    from triumvirate_amd.dataobjs import Binning
    print(f"one-off weights {deg}", flush=True)
    ps = cf.make_test_paramset(ngrid={'x': 32, 'y': 32, 'z': 32},
                               boxsize={'x': 800., 'y': 800., 'z': 800.})
    b = Binning('fourier', 'lin', bin_min=0.01, bin_max=0.11, num_bins=5)
    out = []
    for be in (gpu_backend, oracle_backend):
        d, r = P.survey_pair(2000, 8000, 700., seed=9, weights='one-off')
        out.append(T._compute_3pt_stats_survey_like(
            be, 'bispec', d, r, paramset=ps, degrees=deg, binning=b, form='full'))
    P.compare_results(out[0], out[1], 'bispec', rtol=1e-10)
