"""Host-side logic mirroring the reference's Python/C++ drivers: parameter
validation, binning arithmetic, catalogue geometry, and the C ABI exports.
No GPU needed."""

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import conftest as cf

ROOT = Path(__file__).resolve().parent.parent


def test_form_full_becomes_triu_when_degrees_equal():
    """parameters.cpp:845-849."""
    ps = cf.make_test_paramset(form='full',
                               degrees={'ell1': 0, 'ell2': 0, 'ELL': 0})
    assert ps['shape'] == 'triu'
    ps = cf.make_test_paramset(form='full',
                               degrees={'ell1': 2, 'ell2': 0, 'ELL': 2})
    assert ps['shape'] == 'full'
    ps = cf.make_test_paramset(form='row', idx_bin=1,
                               degrees={'ell1': 0, 'ell2': 0, 'ELL': 0})
    assert ps['shape'] == 'row'


def test_interlacing_is_switched_off_for_three_point():
    """parameters.cpp:1240-1249."""
    ps = cf.make_test_paramset(interlace=True, statistic_type='bispec')
    assert ps['interlace'] is False
    ps = cf.make_test_paramset(interlace='on', statistic_type='powspec')
    assert ps['interlace'] is True


@pytest.mark.parametrize('scheme,order', [('ngp', 1), ('cic', 2), ('tsc', 3),
                                          ('pcs', 4)])
def test_assignment_order(scheme, order):
    assert cf.make_test_paramset(assignment=scheme)['assignment_order'] == order


def test_invalid_parameters_raise():
    from triumvirate_amd.parameters import InvalidParameterError
    with pytest.raises(InvalidParameterError):
        cf.make_test_paramset(form='square')
    with pytest.raises(InvalidParameterError):
        cf.make_test_paramset(assignment='sph')
    with pytest.raises(InvalidParameterError):
        cf.make_test_paramset(form='row', idx_bin=4, num_bins=4)


def test_binning_matches_reference_fixture_centres():
    """dataobjs.cpp:134-160 linear edges: bin_min + dbin * i."""
    b = cf.binning_fourier()
    ext = cf.load_ref_stats('bk000_diag_gpp.txt')
    assert np.array_equal(np.array(b.bin_centres), ext[0]) or \
        np.allclose(b.bin_centres, ext[0], rtol=1e-15)
    assert b.bin_edges[-1] == 0.105
    assert len(b.bin_edges) == 5
    r = cf.binning_config()
    ext = cf.load_ref_stats('zeta000_diag_gpp.txt')
    assert np.allclose(r.bin_centres, ext[0], rtol=1e-15)


def test_periodise_matches_fixture_extents():
    """Fixture header 'Catalogue particle extents' after periodise
    (catalogue.py:648-676)."""
    d = cf.make_test_data_catalogue()
    d.periodise([1000.] * 3)
    got = [d.bounds[a] for a in 'xyz']
    want = [(450.000, 550.000), (456.699, 543.301), (500.000, 500.000)]
    for (lo, hi), (wl, wh) in zip(got, want):
        assert round(lo, 3) == wl and round(hi, 3) == wh


def test_centre_matches_fixture_extents():
    """lpp fixture headers: data extents after centring on the randoms
    (catalogue.py:490-545)."""
    d, r = cf.make_test_data_catalogue(), cf.make_test_rand_catalogue()
    d.centre([1000.] * 3, catalogue_ref=r)
    want_d = [(450.002, 550.002), (471.122, 557.724), (500.017, 500.017)]
    want_r = [(0.015, 999.985), (0.017, 999.983), (0.031, 999.969)]
    for cat, want in ((d, want_d), (r, want_r)):
        for a, (wl, wh) in zip('xyz', want):
            lo, hi = cat.bounds[a]
            assert round(lo, 3) == wl and round(hi, 3) == wh


def test_lines_of_sight_are_unit_vectors():
    d = cf.make_test_data_catalogue()
    los = d.compute_los()
    assert np.allclose(np.linalg.norm(los, axis=1), 1.)
    from triumvirate_amd.catalogue import ParticleCatalogue
    z = ParticleCatalogue([0.], [0.], [0.], nz=1.)
    assert np.array_equal(z.compute_los(), np.zeros((1, 3)))


def _declared_functions():
    text = (ROOT / 'include' / 'trv_gpu.h').read_text()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(trvg_[a-z0-9_]+)\s*\(', text)))


def test_engine_library_exports_every_declared_symbol():
    """libtrvgpu.so loads without a GPU and exports each function of
    include/trv_gpu.h (no compute calls)."""
    from triumvirate_amd import _lib
    lib = _lib.load()
    names = _declared_functions()
    assert len(names) >= 16
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.trvg_abi_version() == 1
    n = C.c_int(-1)
    assert lib.trvg_device_count(C.byref(n)) == 0


def test_engine_reports_missing_device_on_cpu():
    """Without a GPU every compute entry point fails loudly (no fallback)."""
    from triumvirate_amd import _lib
    from triumvirate_amd.threept import CLibBackend
    from triumvirate_amd import _abi
    lib = _lib.load()
    n = C.c_int(0)
    lib.trvg_device_count(C.byref(n))
    if n.value > 0:
        pytest.skip("a HIP device is visible")
    be = CLibBackend(lib)
    d = cf.make_test_data_catalogue()
    x, y, z, nz, ws, wc = d.columns()
    view = _abi.CatalogueView(x, y, z, nz, ws, wc)
    from triumvirate_amd._lib import EngineError
    p = _abi.Params()
    for i in range(3):
        p.ngrid[i] = 16
        p.boxsize[i] = 100.
    p.assignment_order = 3
    p.precision = 64
    with pytest.raises(EngineError):
        be.norm_mesh(view, p, 1.)


def test_result_dims():
    from triumvirate_amd import _abi
    from triumvirate_amd import _lib
    lib = _lib.load()
    for shape, nb, idx, want in [(0, 20, 0, 20), (1, 20, -3, 17), (2, 20, 4, 20),
                                 (3, 20, 0, 400), (4, 20, 0, 210)]:
        d = C.c_int(0)
        assert lib.trvg_result_dim(shape, nb, idx, C.byref(d)) == 0
        assert d.value == want == _abi.result_dim(shape, nb, idx)


def test_mesh_mixed_rejected_for_three_point():
    """parameters.cpp:870-882: 'mesh-mixed' applies to two-point statistics
    only; three-point statistics raise instead of normalising by zero."""
    from triumvirate_amd.parameters import InvalidParameterError
    for st in ('bispec', '3pcf'):
        with pytest.raises(InvalidParameterError, match='mesh-mixed'):
            cf.make_test_paramset(norm_convention='mesh-mixed',
                                  statistic_type=st)
    ps = cf.make_test_paramset(norm_convention='mesh-mixed',
                               statistic_type='powspec')
    assert ps['norm_convention'] == 'mesh-mixed'
    from triumvirate_amd import threept as T
    with pytest.raises(InvalidParameterError):
        T._select_norm({'norm_convention': 'mesh-mixed'}, 1., 2.)


def test_differentiable_inputs_rejected_before_the_engine(monkeypatch):
    """Host pointers, wrong dtypes or shapes must raise ValueError before any
    engine call (they would be out-of-bounds device accesses)."""
    torch = pytest.importorskip('torch')
    from triumvirate_amd import _lib, autodiff
    from triumvirate_amd.dataobjs import Binning

    def _no_engine(*a, **k):
        raise AssertionError('engine called')
    monkeypatch.setattr(_lib, 'load', _no_engine)
    b = Binning('fourier', 'lin', bin_min=0.01, bin_max=0.05, num_bins=2)
    B = autodiff.DifferentiableBispectrum(100., 16, b, precision='fp64')
    pos = torch.zeros(10, 3, dtype=torch.float64)
    with pytest.raises(ValueError, match='CUDA'):
        B(pos)
    with pytest.raises(ValueError, match='CUDA'):
        B.from_mesh(torch.zeros(16, 16, 16, dtype=torch.float64))
    with pytest.raises(ValueError):
        B(pos.numpy())
    with pytest.raises(ValueError):
        autodiff._check_mesh(torch.zeros(16, 16, 8), B.cfg)
    # The dtype / shape / weights rules themselves (independent of device).
    with pytest.raises(ValueError):
        autodiff._check_weights(pos[:, 0], pos)
    with pytest.raises(ValueError):
        autodiff._check_weights(np.ones(10), pos)


def test_unit_weight_columns_are_not_passed_to_the_engine():
    """Default (unit) weight columns are read-only and go to the engine as
    absent columns; weights set later are passed."""
    import numpy as np
    from triumvirate_amd.catalogue import ParticleCatalogue
    x = np.linspace(0., 1., 8)
    c = ParticleCatalogue(x, x, x, nz=1.)
    cols = c.engine_columns()
    assert cols[4] is None and cols[5] is None
    np.testing.assert_array_equal(c['ws'], np.ones(8))
    with pytest.raises(ValueError):
        c['ws'][0] = 2.
    c['ws'] = np.full(8, 2.)
    assert c.engine_columns()[4] is not None and c.engine_columns()[5] is None
    assert c.wstotal == 16.
    c['ws'] = 1.
    assert c.engine_columns()[4] is None
    d = ParticleCatalogue(x, x, x, nz=1., wc=np.ones(8))
    assert d.engine_columns()[5] is not None   # arrays are passed as given
