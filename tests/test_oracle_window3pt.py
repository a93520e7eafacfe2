"""Pins of the oracle's three-point correlation function window
(threept.cpp:2621-3077): the reference fixtures zetaw000_diag,
zetaw202_diag and zetaw000_row0 (random catalogue alone, centred and
periodised, particle normalisation with alpha = 1) and their normalisation
headers; and, for the wide-angle kernel (field.cpp:1727-1762), which no
fixture covers, two checks of what the mathematics fixes: orders (0, 0)
leave every output unchanged, and with a single (L, M) = (0, 0) term the
raw window is linear in the kernel, so orders (i, j) and (j, i) agree.
"""

import re

import numpy as np
import pytest

import conftest as cf

WINDOW_FIXTURES = [
    ((0, 0, 0), 'diag', None, 'zetaw000_diag.txt'),
    ((2, 0, 2), 'diag', None, 'zetaw202_diag.txt'),
    ((0, 0, 0), 'row', 0, 'zetaw000_row0.txt'),
]


def run_window(backend, degrees, form, idx_bin, wa_orders=None, precision=None):
    from triumvirate_amd import threept as T
    ps = cf.make_test_paramset()
    return T._compute_3pcf_window(
        backend, cf.make_test_rand_catalogue(), paramset=ps, degrees=degrees,
        wa_orders=wa_orders, binning=cf.binning_config(), form=form,
        idx_bin=idx_bin, precision=precision)


def check_window_fixture(res, name, rtol=2e-9):
    ext = cf.load_ref_stats(name)
    np.testing.assert_allclose(res['r1_bin'], ext[0], rtol=1e-9)
    np.testing.assert_allclose(res['r1_eff'], ext[1], rtol=1e-9)
    assert np.array_equal(res['npairs_1'], ext[2].astype(int))
    np.testing.assert_allclose(res['r2_bin'], ext[3], rtol=1e-9)
    np.testing.assert_allclose(res['r2_eff'], ext[4], rtol=1e-9)
    assert np.array_equal(res['npairs_2'], ext[5].astype(int))
    for col, key in ((6, 'zeta_raw'), (8, 'zeta_shot')):
        scale = np.max(np.abs(ext[col]))
        assert np.max(np.abs(res[key].real - ext[col])) <= rtol * scale, key
        assert np.max(np.abs(res[key].imag)) <= 1e-6 * scale, key
    txt = (cf.REF_STATS / name).read_text()
    m = re.search(r"alternatives: ([0-9.e+-]+) \(particle\), ([0-9.e+-]+) \(mesh\)", txt)
    np.testing.assert_allclose(res['_norm_factors'][0], float(m.group(1)), rtol=2e-9)
    np.testing.assert_allclose(res['_norm_factors'][1], float(m.group(2)), rtol=2e-9)


@pytest.mark.parametrize('degrees,form,idx_bin,name', WINDOW_FIXTURES)
def test_oracle_window_matches_reference_fixture(degrees, form, idx_bin, name):
    import oracle
    check_window_fixture(run_window(oracle.backend(schedule=0), degrees, form,
                                    idx_bin), name)


def test_oracle_wide_angle_kernel_properties():
    import oracle
    be = oracle.backend(schedule=1)
    base = run_window(be, (0, 0, 0), 'diag', None)
    zero = run_window(be, (0, 0, 0), 'diag', None, wa_orders=(0, 0))
    for key in ('zeta_raw', 'zeta_shot'):
        np.testing.assert_allclose(zero[key], base[key], rtol=1e-13, atol=0)
    a = run_window(be, (0, 0, 0), 'diag', None, wa_orders=(2, 1))
    b = run_window(be, (0, 0, 0), 'diag', None, wa_orders=(1, 2))
    np.testing.assert_allclose(a['zeta_raw'], b['zeta_raw'], rtol=1e-13, atol=0)
    # The kernel suppresses G at large |x|: the raw window changes.
    assert np.max(np.abs(a['zeta_raw'] - base['zeta_raw'])) > \
        1e-3 * np.max(np.abs(base['zeta_raw']))
    # The shot noise does not involve G.
    np.testing.assert_allclose(a['zeta_shot'], base['zeta_shot'], rtol=1e-13, atol=0)
