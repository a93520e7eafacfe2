"""The engine's three-point correlation function window
(trvg_compute_3pcf_window): the reference fixtures through the public
driver (FP64), the oracle on forms, degrees and wide-angle orders no fixture
covers (1e-10), and FP32 fields within 1e-5 of the FP64 window."""

import numpy as np
import pytest

from test_oracle_window3pt import WINDOW_FIXTURES, check_window_fixture, run_window

pytestmark = pytest.mark.gpu


def _engine():
    from triumvirate_amd import _lib
    from triumvirate_amd.threept import CLibBackend
    return CLibBackend(_lib.load(), 'trvg_')


@pytest.mark.parametrize('degrees,form,idx_bin,name', WINDOW_FIXTURES)
def test_engine_window_matches_reference_fixture(degrees, form, idx_bin, name):
    print(f"3pcf window fixture {name}", flush=True)
    check_window_fixture(run_window(_engine(), degrees, form, idx_bin), name)


def _rel(a, b):
    scale = np.max(np.abs(b))
    if scale == 0.:          # e.g. the shot noise of an off-diagonal form
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / scale)


@pytest.mark.parametrize('degrees,form,idx_bin,wa', [
    ((0, 0, 0), 'full', None, None),
    ((1, 1, 0), 'full', None, (1, 0)),
    ((2, 0, 2), 'full', None, (0, 2)),
    ((0, 0, 0), 'off-diag', 1, (1, 1)),
])
def test_engine_window_matches_oracle(degrees, form, idx_bin, wa):
    import oracle
    print(f"3pcf window vs oracle {degrees} {form} {idx_bin} wa={wa}", flush=True)
    got = run_window(_engine(), degrees, form, idx_bin, wa_orders=wa)
    ref = run_window(oracle.backend(schedule=1), degrees, form, idx_bin, wa_orders=wa)
    assert np.array_equal(got['npairs_1'], ref['npairs_1'])
    assert np.array_equal(got['npairs_2'], ref['npairs_2'])
    np.testing.assert_allclose(got['r1_eff'], ref['r1_eff'], rtol=1e-12)
    for key in ('zeta_raw', 'zeta_shot'):
        err = _rel(got[key], ref[key])
        print(f"  {key} {err:.2e}", flush=True)
        assert err < 1e-10, (key, err)


def test_engine_window_fp32():
    print("3pcf window fp32", flush=True)
    a = run_window(_engine(), (2, 0, 2), 'full', None, wa_orders=(1, 1),
                   precision='fp32')
    b = run_window(_engine(), (2, 0, 2), 'full', None, wa_orders=(1, 1))
    assert np.array_equal(a['npairs_1'], b['npairs_1'])
    for key in ('zeta_raw', 'zeta_shot'):
        err = _rel(a[key], b[key])
        print(f"  {key} {err:.2e}", flush=True)
        assert err < 1e-5, (key, err)
