"""Committed golden vectors (tests/golden/oracle_golden.npz, written by
tests/golden/make_golden.py with the oracle's reference schedule) for
reduced, seeded versions of the benchmark configurations."""

import pytest

import conftest as cf
import golden_cases as G
import parity as P

GOLDEN = cf.GOLDEN / 'oracle_golden.npz'


@pytest.mark.parametrize('name', G.CASES)
def test_oracle_dedup_schedule_matches_golden(name):
    """The deduplicated oracle schedule reproduces the reference-schedule
    golden vectors (and the seeded inputs regenerate identically)."""
    import oracle
    res = G.run_case(oracle.backend(schedule=1), name)
    P.compare_results(res, G.load_golden(GOLDEN, name), G.kind_of(name),
                      rtol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize('name', G.CASES)
def test_engine_fp64_matches_golden(name):
    """Engine FP64 vs golden: counts bit-exact, values within 1e-10."""
    from triumvirate_amd import _lib
    from triumvirate_amd.threept import CLibBackend
    print(f"golden fp64 {name}", flush=True)
    res = G.run_case(CLibBackend(_lib.load()), name)
    P.compare_results(res, G.load_golden(GOLDEN, name), G.kind_of(name),
                      rtol=1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize('name', G.CASES)
def test_engine_fp32_error_vs_golden(name):
    """Engine FP32 fields (FP64 sums and shot-noise path) vs golden: counts
    bit-exact, NaN (empty-shell) entries in the same places, max |d| / max |B|
    < 1e-5 over the finite entries of the raw and shot parts."""
    import numpy as np
    from triumvirate_amd import _lib
    from triumvirate_amd.threept import CLibBackend
    print(f"golden fp32 {name}", flush=True)
    res = G.run_case(CLibBackend(_lib.load()), name, precision='fp32')
    ref = G.load_golden(GOLDEN, name)
    cols, raw, shot = P.keys(G.kind_of(name))
    for k in cols:
        if k.startswith(('nmodes', 'npairs')):
            assert np.array_equal(res[k], ref[k])
    for key in (raw, shot):
        nan = np.isnan(ref[key])
        assert np.array_equal(np.isnan(res[key]), nan)
        a, b = res[key][~nan], ref[key][~nan]
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        print(f"  {key} rel err {err:.2e}", flush=True)
        assert err < 1e-5
