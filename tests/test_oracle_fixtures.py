"""The CPU oracle against the reference's own end-to-end fixture files
(reference tests/test_threept.py:16-332, fixtures
tests/test_input/stats/{bk,zeta}*.txt).

Both oracle schedules are checked: 0 follows the reference's loop structure
(threept.cpp), 1 deduplicates shell transforms as the engine does.
"""

import pytest

import parity as P


@pytest.mark.parametrize('schedule', [0, 1])
@pytest.mark.parametrize('case', P.FIXTURE_CASES,
                         ids=[P.fixture_name(*c) for c in P.FIXTURE_CASES])
def test_oracle_reproduces_reference_fixture(case, schedule):
    import oracle
    res = P.run_fixture_case(oracle.backend(schedule=schedule), *case)
    P.check_against_fixture(res, case[0], P.fixture_name(*case))


def test_normalisations_match_fixture_headers():
    """Headers of bk000_diag_{gpp,lpp}.txt record both normalisation
    alternatives: 3.703703704e+16 (particle) for both, 8.142524244e+07
    (mesh, gpp) and 4.417104020e+15 (mesh, lpp)."""
    import oracle
    be = oracle.backend(schedule=1)
    r = P.run_fixture_case(be, 'bispec', (0, 0, 0), 'diag', None, 'gpp')
    part, mesh = r['_norm_factors']
    assert abs(part / 3.703703704e+16 - 1) < 1e-9
    assert abs(mesh / 8.142524244e+07 - 1) < 1e-9
    r = P.run_fixture_case(be, 'bispec', (0, 0, 0), 'diag', None, 'lpp')
    part, mesh = r['_norm_factors']
    assert abs(part / 3.703703704e+16 - 1) < 1e-9
    assert abs(mesh / 4.417104020e+15 - 1) < 1e-9


def test_fixture_mode_counts_sum():
    """Total modes over the four default bins equals the reference's
    test_fieldmesh count (19548 = 460 + 2340 + 5908 + 10840)."""
    import conftest as cf
    ext = cf.load_ref_stats('bk000_diag_gpp.txt')
    assert int(ext[2].sum()) == 19548
