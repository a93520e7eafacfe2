"""GPU parity of the x-first shell transforms at sizes beyond the 32^3
oracle cases.

Shell fields are transformed x pass first (with the shell loader, pruned on
k_y^2 + k_z^2), then y (pruned on k_z), then along the contiguous z lines
(reading only the k_z rows the y pass wrote).  These cases run at 64^3
(line-blocked fast paths of all three passes) and at 32 x 64 x 128 (unequal
axis lengths, equal cells), through the pair-kernel variants: complex x
half-real (survey B_202 terms with M != 0), all-real (M = 0 terms and
periodic-box B_202), half-real x complex (B_022), complex x complex
(B_110, box B_220).

Tolerances: FP64 vs the CPU oracle 1e-10 relative to max |B| of the
multipole; FP32 vs the FP64 engine 1e-4 (as the FP32 bound elsewhere).
The FP32 cases with 256 and 512 cells along z run the register-resident
z pass (the FP64 engine and the oracle use the LDS-staged one); those with
256 cells along y or x check the FP32 strided passes at that length.
"""

import pytest

import conftest as cf
import parity as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def gpu_backend():
    from triumvirate_amd import _lib
    from triumvirate_amd.threept import CLibBackend
    return CLibBackend(_lib.load(), 'trvg_')


@pytest.fixture(scope='module')
def oracle_backend():
    import oracle
    return oracle.backend(schedule=1)


CASES = [
    ((2, 0, 2), 'full', None, 'lpp', (64, 64, 64)),
    ((2, 0, 2), 'row', 2, 'lpp', (64, 64, 64)),
    ((0, 2, 2), 'full', None, 'lpp', (64, 64, 64)),
    ((1, 1, 0), 'full', None, 'lpp', (64, 64, 64)),
    ((2, 0, 2), 'full', None, 'gpp', (64, 64, 64)),
    ((2, 2, 0), 'full', None, 'gpp', (64, 64, 64)),
    ((2, 0, 2), 'full', None, 'lpp', (32, 64, 128)),
]


def _run(be, deg, form, idx, geo, ng, precision=None, weights=True):
    from triumvirate_amd import threept as T
    from triumvirate_amd.dataobjs import Binning
    cell = 12.5
    ps = cf.make_test_paramset(
        assignment='tsc', ngrid={'x': ng[0], 'y': ng[1], 'z': ng[2]},
        boxsize={'x': cell * ng[0], 'y': cell * ng[1], 'z': cell * ng[2]})
    b = Binning('fourier', 'lin', bin_min=0.01, bin_max=0.16, num_bins=6)
    kw = {} if precision is None else {'precision': precision}
    if geo == 'gpp':
        d = P.uniform_box_catalogue(4000, cell * ng[0], seed=5, weights=False)
        return T._compute_3pt_stats_sim_like(be, 'bispec', d, paramset=ps, degrees=deg,
                                             binning=b, form=form, idx_bin=idx, **kw)
    d, r = P.survey_pair(2000, 8000, 0.9 * cell * min(ng), seed=9, weights=weights)
    return T._compute_3pt_stats_survey_like(be, 'bispec', d, r, paramset=ps, degrees=deg,
                                            binning=b, form=form, idx_bin=idx, **kw)


@pytest.mark.parametrize('case', CASES, ids=['-'.join(map(str, c)) for c in CASES])
def test_xfirst_matches_oracle(gpu_backend, oracle_backend, case):
    print(f"xfirst-case {case}", flush=True)
    a = _run(gpu_backend, *case)
    b = _run(oracle_backend, *case)
    P.compare_results(a, b, 'bispec', rtol=1e-10)


def test_xfirst_fp32_bound(gpu_backend):
    import numpy as np
    case = ((2, 0, 2), 'full', None, 'lpp', (64, 64, 64))
    print("xfirst fp32 bound", flush=True)
    a = _run(gpu_backend, *case, precision='fp32')
    b = _run(gpu_backend, *case, precision='fp64')
    for key in ('bk_raw', 'bk_shot'):
        err = np.max(np.abs(a[key] - b[key])) / np.max(np.abs(b[key]))
        print(f"  {key} {err:.2e}", flush=True)
        assert err < 1e-4


@pytest.mark.parametrize('ng', [(32, 32, 256), (16, 16, 512), (16, 256, 16),
                                (256, 16, 16)])
@pytest.mark.parametrize('geo', ['lpp', 'gpp'])
def test_fp32_long_axes(gpu_backend, ng, geo):
    import numpy as np
    case = ((2, 0, 2), 'full', None, geo, ng)
    print(f"fp32 long axes {geo} {ng}", flush=True)
    a = _run(gpu_backend, *case, precision='fp32')
    b = _run(gpu_backend, *case, precision='fp64')
    for key in ('bk_raw', 'bk_shot'):
        err = np.max(np.abs(a[key] - b[key])) / np.max(np.abs(b[key]))
        print(f"  {key} {err:.2e}", flush=True)
        assert err < 1e-4


@pytest.mark.parametrize('deg', [(0, 0, 0), (2, 0, 2), (0, 2, 2)])
def test_uniform_weight_survey_fields(gpu_backend, oracle_backend, deg):
    """Catalogues whose particles share one weight (per catalogue) take the
    engine's single-accumulation path for the survey pair (dn, N): M = 0
    and M != 0 terms against the oracle at 1e-10."""
    print(f"uniform weights {deg}", flush=True)
    case = (deg, 'full', None, 'lpp', (64, 64, 64))
    a = _run(gpu_backend, *case, weights='uniform')
    b = _run(oracle_backend, *case, weights='uniform')
    P.compare_results(a, b, 'bispec', rtol=1e-10)
