"""The C++ drop-in header (include/triumvirate_gpu.hpp) compiles and links
against libtrvgpu.so with the reference's trv:: names and signatures (CPU),
and reproduces the periodic-box fixture when run on a GPU."""

import subprocess
from pathlib import Path

import numpy as np
import pytest

import conftest as cf

ROOT = Path(__file__).resolve().parent.parent
LIBDIR = ROOT / 'triumvirate_amd'


def _build(tmp_path):
    exe = tmp_path / 'dropin_main'
    cmd = ['g++', '-std=c++17', '-O1', '-I', str(ROOT / 'include'),
           str(ROOT / 'tests' / 'cpp' / 'dropin_main.cpp'), '-o', str(exe),
           '-L', str(LIBDIR), '-ltrvgpu', f'-Wl,-rpath,{LIBDIR}']
    subprocess.run(cmd, check=True)
    return exe


def test_dropin_header_compiles_and_links(tmp_path):
    exe = _build(tmp_path)
    assert exe.exists()


@pytest.mark.gpu
def test_dropin_reproduces_fixture(tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([str(exe), str(cf.REF_STATS / 'test_data_catalogue.txt')],
                         check=True, capture_output=True, text=True,
                         timeout=120).stdout
    got = np.array([[float(v) for v in ln.split()] for ln in out.splitlines()]).T
    ext = cf.load_ref_stats('bk000_diag_gpp.txt')
    assert np.array_equal(got[2], ext[2]) and np.array_equal(got[5], ext[5])
    np.testing.assert_allclose(got[1], ext[1], rtol=1e-9)
    for col in (6, 8):
        err = np.max(np.abs(got[col] - ext[col])) / np.max(np.abs(ext[col]))
        assert err < 2e-9, (col, err)
