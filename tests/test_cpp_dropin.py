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


@pytest.mark.gpu
def test_dropin_powspec_reproduces_fixture(tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([str(exe), str(cf.REF_STATS / 'test_data_catalogue.txt'),
                          'powspec'], check=True, capture_output=True, text=True,
                         timeout=120).stdout
    got = np.array([[float(v) for v in ln.split()] for ln in out.splitlines()]).T
    ext = cf.load_ref_stats('pk0_gpp.txt')
    assert np.array_equal(got[2], ext[2])
    np.testing.assert_allclose(got[1], ext[1], rtol=1e-9)
    for col in (3, 5):
        err = np.max(np.abs(got[col] - ext[col])) / np.max(np.abs(ext[col]))
        assert err < 2e-9, (col, err)


def test_dropin_binning_schemes_match_python(tmp_path):
    """trv::Binning of the C++ drop-in (every scheme, the grid-scale padding
    of the ParameterSet constructor and the grid-scale default binning)
    equals the Python Binning, which restates src/dataobjs.cpp:134-250."""
    from triumvirate_amd.dataobjs import Binning
    exe = tmp_path / 'binning_main'
    subprocess.run(['g++', '-std=c++17', '-O1', '-I', str(ROOT / 'include'),
                    str(ROOT / 'tests' / 'cpp' / 'binning_main.cpp'), '-o',
                    str(exe)], check=True)
    lines = subprocess.run([str(exe)], check=True, capture_output=True,
                           text=True).stdout.splitlines()
    expect = []
    for sch in ('lin', 'log', 'linpad', 'logpad'):
        expect.append(Binning('fourier', sch, 0.01, 0.2, 12))
        expect.append(Binning('config', sch, 20., 180., 10))
    expect.append(Binning('fourier', 'linpad', 0., 0.2, 15, boxsize_max=1500.,
                          ngrid_min=64))
    assert len(lines) == len(expect) + 1
    for ln, b in zip(lines, expect):
        head, cen, wid = ln.split('|')
        vals = head.split()
        assert vals[0] == b.scheme
        assert np.array_equal(np.array(vals[1:], float), np.asarray(b.bin_edges))
        assert np.array_equal(np.array(cen.split(), float), np.asarray(b.bin_centres))
        assert np.array_equal(np.array(wid.split(), float), np.asarray(b.bin_widths))
    # Grid-scale default (dataobjs.cpp:80-104): 32 linear bins from 0 of
    # width L / N, up to L / 2 + half a width.
    head, cen, wid = lines[-1].split('|')
    edges = np.array(head.split()[1:], float)
    w = 1500. / 64
    assert len(edges) == 33 and edges[0] == 0.
    np.testing.assert_allclose(np.diff(edges), w * (750. + w / 2) / (32 * w),
                               rtol=1e-12)
