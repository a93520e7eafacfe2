#!/usr/bin/env python
"""Benchmark: bispectrum bin-pair evaluations per second at a 256^3 mesh.

Workload (one step): survey-like, local plane-parallel B_000 (form full ->
upper triangle, 210 pairs) + B_202 (full, 400 pairs) over 20 linear k-bins
in [0.005, 0.205] h/Mpc; 1e6 lognormal data points and 5e6 uniform randoms
in a quarter-sky shell r in [500, 1500] Mpc/h; box L = 3200 Mpc/h centred on
the randoms; 256^3 mesh, PCS assignment (interlacing off, as the reference
forces for three-point statistics).  Fields in FP32, sums in FP64.  Each
step is one engine call with the catalogues in host memory and the results
returned to host memory; the value is the median over the steps, beside
one cold call, FP64 calls, device-resident catalogues and the
differentiable B_000 forward + backward time.

Prints ONE JSON line with the metric, throughput, per-kernel roofline and a
CPU baseline (the oracle's reference-schedule per-pair loop on host cores).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--ngrid 256]
"""

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_inputs(nd, nr, ngen):
    from triumvirate_amd import mocks
    from triumvirate_amd.catalogue import ParticleCatalogue
    t = time.time()
    xd, yd, zd, nbar = mocks.lognormal_survey(nd, ngrid=ngen, seed=2)
    xr, yr, zr = mocks.uniform_shell_randoms(nr, seed=3)
    data = ParticleCatalogue(xd, yd, zd, nz=nbar)
    rand = ParticleCatalogue(xr, yr, zr, nz=nbar)
    los_d, los_r = data.compute_los(), rand.compute_los()
    data.centre([3200.] * 3, catalogue_ref=rand)
    log(f"inputs: {data.ntotal} data, {rand.ntotal} randoms "
        f"({time.time() - t:.1f} s)")
    return data, rand, los_d, los_r


def traffic_for(kernel_class):
    """HBM bytes per launch of a kernel class from the committed counter
    summary (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a plain copy
    kernel), or None when no summary covers it."""
    path = os.path.join(ROOT, 'profiles', 'hbm_traffic_bispec_b000_b202_256.md')
    if not os.path.exists(path):
        return None
    for ln in open(path):
        parts = ln.split()
        if len(parts) == 3 and parts[0] == 'traffic_bytes_per_launch' \
                and parts[1] == kernel_class:
            return float(parts[2])
    return None


def params_struct(ngrid, degrees, shape, prec):
    from triumvirate_amd import _abi
    p = _abi.Params()
    for i in range(3):
        p.ngrid[i] = ngrid
        p.boxsize[i] = 3200.
    p.assignment_order = 4
    p.interlace = 0
    p.ell1, p.ell2, p.ELL = degrees
    p.shape = _abi.SHAPE_CODES[shape]
    p.idx_bin = 0
    p.precision = prec
    p.verbose = 20
    p.device = 0
    return p


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--ngrid', type=int, default=256)
    ap.add_argument('--ndata', type=int, default=1_000_000)
    ap.add_argument('--nrand', type=int, default=5_000_000)
    ap.add_argument('--workload', choices=('bispec', 'grad', '3pcf', 'mocks'),
                    default='bispec',
                    help='bispec: B_000 + B_202 survey at 256^3 (default); '
                         'grad: differentiable B_000 fwd+bwd, 1e6 particles, '
                         '256^3 CIC; 3pcf: zeta_000 + zeta_110 survey at 512^3; '
                         'mocks: 256 periodic lognormal mocks, batched, 128^3')
    ap.add_argument('--nmocks', type=int, default=256)
    ap.add_argument('--batch', type=int, default=64)
    ap.add_argument('--skip-fp64', action='store_true',
                    help='skip the FP64 steps (profiler runs)')
    ap.add_argument('--skip-grad', action='store_true',
                    help='skip the differentiable fwd+bwd figure')
    ap.add_argument('--cpu-pairs', type=int, default=40,
                    help='bin pairs in the CPU-baseline sample (0: skip)')
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; multi-GPU figures are "
              "not produced (use --gpus 1).", file=sys.stderr)
        sys.exit(2)

    from triumvirate_amd import _abi, _lib
    from triumvirate_amd.dataobjs import Binning
    lib = _lib.load()
    ndev = C.c_int(0)
    lib.trvg_device_count(C.byref(ndev))
    if ndev.value < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)

    if args.workload == 'grad':
        print(json.dumps(bench_grad(args, lib)))
        return
    if args.workload == '3pcf':
        print(json.dumps(bench_3pcf(args, lib)))
        return
    if args.workload == 'mocks':
        print(json.dumps(bench_mocks(args, lib)))
        return
    print(json.dumps(bench_bispec(args, lib)))


def bench_bispec(args, lib):
    from triumvirate_amd import _abi
    from triumvirate_amd.dataobjs import Binning
    data, rand, los_d, los_r = make_inputs(args.ndata, args.nrand, args.ngrid)
    kb = Binning('fourier', 'lin', bin_min=0.005, bin_max=0.205, num_bins=20)
    kbv = _abi.BinningView(kb.bin_edges, kb.bin_centres)

    def check(st):
        if st != 0:
            raise RuntimeError(lib.trvg_last_error().decode())

    # Unit weights: the weight columns are not passed (absent = 1).
    dview = _abi.CatalogueView(*data.engine_columns())
    rview = _abi.CatalogueView(*rand.engine_columns())

    # Particle normalisation (reference default convention).
    alpha = data.wstotal / rand.wstotal
    norm = C.c_double()
    check(lib.trvg_calc_bispec_normalisation_from_particles(
        C.byref(rview.struct), alpha, C.byref(norm)))

    multipoles = [((0, 0, 0), 'triu'), ((2, 0, 2), 'full')]
    pairs_per_step = 210 + 400

    # Observer position in the box frame: the catalogues were centred by
    # an offset d after their lines of sight were taken (positions r - d).
    observer = -np.asarray(data._offset, dtype=np.float64)

    def step(prec, handles=None, los_arrays=False):
        """One engine call computing both multipoles on shared fields:
        host-resident catalogues in (the default; lines of sight formed on
        the device from the observer position, or with los_arrays passed
        as n x 3 host arrays), results out on the host; with handles,
        device-resident catalogues."""
        plist = [params_struct(args.ngrid, deg, shape, prec)
                 for deg, shape in multipoles]
        res = [_abi.ResultBuffers(_abi.result_dim(p.shape, 20, 0), 'bispec')
               for p in plist]
        arr = (_abi.Params * len(plist))(*plist)
        outs = (_abi.BispecResult * len(plist))(*[r.struct for r in res])
        norms = np.full(len(plist), norm.value)
        if handles is None and los_arrays:
            check(lib.trvg_compute_bispec_multi(
                C.byref(dview.struct), C.byref(rview.struct), _abi.dptr(los_d),
                _abi.dptr(los_r), arr, len(plist), C.byref(kbv.struct),
                _abi.dptr(norms), outs))
        elif handles is None:
            check(lib.trvg_compute_bispec_multi_observer(
                C.byref(dview.struct), C.byref(rview.struct), _abi.dptr(observer),
                arr, len(plist), C.byref(kbv.struct), _abi.dptr(norms), outs))
        else:
            check(lib.trvg_compute_bispec_multi_dev(
                handles[0], handles[1], arr, len(plist), C.byref(kbv.struct),
                _abi.dptr(norms), outs))
        out = {}
        for (deg, _), r, o in zip(multipoles, res, outs):
            r.struct.dim = o.dim
            out[deg] = r.as_dict()
        return out

    def timed(fn, n):
        ts = []
        for _ in range(n):
            t = time.perf_counter()
            out = fn()
            ts.append(time.perf_counter() - t)
        return ts, out

    # Cold run: the first engine call of the process (tables, splines, shell
    # plans, device pool and pinned staging all built inside it).
    t = time.perf_counter()
    step(32)
    t_cold = time.perf_counter() - t
    log(f"cold FP32 call {t_cold * 1e3:.1f} ms")
    for _ in range(args.warmup):
        step(32)

    # Breakdown step (outside the timed region): every kernel class timed.
    ncls = lib.trvg_profile_class_count()
    names = [lib.trvg_profile_class_name(kc).decode() for kc in range(ncls)]

    def read_profile(nsteps):
        out = {}
        for kc in range(ncls):
            ms, n, by = C.c_double(), C.c_longlong(), C.c_double()
            check(lib.trvg_profile_read(kc, C.byref(ms), C.byref(n), C.byref(by)))
            if n.value:
                out[names[kc]] = {
                    'class': kc,
                    'ms_per_step': ms.value / nsteps,
                    'launches_per_step': n.value / nsteps,
                    'avg_us': 1e3 * ms.value / n.value,
                    'bytes_per_launch': by.value / n.value,
                    'achieved_GBs': (by.value / n.value) /
                    (ms.value / n.value * 1e-3) / 1e9 if by.value > 0 else None,
                }
        return out

    check(lib.trvg_profile_only(-1))
    check(lib.trvg_profile_enable(1))
    t = time.perf_counter()
    step(32)
    ms_breakdown = (time.perf_counter() - t) * 1e3
    kernels = read_profile(1)
    check(lib.trvg_profile_enable(0))
    # The roofline kernel: the HBM-bound class with the most time per step.
    dom = max((k for k in kernels if kernels[k]['achieved_GBs']),
              key=lambda k: kernels[k]['ms_per_step'])

    # Timed region: K host-resident FP32 calls, each timed; HIP events only
    # around the roofline kernel's launches.
    log(f"timing {args.steps} FP32 calls, host-resident catalogues "
        f"(events on {dom}) ...")
    check(lib.trvg_profile_only(kernels[dom]['class']))
    check(lib.trvg_profile_enable(1))
    t32, r32 = timed(lambda: step(32), args.steps)
    kd = read_profile(args.steps)[dom]
    check(lib.trvg_profile_enable(0))
    check(lib.trvg_profile_only(-1))
    ms_per_step = float(np.median(t32)) * 1e3
    log(f"FP32 call median {ms_per_step:.2f} ms "
        f"(min {min(t32) * 1e3:.2f}, max {max(t32) * 1e3:.2f})")

    # FP64 (the drop-in precision), host-resident: one warm-up, median of >= 5.
    r64, t64 = None, None
    if not args.skip_fp64:
        step(64)
        t64, r64 = timed(lambda: step(64), max(5, args.steps))
        log(f"FP64 call median {np.median(t64) * 1e3:.2f} ms")

    # Host-resident with the lines of sight as n x 3 host arrays (the
    # reference's calling convention): secondary figure, and the check that
    # both forms agree.
    step(32, los_arrays=True)
    tlos, rlos = timed(lambda: step(32, los_arrays=True), args.steps)
    los_agree = max(
        float(np.max(np.abs(rlos[d][k] - r32[d][k])) / np.max(np.abs(r32[d][k])))
        for d, _ in multipoles for k in ('bk_raw', 'bk_shot'))
    log(f"host LOS arrays: FP32 call median {np.median(tlos) * 1e3:.2f} ms, "
        f"max rel diff vs device LOS {los_agree:.1e}")

    # Device-resident catalogues (secondary figure): uploads outside.
    t = time.perf_counter()
    handles = [_upload(lib, data, los_d), _upload(lib, rand, los_r)]
    t_up = time.perf_counter() - t
    step(32, handles)
    tdev, _ = timed(lambda: step(32, handles), args.steps)
    for h in handles:
        lib.trvg_catalogue_free(h)
    log(f"device-resident FP32 call median {np.median(tdev) * 1e3:.2f} ms, "
        f"upload of both catalogues {t_up * 1e3:.1f} ms")

    # Plain copy kernel of 1 GiB: the scale for hardware byte counters.
    cal_ms = C.c_double()
    check(lib.trvg_copy_calibration(1 << 30, C.byref(cal_ms)))

    # FP32 error against FP64, per multipole.
    errs = {}
    for deg, _ in (multipoles if r64 else []):
        tag = 'B' + ''.join(map(str, deg))
        e = {}
        for key in ('bk_raw', 'bk_shot'):
            a, b = r32[deg][key], r64[deg][key]
            e[key] = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        est32 = r32[deg]['bk_raw'] - r32[deg]['bk_shot']
        est64 = r64[deg]['bk_raw'] - r64[deg]['bk_shot']
        e['bk'] = float(np.max(np.abs(est32 - est64)) / np.max(np.abs(est64)))
        e['nmodes_equal'] = bool(np.array_equal(r32[deg]['nmodes_1'],
                                                r64[deg]['nmodes_1']))
        errs[tag] = e

    roofline = {
        'bound': 'hbm', 'kernel': dom,
        'achieved': round(kd['achieved_GBs'], 1), 'peak': HBM_PEAK_GBS,
        'unit': 'GB/s', 'frac': round(kd['achieved_GBs'] / HBM_PEAK_GBS, 4),
        'traffic': traffic_for(dom),
        'bytes_per_launch': kd['bytes_per_launch'],
        'avg_launch_us': round(kd['avg_us'], 2),
        'launches_per_step': kd['launches_per_step'],
    }

    # Differentiable B_000 (1e6 particles, 256^3 CIC): forward + backward
    # time against the plain GPU FP64 forward.
    fwd_bwd = None
    if not args.skip_grad:
        fwd_bwd = grad_timings(lib, args.steps, fd_particles=0)

    cpu = cpu_baseline(args, kbv, pairs_per_step) if args.cpu_pairs > 0 else None

    value = pairs_per_step / (ms_per_step / 1e3)
    line = {
        'metric': 'B_ℓ bin-pair evals/s at 256³ mesh, FP32 vs FP64 at stated '
                  'err; fwd+bwd time',
        'value': round(value, 2),
        'unit': 'bin-pair evals/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(ms_per_step, 3),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic (seeded lognormal mock + uniform randoms)',
        'config': {
            'workload': 'survey-like lpp B_000 (full->triu, 210 pairs) + '
                        'B_202 (full, 400 pairs), 20 lin k-bins '
                        '[0.005,0.205], 1e6 lognormal data + 5e6 randoms, '
                        'L=3200, PCS, interlace off; one engine call per '
                        'step with host-resident catalogues in (lines of '
                        'sight formed on the device from the observer '
                        'position) and results out (median of the steps)',
            'ngrid': args.ngrid, 'pairs_per_step': pairs_per_step,
            'n_data': int(data.ntotal), 'n_rand': int(rand.ntotal),
            'precision': 'fp32 fields, fp64 sums',
        },
        'roofline': roofline,
        'cpu_baseline': cpu,
        'step_ms': [round(t * 1e3, 3) for t in t32],
        'cold_ms': round(t_cold * 1e3, 3),
        'fp64': {'value': round(pairs_per_step / float(np.median(t64)), 2),
                 'ms_per_step': round(float(np.median(t64)) * 1e3, 3),
                 'step_ms': [round(t * 1e3, 3) for t in t64]} if t64 else None,
        'fp32_rel_err': errs,
        'host_los_arrays': {
            'value': round(pairs_per_step / float(np.median(tlos)), 2),
            'ms_per_step': round(float(np.median(tlos)) * 1e3, 3),
            'max_rel_diff_vs_device_los': los_agree},
        'device_resident': {
            'value': round(pairs_per_step / float(np.median(tdev)), 2),
            'ms_per_step': round(float(np.median(tdev)) * 1e3, 3),
            'upload_ms': round(t_up * 1e3, 1)},
        'fwd_bwd': fwd_bwd,
        'ms_breakdown_step': round(ms_breakdown, 3),
        'copy_calibration': {'bytes': 1 << 30, 'ms': round(cal_ms.value, 4),
                             'GBs': round(2 * (1 << 30) / cal_ms.value / 1e6, 1)},
        'kernels': kernels,
    }
    return line


def cpu_model():
    try:
        for ln in open('/proc/cpuinfo'):
            if ln.startswith('model name'):
                return ln.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def cpu_baseline(args, kbv, pairs_per_step):
    """The oracle's reference-schedule per-pair loop body on host cores."""
    import oracle
    olib = oracle.load()
    threads = int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))
    olib.trvo_set_threads(threads)
    p = params_struct(args.ngrid, (0, 0, 0), 'full', 64)
    sec = C.c_double()
    log(f"CPU baseline sample: {args.cpu_pairs} pairs on {threads} threads")
    if olib.trvo_bench_pairs(C.byref(p), C.byref(kbv.struct), args.cpu_pairs,
                             C.byref(sec)) != 0:
        return None
    per_pair = sec.value / args.cpu_pairs
    # Pair-loop work per step: 210 pairs x 1 term (B_000) + 400 pairs x 3
    # terms (B_202); fixed per-term costs are not counted.
    work = 210 * 1 + 400 * 3
    fft = fft_calibration(olib, args.ngrid, threads)
    # The per-pair body holds three mesh transforms (two shell IFFTs and the
    # shot-noise IFFT); with an FFTW-class transform in their place:
    per_pair_fftw = per_pair - 3. * (fft['oracle_s'] - fft['torch_s'])
    return {
        'value': round(pairs_per_step / (per_pair * work), 5),
        'unit': 'bin-pair evals/s', 'cores': threads, 'kind': 'port',
        'cpu_model': cpu_model(),
        'sample': (f"reference-schedule per-pair loop body (2 shell IFFTs + "
                   f"triple sum + shot-noise IFFT + spline pass), FP64, "
                   f"{args.ngrid}^3, {args.cpu_pairs} pairs, {sec.value:.1f} "
                   f"s; per-term assignment/FFT costs excluded (favours the "
                   f"CPU)"),
        'seconds_per_pair': round(per_pair, 3),
        'fft_calibration': fft,
        'value_fftw_class': (round(pairs_per_step / (per_pair_fftw * work), 5)
                             if per_pair_fftw > 0 else None),
    }


def fft_calibration(olib, n, threads):
    """One n^3 complex FP64 transform by the oracle's radix-2 FFT and by
    torch's CPU FFT (an FFTW-class library) on the same threads."""
    import torch
    g = np.random.default_rng(0)
    a = g.standard_normal((n, n, n)) + 1j * g.standard_normal((n, n, n))
    buf = np.ascontiguousarray(a).view(np.float64).copy()
    ng = (C.c_int * 3)(n, n, n)
    olib.trvo_fft3d.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int]
    olib.trvo_fft3d.restype = None
    ptr = buf.ctypes.data_as(C.POINTER(C.c_double))
    olib.trvo_fft3d(ptr, ng, +1)
    t = time.perf_counter()
    olib.trvo_fft3d(ptr, ng, +1)
    t_oracle = time.perf_counter() - t
    torch.set_num_threads(threads)
    ta = torch.from_numpy(a)
    torch.fft.ifftn(ta)
    ts = []
    for _ in range(3):
        t = time.perf_counter()
        torch.fft.ifftn(ta)
        ts.append(time.perf_counter() - t)
    t_torch = float(np.median(ts))
    log(f"FFT {n}^3 FP64 on {threads} threads: oracle {t_oracle:.3f} s, "
        f"torch {t_torch:.3f} s")
    return {'n': n, 'threads': threads, 'oracle_s': round(t_oracle, 4),
            'torch_s': round(t_torch, 4)}


def _check(lib, st):
    if st != 0:
        raise RuntimeError(lib.trvg_last_error().decode())


def _upload(lib, cat, los):
    """Device-resident copy of a catalogue (engine handle)."""
    from triumvirate_amd import _abi
    x, y, z, nz, ws, wc = cat.columns()
    view = _abi.CatalogueView(x, y, z, nz, ws, wc)
    h = C.c_void_p()
    _check(lib, lib.trvg_catalogue_upload(C.byref(view.struct),
                                          _abi.dptr(los), 0, C.byref(h)))
    return h


def _median_time(fn, reps):
    """Median wall time of fn() over reps calls (each call is synchronous)
    and the last output."""
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts)), out


def grad_timings(lib, steps, fd_particles=64):
    """Differentiable B_000 diag: 1e6 lognormal particles in a periodic
    L = 1000 Mpc/h box, CIC on a 256^3 FP32 mesh, 15 bins [0.01, 0.16];
    loss = sum_b w_b Re B_000(k_b, k_b), w_b ~ N(0, 1) (seed 5).  Times the
    forward and forward+backward (medians) against the plain GPU FP64
    forward of the same statistic and, with fd_particles > 0, checks
    position gradients against central finite differences (eps = 1e-3
    Mpc/h, FP64 mesh)."""
    import torch
    from triumvirate_amd import _abi, mocks
    from triumvirate_amd import threept as T
    from triumvirate_amd.autodiff import DifferentiableBispectrum
    from triumvirate_amd.catalogue import ParticleCatalogue
    from triumvirate_amd.dataobjs import Binning
    from triumvirate_amd.parameters import ParameterSet
    L, ng = 1000., 256
    x, y, z = mocks.lognormal_box(1_000_000, L, 256, seed=4)
    cat = ParticleCatalogue(x, y, z, nz=len(x) / L ** 3)
    kb = Binning('fourier', 'lin', bin_min=0.01, bin_max=0.16, num_bins=15)
    ps = ParameterSet(param_dict={
        'boxsize': {'x': L, 'y': L, 'z': L},
        'ngrid': {'x': ng, 'y': ng, 'z': ng}, 'assignment': 'cic',
        'norm_convention': 'particle',
        'degrees': {'ell1': None, 'ell2': None, 'ELL': 0}})
    backend = T.CLibBackend(lib)
    ref = T._compute_3pt_stats_sim_like(backend, 'bispec', cat, paramset=ps,
                                        degrees=(0, 0, 0), binning=kb,
                                        form='diag')
    norm = ref['_norm_factors'][0]
    px, py, pz, _, _, _ = cat.columns()        # re-centred by the driver
    pos = torch.tensor(np.stack([px, py, pz], axis=1), device='cuda')
    log(f"grad inputs: {pos.shape[0]} particles")

    # Plain GPU FP64 forward of the same statistic (device-resident
    # catalogue, as in the timed forward below).
    h = _upload(lib, cat, None)
    p = T._params_struct(T._amalgamate_parameters(
        paramset=ps, degrees=(0, 0, 0), binning=kb, form='diag',
        catalogue_type='sim', statistic_type='bispec'))
    res = _abi.ResultBuffers(_abi.result_dim(p.shape, 15, 0), 'bispec')
    kbv = _abi.BinningView(kb.bin_edges, kb.bin_centres)

    def plain64():
        _check(lib, lib.trvg_compute_bispec_dev(h, None, C.byref(p),
                                                C.byref(kbv.struct), norm,
                                                C.byref(res.struct)))
    plain64()
    t_plain, _ = _median_time(plain64, max(5, steps))
    lib.trvg_catalogue_free(h)

    wb = torch.tensor(np.random.default_rng(5).standard_normal(15), device='cuda')
    B32 = DifferentiableBispectrum(L, ng, kb, 'cic', 'fp32', norm)

    def fwd():
        with torch.no_grad():
            out = B32(pos)
        torch.cuda.synchronize()
        return out

    def fwd_bwd():
        pp = pos.clone().requires_grad_(True)
        (B32(pp) * wb).sum().backward()
        torch.cuda.synchronize()
        return pp.grad

    fwd()
    fwd_bwd()
    t_fwd, bk32 = _median_time(fwd, max(5, steps))
    t_fb, g32 = _median_time(fwd_bwd, max(5, steps))
    bk32 = bk32.cpu().numpy()
    ref_re = ref['bk_raw'].real
    err_fwd = float(np.max(np.abs(bk32 - ref_re)) / np.max(np.abs(ref_re)))

    # FP64 differentiable path: the gradient reference.
    B64 = DifferentiableBispectrum(L, ng, kb, 'cic', 'fp64', norm)
    pp = pos.clone().requires_grad_(True)
    (B64(pp) * wb).sum().backward()
    g64 = pp.grad.cpu().numpy()
    err_g32 = float(np.max(np.abs(g32.cpu().numpy() - g64)) / np.max(np.abs(g64)))
    out = {
        'workload': 'periodic box, 1e6 lognormal particles (seed 4), L=1000, '
                    '256^3 CIC, FP32 mesh, 15 diag k-bins [0.01,0.16], loss = '
                    'sum_b w_b Re B_000(k_b,k_b) (medians of the steps)',
        'n_particles': int(pos.shape[0]),
        'fwd_bwd_ms': round(t_fb * 1e3, 3),
        'forward_ms': round(t_fwd * 1e3, 3),
        'plain_fp64_forward_ms': round(t_plain * 1e3, 3),
        'fwd_bwd_over_plain_fp64': round(t_fb / t_plain, 3),
        'fp32_forward_rel_err_vs_fp64_plain': err_fwd,
        'fp32_grad_rel_err_vs_fp64': err_g32,
    }
    if fd_particles > 0:
        rng = np.random.default_rng(6)
        idx = rng.choice(pos.shape[0], fd_particles, replace=False)
        eps = 1e-3
        fd_err = 0.
        gscale = np.max(np.abs(g64[idx]))
        t = time.perf_counter()
        with torch.no_grad():
            for i in idx:
                for a in range(3):
                    q = pos.clone()
                    q[i, a] += eps
                    lp = float((B64(q) * wb).sum())
                    q[i, a] -= 2 * eps
                    lm = float((B64(q) * wb).sum())
                    fd = (lp - lm) / (2 * eps)
                    fd_err = max(fd_err, abs(fd - g64[i, a]) / gscale)
        log(f"FD check {time.perf_counter() - t:.1f} s, max err {fd_err:.2e}")
        out['fd_check'] = {'particles': fd_particles, 'eps': eps,
                           'precision': 'fp64',
                           'max_err_rel_to_max_grad': fd_err}
    return out


def bench_grad(args, lib):
    """Differentiable B_000 (config of grad_timings) with the finite-
    difference check on 64 particles."""
    g = grad_timings(lib, args.steps, fd_particles=64)
    return {
        'metric': 'differentiable B_000 diag fwd+bwd time (1e6 particles, '
                  '256^3 CIC, FP32 mesh)',
        'value': g['fwd_bwd_ms'], 'unit': 'ms', 'n_gpus': 1,
        'steps': args.steps, 'warmup': 1, 'ms_per_step': g['fwd_bwd_ms'],
        'higher_is_better': False, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32', 'data': 'synthetic (seeded lognormal box)',
        'config': {'workload': g['workload'],
                   'n_particles': g['n_particles']},
        **{k: v for k, v in g.items() if k not in ('workload', 'n_particles',
                                                   'fwd_bwd_ms')},
    }


def bench_3pcf(args, lib):
    """Survey-like zeta_000 and zeta_110 (full form -> upper triangle, 465
    pairs each), 30 linear r-bins [10, 200] Mpc/h, 1e7 lognormal data + 3e7
    randoms in the quarter-sky shell, L = 3200 Mpc/h, 512^3 TSC, FP32
    fields with FP64 sums (one FP64 run for the error)."""
    from triumvirate_amd import _abi
    from triumvirate_amd.dataobjs import Binning
    data, rand, los_d, los_r = make_inputs(10_000_000, 30_000_000, 256)
    rb = Binning('config', 'lin', bin_min=10., bin_max=200., num_bins=30)
    rbv = _abi.BinningView(rb.bin_edges, rb.bin_centres)
    t_up = time.perf_counter()
    hd, hr = _upload(lib, data, los_d), _upload(lib, rand, los_r)
    t_up = time.perf_counter() - t_up
    x, y, z, nz, ws, wc = rand.columns()
    rv = _abi.CatalogueView(x, y, z, nz, ws, wc)
    norm = C.c_double()
    _check(lib, lib.trvg_calc_bispec_normalisation_from_particles(
        C.byref(rv.struct), data.wstotal / rand.wstotal, C.byref(norm)))
    ng = 512

    def run(deg, prec):
        # full form with l1 == l2 -> upper triangle (parameters.cpp:830-850)
        p = params_struct(ng, deg, 'triu' if deg[0] == deg[1] else 'full', prec)
        for i in range(3):
            p.boxsize[i] = 3200.
        p.assignment_order = 3
        p.verbose = 20
        res = _abi.ResultBuffers(_abi.result_dim(p.shape, 30, 0), '3pcf')
        _check(lib, lib.trvg_compute_3pcf_dev(hd, hr, C.byref(p),
                                              C.byref(rbv.struct), norm.value,
                                              C.byref(res.struct)))
        return res.as_dict()

    degs = [(0, 0, 0), (1, 1, 0)]
    r64 = {d: run(d, 64) for d in degs}
    log("FP64 done")
    for d in degs:
        run(d, 32)
    t = time.perf_counter()
    for _ in range(args.steps):
        r32 = {d: run(d, 32) for d in degs}
    ms = (time.perf_counter() - t) * 1e3 / args.steps
    pairs = sum(r32[d]['zeta_raw'].shape[0] for d in degs)
    errs = {}
    for d in degs:
        tag = 'zeta' + ''.join(map(str, d))
        errs[tag] = {k: float(np.max(np.abs(r32[d][k] - r64[d][k])) /
                               np.max(np.abs(r64[d][k])))
                     for k in ('zeta_raw', 'zeta_shot')}
        errs[tag]['npairs_equal'] = bool(np.array_equal(r32[d]['npairs_1'],
                                                        r64[d]['npairs_1']))
    lib.trvg_catalogue_free(hd)
    lib.trvg_catalogue_free(hr)
    return {
        'metric': 'zeta_l bin-pair evals/s at 512^3 mesh (FP32 fields)',
        'value': round(pairs / (ms / 1e3), 2), 'unit': 'bin-pair evals/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': 1,
        'ms_per_step': round(ms, 3), 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32',
        'data': 'synthetic (seeded lognormal mock + uniform randoms)',
        'config': {'workload': 'survey-like lpp zeta_000 + zeta_110 (full -> '
                               'triu, 465 pairs each), 30 lin r-bins [10,200], '
                               '1e7 lognormal data + 3e7 randoms, L=3200, '
                               '512^3 TSC, interlace off',
                   'pairs_per_step': pairs, 'n_data': int(data.ntotal),
                   'n_rand': int(rand.ntotal)},
        'fp32_rel_err': errs, 'upload_ms': round(t_up * 1e3, 1),
    }


def bench_mocks(args, lib):
    """256 periodic-box lognormal mocks (seeds 0..255, Poisson mean 3e5,
    uneven sizes), L = 1000 Mpc/h, 128^3 TSC, B_000 (full -> upper triangle,
    120 pairs) + B_202 (full, 225 pairs), 15 linear k-bins [0.01, 0.16]:
    batched plan (FP32 fields, FP64 sums, autotuned, graph-replayed) against
    the plain GPU FP64 single-catalogue path on a sample of mocks."""
    from triumvirate_amd import _abi, mocks
    from triumvirate_amd.batch import BatchedBispectrum
    from triumvirate_amd.catalogue import ParticleCatalogue
    from triumvirate_amd.dataobjs import Binning
    from triumvirate_amd.threept import _multipole_structs
    from triumvirate_amd.parameters import ParameterSet
    L, ng = 1000., 128
    t = time.perf_counter()
    cats = [mocks.lognormal_box(300_000, L, 128, seed=s) for s in range(args.nmocks)]
    sizes = [len(c[0]) for c in cats]
    log(f"mocks: {len(cats)} catalogues, {min(sizes)}..{max(sizes)} particles "
        f"({time.perf_counter() - t:.1f} s)")
    kb = Binning('fourier', 'lin', bin_min=0.01, bin_max=0.16, num_bins=15)
    mps = [((0, 0, 0), 'full'), ((2, 0, 2), 'full')]
    bb = BatchedBispectrum(L, ng, kb, mps, 'tsc', 'fp32', batch=args.batch,
                           capacity=args.batch * max(sizes))
    t = time.perf_counter()
    prepared = bb.prepare(cats)                       # periodise + pack (host)
    t_prep = time.perf_counter() - t
    resident = bb.upload(prepared)                    # catalogues in HBM
    bb.run(resident[:1])                              # tune + capture
    t = time.perf_counter()
    res = bb.run(resident)
    t_batch = time.perf_counter() - t
    t = time.perf_counter()
    bb.run(prepared)                                  # from host arrays
    t_host_inputs = time.perf_counter() - t
    pairs = sum(r['bk_raw'].shape[0] for r in res[0]) * len(cats)

    # Plain GPU FP64 path: one catalogue per call (device-resident upload
    # excluded), both multipoles per call, on a sample of mocks.
    ps = ParameterSet(param_dict={
        'boxsize': {'x': L, 'y': L, 'z': L},
        'ngrid': {'x': ng, 'y': ng, 'z': ng}, 'assignment': 'tsc',
        'norm_convention': 'particle', 'precision': 'fp64',
        'degrees': {'ell1': 0, 'ell2': 0, 'ELL': 0}, 'form': 'diag',
        'range': [0.01, 0.16], 'num_bins': 15})
    plist = _multipole_structs(ps, mps)
    kbv = _abi.BinningView(kb.bin_edges, kb.bin_centres)
    nsample = min(8, len(cats))
    t_plain, err = 0., {'B000': 0., 'B202': 0.}
    for i in range(nsample):
        pc = ParticleCatalogue(*[a.copy() for a in cats[i]])
        pc.periodise([L, L, L])
        pc.compute_mean_density(boxsize=[L, L, L])
        h = _upload(lib, pc, None)
        x, y, z, nz, ws, wc = pc.columns()
        view = _abi.CatalogueView(x, y, z, nz, ws, wc)
        norm = C.c_double()
        _check(lib, lib.trvg_calc_bispec_normalisation_from_particles(
            C.byref(view.struct), 1., C.byref(norm)))
        rb = [_abi.ResultBuffers(_abi.result_dim(p.shape, 15, 0), 'bispec')
              for p in plist]
        arr = (_abi.Params * 2)(*plist)
        outs = (_abi.BispecResult * 2)(*[r.struct for r in rb])
        norms = np.full(2, norm.value)
        t = time.perf_counter()
        _check(lib, lib.trvg_compute_bispec_multi_dev(
            h, None, arr, 2, C.byref(kbv.struct), _abi.dptr(norms), outs))
        if i > 0:                                      # first call warms up
            t_plain += time.perf_counter() - t
        lib.trvg_catalogue_free(h)
        for q, tag in enumerate(err):
            rb[q].struct.dim = outs[q].dim
            ref = rb[q].as_dict()
            for key in ('bk_raw', 'bk_shot'):
                e = float(np.max(np.abs(res[i][q][key] - ref[key])) /
                          np.max(np.abs(ref[key])))
                err[tag] = max(err[tag], e)
    per_mock_plain = t_plain / max(1, nsample - 1)
    value = pairs / t_batch
    return {
        'metric': 'batched-mock B_l bin-pair evals/s (256 periodic lognormal '
                  'mocks, 128^3, FP32 fields)',
        'value': round(value, 1), 'unit': 'bin-pair evals/s', 'n_gpus': 1,
        'steps': 1, 'warmup': 1, 'ms_per_step': round(t_batch * 1e3, 3),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32', 'data': 'synthetic (seeded lognormal boxes)',
        'config': {'workload': f'{len(cats)} periodic lognormal mocks (seeds '
                               '0..N-1, Poisson mean 3e5), L=1000, 128^3 TSC, '
                               'B_000 (triu 120) + B_202 (full 225), 15 lin '
                               'k-bins [0.01,0.16]',
                   'n_mocks': len(cats), 'batch': args.batch,
                   'particles_min': min(sizes), 'particles_max': max(sizes),
                   'pairs_per_step': pairs},
        'mocks_per_s': round(len(cats) / t_batch, 2),
        'host_prepare_ms': round(t_prep * 1e3, 1),
        'from_host_arrays_ms': round(t_host_inputs * 1e3, 1),
        'plain_fp64_ms_per_mock': round(per_mock_plain * 1e3, 3),
        'speedup_vs_plain_fp64': round(per_mock_plain * len(cats) / t_batch, 2),
        'fp32_rel_err_vs_plain_fp64': err,
        'tuning': bb.tuning,
    }


if __name__ == '__main__':
    main()
