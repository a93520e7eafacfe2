/*
 * trv_gpu.h -- C ABI of the MI355X (gfx950) three-point clustering engine.
 *
 * Every entry point replaces one function of Triumvirate's three-point path
 * (cited below as reference file:line).  Plain pointers and sizes only: no
 * C++ or torch types cross this boundary, so a Cython `extern`, a ctypes
 * stub or a C++ program can bind it unchanged.
 *
 * Conventions
 *   - Caller owns every input and output array (as with the reference's
 *     Cython layer, src/triumvirate/_threept.pyx:137-161).  The library owns
 *     all device memory it allocates and frees it before returning, except
 *     where an entry point documents a persistent workspace.
 *   - Complex outputs are interleaved (re, im) doubles, i.e. the memory layout
 *     of std::complex<double>[dim].
 *   - Mesh index order is row-major with the last (z) index fastest, as
 *     MeshField::ret_grid_index (src/triumvirate/src/field.cpp:533-538).
 *   - Return value: TRVG_OK (0) or an error code; the message of the last
 *     error on the calling thread is available from trvg_last_error().
 *     Parameter errors map to the reference's trv::sys::InvalidParameterError,
 *     data errors to trv::sys::InvalidDataError (include/monitor.hpp:731-779);
 *     device errors are reported, never silently ignored.
 */
#ifndef TRV_GPU_H_INCLUDED
#define TRV_GPU_H_INCLUDED

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TRVG_ABI_VERSION 1

/* Status codes. */
enum {
  TRVG_OK = 0,
  TRVG_ERR_PARAM = 1,     /* invalid parameter (InvalidParameterError) */
  TRVG_ERR_DATA = 2,      /* invalid data (InvalidDataError) */
  TRVG_ERR_DEVICE = 3,    /* HIP runtime / kernel launch failure */
  TRVG_ERR_ALLOC = 4,     /* device or host allocation failure */
  TRVG_ERR_INTERNAL = 5
};

/* Binning shapes, i.e. ParameterSet::shape after validation
 * (src/triumvirate/src/parameters.cpp:830-850: form "full" with
 * ell1 == ell2 becomes shape "triu"). */
enum {
  TRVG_SHAPE_DIAG = 0,
  TRVG_SHAPE_OFFDIAG = 1,
  TRVG_SHAPE_ROW = 2,
  TRVG_SHAPE_FULL = 3,
  TRVG_SHAPE_TRIU = 4
};

/* Field precision of assignment, FFTs and Y_lm products.  Bin sums,
 * normalisation, shot-noise amplitudes and Wigner-3j coupling are always
 * accumulated in FP64. */
enum {
  TRVG_PREC_FP64 = 64,
  TRVG_PREC_FP32 = 32
};

/* A particle catalogue viewed in place.  Element i of a column is
 * col[i * stride].  stride = 1 for separate (SoA) arrays; stride = 7 for
 * the reference's 56-byte ParticleData {pos[3], nz, ws, wc, w}
 * (include/particles.hpp:63-69), with x = &pdata[0].pos[0], etc.
 * The total weight is w = ws * wc as in ParticleCatalogue::load_particle_data
 * (src/particles.cpp:553).  nz may be NULL where unused. */
typedef struct trvg_catalogue {
  long long ntotal;
  long long stride;
  const double* x;
  const double* y;
  const double* z;
  const double* nz;
  const double* ws;
  const double* wc;
} trvg_catalogue;

/* The numerical subset of trv::ParameterSet (include/parameters.hpp:58-240)
 * that the three-point path reads, after ParameterSet::validate. */
typedef struct trvg_params {
  double boxsize[3];
  int ngrid[3];
  int assignment_order;   /* 1 ngp, 2 cic, 3 tsc, 4 pcs */
  int interlace;          /* forced to 0 for three-point statistics */
  int ell1, ell2, ELL;
  int shape;              /* TRVG_SHAPE_* */
  int idx_bin;
  int precision;          /* TRVG_PREC_FP64 (drop-in) or TRVG_PREC_FP32 */
  int verbose;
  int device;             /* HIP device ordinal; -1 = current device */
} trvg_params;

/* trv::Binning (include/dataobjs.hpp:58): edges has num_bins + 1 entries,
 * centres has num_bins entries. */
typedef struct trvg_binning {
  int num_bins;
  const double* bin_edges;
  const double* bin_centres;
} trvg_binning;

/* trv::BispecMeasurements (include/dataobjs.hpp:249-265).  The caller sets
 * `dim` to the capacity of each array (see trvg_result_dim); the library
 * sets it to the filled length.  bk_raw/bk_shot hold 2*dim doubles. */
typedef struct trvg_bispec_result {
  int dim;
  double* k1_bin;
  double* k2_bin;
  double* k1_eff;
  double* k2_eff;
  int* nmodes_1;
  int* nmodes_2;
  double* bk_raw;
  double* bk_shot;
} trvg_bispec_result;

/* trv::ThreePCFMeasurements (include/dataobjs.hpp:267-283). */
typedef struct trvg_3pcf_result {
  int dim;
  double* r1_bin;
  double* r2_bin;
  double* r1_eff;
  double* r2_eff;
  int* npairs_1;
  int* npairs_2;
  double* zeta_raw;
  double* zeta_shot;
} trvg_3pcf_result;

/* trv::PowspecMeasurements (include/dataobjs.hpp:202-211).  pk_raw and
 * pk_shot hold 2*dim doubles (real, imaginary). */
typedef struct trvg_powspec_result {
  int dim;
  double* kbin;
  double* keff;
  int* nmodes;
  double* pk_raw;
  double* pk_shot;
} trvg_powspec_result;

/* trv::TwoPCFMeasurements and trv::TwoPCFWindowMeasurements
 * (include/dataobjs.hpp:217-238).  xi holds 2*dim doubles. */
typedef struct trvg_2pcf_result {
  int dim;
  double* rbin;
  double* reff;
  int* npairs;
  double* xi;
} trvg_2pcf_result;

/* ---------------------------------------------------------------------- */
/* Library information and errors                                          */
/* ---------------------------------------------------------------------- */

int trvg_abi_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* trvg_last_error(void);

/* Number of visible HIP devices (0 when none); never initialises a context
 * beyond what hipGetDeviceCount does. */
int trvg_device_count(int* count);

/* Data-vector dimension for a shape and bin count: the dv_dim switch of
 * compute_bispec (src/threept.cpp:275-298). */
int trvg_result_dim(int shape, int num_bins, int idx_bin, int* dim);

/* ---------------------------------------------------------------------- */
/* Normalisation                                                           */
/* ---------------------------------------------------------------------- */

/* Replaces trv::calc_bispec_normalisation_from_particles
 * (include/threept.hpp:138-140; src/threept.cpp:96-136):
 * 1 / (alpha * sum_p ws nz^2 wc^3).  Fails with TRVG_ERR_DATA if the sum
 * is zero (all nz zero). */
int trvg_calc_bispec_normalisation_from_particles(
  const trvg_catalogue* particles, double alpha, double* norm_factor);

/* Replaces trv::calc_bispec_normalisation_from_mesh
 * (include/threept.hpp:150-152; src/threept.cpp:138-149, field.cpp:2017-2065):
 * 1 / (V_cell sum_x n(x)^3) / alpha^3 with n assigned on the GPU. */
int trvg_calc_bispec_normalisation_from_mesh(
  const trvg_catalogue* particles, const trvg_params* params, double alpha,
  double* norm_factor);

/* ---------------------------------------------------------------------- */
/* Estimators                                                              */
/* ---------------------------------------------------------------------- */

/* Replaces trv::compute_bispec (include/threept.hpp:237-242;
 * src/threept.cpp:248-1012): survey-type data + random catalogues, local
 * plane-parallel lines of sight.  los_data / los_rand are ntotal x 3
 * doubles, the memory layout of trv::LineOfSight[] (dataobjs.hpp:186). */
int trvg_compute_bispec(
  const trvg_catalogue* data, const trvg_catalogue* rand,
  const double* los_data, const double* los_rand,
  const trvg_params* params, const trvg_binning* kbinning,
  double norm_factor, trvg_bispec_result* out);

/* Replaces trv::compute_3pcf (include/threept.hpp:257-262;
 * src/threept.cpp:1014-1471). */
int trvg_compute_3pcf(
  const trvg_catalogue* data, const trvg_catalogue* rand,
  const double* los_data, const double* los_rand,
  const trvg_params* params, const trvg_binning* rbinning,
  double norm_factor, trvg_3pcf_result* out);

/* Replaces trv::compute_bispec_in_gpp_box (include/threept.hpp:274-278;
 * src/threept.cpp:1473-2188): periodic box, global plane-parallel. */
int trvg_compute_bispec_in_gpp_box(
  const trvg_catalogue* data,
  const trvg_params* params, const trvg_binning* kbinning,
  double norm_factor, trvg_bispec_result* out);

/* Replaces trv::compute_3pcf_in_gpp_box (include/threept.hpp:290-294;
 * src/threept.cpp:2190-2619). */
int trvg_compute_3pcf_in_gpp_box(
  const trvg_catalogue* data,
  const trvg_params* params, const trvg_binning* rbinning,
  double norm_factor, trvg_3pcf_result* out);


/* Replaces trv::compute_3pcf_window (include/threept.hpp:310-314;
 * src/threept.cpp:2621-3077): the three-point correlation function window
 * of one random catalogue (fields of that catalogue alone scaled by
 * `alpha`), with the wide-angle power-law kernel |x|^-(i_wa + j_wa) on the
 * G field when wide_angle is non-zero (field.cpp:1727-1762; the reference
 * reads the orders from ParameterSet::i_wa, j_wa). */
int trvg_compute_3pcf_window(
  const trvg_catalogue* rand, const double* los_rand,
  const trvg_params* params, const trvg_binning* rbinning, double alpha,
  double norm_factor, int wide_angle, int i_wa, int j_wa,
  trvg_3pcf_result* out);

/* ---------------------------------------------------------------------- */
/* Two-point statistics (include/twopt.hpp; src/twopt.cpp).  params->ELL is */
/* the multipole degree; interlacing follows params->interlace.  Result    */
/* capacity: `dim` >= num_bins.                                            */
/* ---------------------------------------------------------------------- */

/* Replaces trv::calc_powspec_normalisation_from_particles
 * (include/twopt.hpp:89-91; src/twopt.cpp:56-96): 1 / (alpha sum ws nz wc^2). */
int trvg_calc_powspec_normalisation_from_particles(
  const trvg_catalogue* particles, double alpha, double* norm_factor);

/* Replaces trv::calc_powspec_normalisation_from_mesh (include/twopt.hpp:
 * 101-103; src/twopt.cpp:98-109): 1 / (V_cell sum_x n(x)^2) / alpha^2. */
int trvg_calc_powspec_normalisation_from_mesh(
  const trvg_catalogue* particles, const trvg_params* params, double alpha,
  double* norm_factor);

/* Replaces trv::calc_powspec_normalisation_from_meshes (include/twopt.hpp:
 * 122-125; src/twopt.cpp:111-228) for catalogues already aligned in the
 * box of `params`: 1 / (alpha V_cell sum_x n_d(x) n_r(x)).  The reference's
 * re-alignment (centre or pad in that box) is the caller's step; the
 * Python and C++ layers of this library perform it. */
int trvg_calc_powspec_normalisation_from_meshes(
  const trvg_catalogue* data, const trvg_catalogue* rand,
  const trvg_params* params, double alpha, double* norm_factor);

/* Replaces trv::calc_powspec_shotnoise_from_particles (include/twopt.hpp:
 * 169-171; src/twopt.cpp:271-293): sum ws^2 wc^2. */
int trvg_calc_powspec_shotnoise_from_particles(
  const trvg_catalogue* particles, double alpha, double* shotnoise);

/* Replaces trv::compute_powspec (include/twopt.hpp:246-251;
 * src/twopt.cpp:388-496). */
int trvg_compute_powspec(
  const trvg_catalogue* data, const trvg_catalogue* rand,
  const double* los_data, const double* los_rand,
  const trvg_params* params, const trvg_binning* kbinning,
  double norm_factor, trvg_powspec_result* out);

/* Replaces trv::compute_corrfunc (include/twopt.hpp:266-271;
 * src/twopt.cpp:498-607). */
int trvg_compute_corrfunc(
  const trvg_catalogue* data, const trvg_catalogue* rand,
  const double* los_data, const double* los_rand,
  const trvg_params* params, const trvg_binning* rbinning,
  double norm_factor, trvg_2pcf_result* out);

/* Replaces trv::compute_powspec_in_gpp_box (include/twopt.hpp:283-287;
 * src/twopt.cpp:609-701). */
int trvg_compute_powspec_in_gpp_box(
  const trvg_catalogue* data, const trvg_params* params,
  const trvg_binning* kbinning, double norm_factor, trvg_powspec_result* out);

/* Replaces trv::compute_corrfunc_in_gpp_box (include/twopt.hpp:299-303;
 * src/twopt.cpp:703-793). */
int trvg_compute_corrfunc_in_gpp_box(
  const trvg_catalogue* data, const trvg_params* params,
  const trvg_binning* rbinning, double norm_factor, trvg_2pcf_result* out);

/* Replaces trv::compute_corrfunc_window (include/twopt.hpp:317-321;
 * src/twopt.cpp:795-901). */
int trvg_compute_corrfunc_window(
  const trvg_catalogue* rand, const double* los_rand,
  const trvg_params* params, const trvg_binning* rbinning, double alpha,
  double norm_factor, trvg_2pcf_result* out);


/* ---------------------------------------------------------------------- */
/* Engine extensions (no counterpart in the reference)                     */
/* ---------------------------------------------------------------------- */

/* Device-resident catalogue: uploaded once, reused by any number of
 * estimator calls (for callers that keep particles in HBM, e.g. repeated
 * measurements and benchmarks).  `los` (ntotal x 3) is required for
 * survey-type (local plane-parallel) use and may be NULL otherwise.
 * device = -1 uses the current device. */
typedef struct trvg_device_catalogue trvg_device_catalogue;

int trvg_catalogue_upload(const trvg_catalogue* cat, const double* los,
                          int device, trvg_device_catalogue** out);

/* As trvg_catalogue_upload with the lines of sight formed on the device
 * from the positions: los_i = (r_i - observer) / |r_i - observer| (a zero
 * norm taken as 1), observer[3] being the observer's position in the
 * coordinates of `cat`.  For a catalogue aligned in the box by an offset
 * (ParticleCatalogue::offset_coords, positions r - d), observer = -d
 * gives the lines of sight the reference's driver computes before
 * alignment (threept.py:525-528, catalogue.py:437-458), up to the
 * rounding of the shift; no n x 3 array crosses the bus. */
int trvg_catalogue_upload_observer(const trvg_catalogue* cat, const double* observer,
                                   int device, trvg_device_catalogue** out);
int trvg_catalogue_free(trvg_device_catalogue* cat);

/* As trvg_calc_bispec_normalisation_from_mesh on a device-resident
 * catalogue (the survey drivers compute both normalisations and the
 * estimator from one upload of the randoms). */
int trvg_calc_bispec_normalisation_from_mesh_dev(
  const trvg_device_catalogue* particles, const trvg_params* params,
  double alpha, double* norm_factor);

/* As trvg_compute_bispec / trvg_compute_bispec_in_gpp_box (rand == NULL)
 * on device-resident catalogues.  A device catalogue must not be used by
 * two calls at the same time (the engine reorders its columns). */
int trvg_compute_bispec_dev(const trvg_device_catalogue* data,
                            const trvg_device_catalogue* rand,
                            const trvg_params* params,
                            const trvg_binning* kbinning, double norm_factor,
                            trvg_bispec_result* out);

/* As trvg_compute_3pcf / trvg_compute_3pcf_in_gpp_box (rand == NULL). */
int trvg_compute_3pcf_dev(const trvg_device_catalogue* data,
                          const trvg_device_catalogue* rand,
                          const trvg_params* params,
                          const trvg_binning* rbinning, double norm_factor,
                          trvg_3pcf_result* out);

/* Several bispectrum multipoles in one call on the same catalogues:
 * params[i] give (ell1, ell2, ELL, shape, idx_bin) per multipole and must
 * agree on box, mesh, assignment and precision.  The multipoles share the
 * common fields, shell statistics and shell fields (e.g. the (0, 0) shells
 * of B_000 are the l2 = 0 side of B_202).  rand == NULL selects the
 * periodic-box estimator.  outs[i] receives multipole i (norm_factors[i]). */
int trvg_compute_bispec_multi_dev(const trvg_device_catalogue* data,
                                  const trvg_device_catalogue* rand,
                                  const trvg_params* params, int nmultipoles,
                                  const trvg_binning* kbinning,
                                  const double* norm_factors,
                                  trvg_bispec_result* outs);

/* Host-array form of trvg_compute_bispec_multi_dev (rand/los NULL for the
 * periodic box). */
int trvg_compute_bispec_multi(const trvg_catalogue* data,
                              const trvg_catalogue* rand,
                              const double* los_data, const double* los_rand,
                              const trvg_params* params, int nmultipoles,
                              const trvg_binning* kbinning,
                              const double* norm_factors,
                              trvg_bispec_result* outs);

/* trvg_compute_bispec_multi for survey catalogues with the lines of sight
 * formed on the device from an observer position (as
 * trvg_catalogue_upload_observer). */
int trvg_compute_bispec_multi_observer(const trvg_catalogue* data,
                                       const trvg_catalogue* rand,
                                       const double* observer,
                                       const trvg_params* params, int nmultipoles,
                                       const trvg_binning* kbinning,
                                       const double* norm_factors,
                                       trvg_bispec_result* outs);

/* Frees the cached device work buffers (they are otherwise kept for reuse
 * by later calls). */
int trvg_release_workspace(void);

/* Kernel-class profiler.  When enabled (which also clears the totals),
 * every kernel launch is bracketed by HIP events on the engine's stream and
 * the elapsed time, launch count and algorithmic bytes (the compulsory HBM
 * traffic of the launch) are accumulated per kernel class.  Classes are
 * numbered 0 .. trvg_profile_class_count() - 1. */
int trvg_profile_enable(int on);
/* Restricts the profiler to one kernel class (-1: all classes), so that a
 * timed run carries events only around the launches being measured. */
int trvg_profile_only(int kernel_class);
int trvg_profile_read(int kernel_class, double* total_ms, long long* launches,
                      double* algorithmic_bytes);
const char* trvg_profile_class_name(int kernel_class);
int trvg_profile_class_count(void);

/* Component entry points (diagnostics and parity tests). */

/* Mesh assignment of a catalogue with per-particle weights
 * (mode 0: 1; 1: w; 2: w Y_lm(los); 3: w^2 conj(Y_lm(los))), field.cpp
 * MeshField::assign_weighted_field_to_mesh (src/field.cpp:569-1112).
 * mesh_out: 2 * nmesh doubles (interleaved complex). */
int trvg_mesh_assign(const trvg_catalogue* cat, const double* los,
                     const trvg_params* params, int mode, int ell, int m,
                     double* mesh_out);

/* In-place unnormalised 3-D FFT of an interleaved complex host mesh
 * (sign -1 forward, +1 backward), computed on the device in
 * params->precision. */
int trvg_fft3d(const trvg_params* params, double* mesh, int sign);

/* ------------------------------------------------------------------------
 * Batched periodic-box bispectrum (no reference counterpart: the reference
 * computes one catalogue per call).  A plan fixes the mesh, binning and a
 * list of multipoles (`multipoles[q]`: geometry, assignment, precision and
 * degrees / shape / idx_bin of multipole q; all share the mesh) for batches
 * of up to `batch` catalogues holding at most `capacity` particles in
 * total.  Each result equals compute_bispec_in_gpp_box
 * (threept.cpp:1473-2188) of that catalogue with the particle
 * normalisation (nz = N / V, threept.cpp:96-136).  With autotune != 0 the
 * first run times candidate launch shapes and keeps the fastest; with
 * use_graph != 0 the whole batch is captured into a HIP graph on the first
 * run and replayed afterwards.
 * ------------------------------------------------------------------------ */
typedef struct trvg_batch_plan trvg_batch_plan;

int trvg_batch_plan_create(const trvg_params* multipoles, int nmultipoles,
                           const trvg_binning* kbinning, int batch,
                           long long capacity, int autotune, int use_graph,
                           trvg_batch_plan** plan);

/* Positions of catalogue b are x/y/z[o_b .. o_b + counts[b]) with
 * o_b = counts[0] + ... + counts[b-1] (host arrays, already in [0, L));
 * outs[b * nmultipoles + q] receives multipole q of catalogue b. */
int trvg_batch_run(trvg_batch_plan* plan, int ncatalogues,
                   const long long* counts, const double* x, const double* y,
                   const double* z, trvg_bispec_result* outs);

/* As trvg_batch_run with x / y / z DEVICE pointers (catalogues already in
 * device memory, e.g. resident for many batches). */
int trvg_batch_run_device(trvg_batch_plan* plan, int ncatalogues,
                          const long long* counts, const double* x,
                          const double* y, const double* z,
                          trvg_bispec_result* outs);

/* tuned[5]: assignment, pair-reduction and binned-statistics blocks per
 * catalogue, pair row-block cap, and whether the graph is captured. */
int trvg_batch_plan_info(const trvg_batch_plan* plan, int* tuned);
int trvg_batch_plan_free(trvg_batch_plan* plan);

/* ------------------------------------------------------------------------
 * Differentiable periodic-box bispectrum (no reference counterpart: the
 * reference computes forward values only).  The forward chain is that of
 * compute_bispec_in_gpp_box (threept.cpp:1473-2188) for B_000 in the diag
 * form, from a real mesh delta(x) instead of a catalogue; the adjoints are
 * hand-written.  All pointers except bk / nmodes / grad_bk are DEVICE
 * pointers; meshes are real, n0 x n1 x n2 row-major, in params->precision
 * (float for TRVG_PREC_FP32, double otherwise); positions are n x 3 FP64.
 * Calls are synchronous on the device.
 * ------------------------------------------------------------------------ */

/* mesh (+)= sum_p w_p W(x - x_p) / V_cell with the assignment scheme of
 * params (field.cpp:569-1112 index and window rules); weights may be NULL
 * (unit weights); accumulate = 0 zeroes the mesh first. */
int trvg_diff_assign(const trvg_params* params, const double* pos,
                     const double* weights, long long n, void* mesh,
                     int accumulate);

/* Adjoint of trvg_diff_assign: grad_pos (n x 3) and grad_weights (n), each
 * optional, from the mesh cotangent grad_mesh. */
int trvg_diff_assign_backward(const trvg_params* params, const double* pos,
                              const double* weights, long long n,
                              const void* grad_mesh, double* grad_pos,
                              double* grad_weights);

/* B_b = norm V_cell / n_b^2 sum_x F_b(x)^2 G(x) for every k-bin b (host
 * outputs bk[num_bins], optional nmodes[num_bins]); F_b and G are the
 * real parts of the shell field and compensated field of delta. */
int trvg_diff_bispec_diag(const trvg_params* params,
                          const trvg_binning* kbinning, double norm_factor,
                          const void* mesh, double* bk, int* nmodes);

/* The forward of trvg_diff_bispec_diag without any device allocation.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Every buffer beyond `mesh` and the host outputs lives
 * in one fixed device array of the library (trvg_noalloc_scratch_bytes(),
 * 64 MiB, held on each device the library runs on until the process ends):
 * FP32 meshes up to 2^21 cells and FP64 meshes up to 2^20 cells, at most 256
 * bins; larger inputs fail with TRVG_ERR_PARAM before any device work.
 * Runs on the calling thread's default stream and returns with the results
 * on the host.  Calls that overlap in time on one device share the array
 * and race. */
int trvg_diff_bispec_diag_noalloc(const trvg_params* params,
                                  const trvg_binning* kbinning,
                                  double norm_factor, const void* mesh,
                                  double* bk, int* nmodes);
long long trvg_noalloc_scratch_bytes(void);

/* grad_mesh = sum_b grad_bk[b] dB_b/ddelta (shells recomputed, not
 * stored). */
int trvg_diff_bispec_diag_backward(const trvg_params* params,
                                   const trvg_binning* kbinning,
                                   double norm_factor, const void* mesh,
                                   const double* grad_bk, void* grad_mesh);

/* Streams `bytes` (rounded down to 16-byte units) from one device buffer
 * to another with a streaming copy kernel (four 16-byte loads in flight
 * per thread, nontemporal loads and stores, one pass) and returns its
 * duration in milliseconds (best of three): the device's practical HBM
 * copy rate, and the reference point for scaling hardware byte counters
 * (a copy moves exactly `bytes` in and `bytes` out). */
int trvg_copy_calibration(long long bytes, double* ms);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif  /* TRV_GPU_H_INCLUDED */
