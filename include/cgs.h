/* cgs.h -- C-ABI of libcgs_hip.so, the MI355X (gfx950) CG-SENSE engine.
 *
 * Drop-in boundary for the hot path of expRad/4d_lung's CGSense program: the
 * encoding operator E (Reco::nufftFWD), its adjoint E^H (Reco::nufftBWD), the
 * centred 3D FFT (Fourier::transformFWD/BWD) and the CG-SENSE loop
 * (Reco::cgSenseDynamic).  Plain pointers and sizes only; every call returns a
 * status (0 = OK) and never exits the process -- where the reference prints and
 * calls exit(EXIT_FAILURE) (Reco.cpp:194-195, 266-267, 333-334, 475-476) this
 * library returns a non-zero status and keeps a message for cgs_last_error().
 *
 * Layouts are the reference's (host side):
 *   complex values   float[2] interleaved (fftwf_complex)
 *   images / maps    [x][y][z], z fastest; several coils: [c][x][y][z] (IO.cpp:25-51)
 *   k-space data     [c][i] with i = ix*(sY*sZ) + iy*sZ + iz, readout ix slowest,
 *                    zero-filled where a line is not acquired (IO.cpp:55-98)
 *   trajectory       [i][3] doubles in [0,1] per axis (Trajectory.h:32)
 *   density weights  [i] floats in the same sample order (IO.cpp:129-156)
 *
 * A context is used from one host thread at a time.  Host buffers are only
 * borrowed for the duration of a call.
 */
#ifndef CGS_H
#define CGS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cgs_ctx cgs_ctx;
typedef struct cgs_fft cgs_fft;

#define CGS_OK 0
#define CGS_ERR_ARG 1         /* invalid argument or unsupported size */
#define CGS_ERR_HIP 2         /* HIP runtime error */
#define CGS_ERR_STATE 3       /* call out of order (e.g. no trajectory yet) */
#define CGS_ERR_NONFINITE 4   /* NaN/Inf in q[0] during CG (Reco.cpp:264-268) */
#define CGS_ERR_MEMORY 5      /* device allocation failed */

#define CGS_MAX_COILS 256     /* coils per context (memory permitting) */

#define CGS_SOLVE_PLAIN 0     /* one bin at a time, host reads the residual every iteration */
#define CGS_SOLVE_BATCHED 1   /* all bins per launch, device-side loop control, hipGraph replay */

/* Message of the last failing call on this context (never NULL). */
const char* cgs_last_error(const cgs_ctx* h);

/* Reco::setGridSize / setSampling / setNCh + Reco::init (Reco.cpp:374-400,
 * 415-427): grid = oversampled grid (x,y,z), samp = (sX, sY, sZ).  KB constants
 * default to the reference's w = 4.25, b = 14.9086, order 5 (Reco.cpp:378-380),
 * imgRange to the grid (Reco.cpp:397). */
int cgs_create(const int grid[3], const int samp[3], int ncoil, cgs_ctx** out);
/* Reco::~Reco (Reco.cpp:29-61). */
int cgs_destroy(cgs_ctx* h);

/* ConvKernel::setW / setB / setKernelOrder (Reco.h:54-56). */
int cgs_set_kernel(cgs_ctx* h, float w, float b, int order);
/* reduceFOV box (Reco.cpp:626-671; imgRange Reco.h:195, offcenter Reco.cpp:429-434). */
int cgs_set_img_range(cgs_ctx* h, const int range[3], const int offcenter[3]);
/* Gridding tile shape index (0..cgs_num_tile_configs()-1), chosen by the autotuner. */
int cgs_set_tile_config(cgs_ctx* h, int tcfg);
int cgs_num_tile_configs(void);
int cgs_get_tile_config(const cgs_ctx* h, int* tcfg, int dims[3]);

/* Trajectory::kSpaceTraj, as Reco::setTrajectory leaves it (Reco.cpp:324-336). */
int cgs_set_trajectory(cgs_ctx* h, const double* ktraj);
/* ConvKernel::calculateWeights (Reco.cpp:698-787): every sample active. */
int cgs_calculate_weights(cgs_ctx* h);
/* ConvKernel::excludeMissingData (Reco.cpp:681-696): samples whose coil-0
 * value in data ([c][i] complex) is exactly 0+0i are dropped. */
int cgs_exclude_missing_data(cgs_ctx* h, const float* data);
/* ConvKernel::calculateKernelFT (Reco.cpp:560-601). */
int cgs_calculate_kernel_ft(cgs_ctx* h);
int cgs_get_kernel_ft(cgs_ctx* h, float* kft /* Ng floats (real part) */);
/* Reco::denCorr (Reco.h:86); NULL = Density::NONE. */
int cgs_set_density(cgs_ctx* h, const float* dcf);
/* Reco::sMaps (IO::importSensitivityMaps, IO.cpp:25-51): [c][x][y][z] complex. */
int cgs_set_maps(cgs_ctx* h, const float* maps);

/* Reco::nufftFWD (Reco.cpp:521-553): img [c][Ng] complex -> data [c][i]
 * complex (samples that are not active are written as 0). */
int cgs_nufft_fwd(cgs_ctx* h, const float* img, float* data);
/* Reco::nufftBWD (Reco.cpp:468-519): data [c][i] -> img [c][Ng]. */
int cgs_nufft_bwd(cgs_ctx* h, const float* data, float* img);
/* One CG system matrix product q = sum_c conj(S_c) E^H W E (S_c p)
 * (Reco.cpp:232-262); p, q: Ng complex. */
int cgs_apply_normal(cgs_ctx* h, const float* p, float* q);

/* Reco::cgSenseDynamic (Reco.cpp:175-321) started from coil images img
 * [c][Ng] (the reference's Reco::img after nufftBWD).  At most max_iter + 1
 * passes (the reference's nMaxIter = 15, Reco.cpp:228).  x_out: Ng complex;
 * resid_out: max_iter + 1 doubles (Reco::cgResiduals); iters_out: passes made
 * (Reco::numPerformedCGIterations). */
int cgs_cg_sense_images(cgs_ctx* h, const float* img, double thresh, int max_iter, float* x_out,
                        double* resid_out, int* iters_out);

/* ---- respiratory bins (all bins share trajectory, maps and KB tables) ---- */
/* Number of datasets (respiratory bins) held by the context (default 1). */
int cgs_set_num_bins(cgs_ctx* h, int nbins);
/* Bin data in the reference's zero-filled layout (one state{n}.bin,
 * CGSense.cpp:200-226): active samples = coil-0 value != 0; dcf may be NULL
 * (then the context's density weights, if any, are used). */
int cgs_set_bin_data(cgs_ctx* h, int bin, const float* data, const float* dcf);
/* Largest number of bins one batched launch covers (0, the default: as many
 * as the device memory holds coil grids for; the bins are split into equal
 * groups).  Takes effect at the next cgs_cg_prepare.  No reference
 * counterpart: the reference solves one state at a time (CGSense.cpp:200-282). */
int cgs_set_bin_group(cgs_ctx* h, int max_bins);
/* Bins per batched launch chosen by the last cgs_cg_prepare. */
int cgs_get_bin_group(const cgs_ctx* h, int* bins_per_launch);
/* Column pruning of the bins' x passes: column groups of the grid that no
 * sample's KB window reaches are skipped by the forward x pass and written as
 * zeros by the inverse one, and gridding skips the tiles of those columns
 * (same images; pays for the sparse line sets of respiratory bins).  mode 0:
 * off; 1 (default): on when it skips at least a tenth of the groups; 2: on.
 * Takes effect when the bins are next packed.  No reference counterpart. */
int cgs_set_pruning(cgs_ctx* h, int mode);
/* 1 if the packed bins use column pruning. */
int cgs_get_pruning(const cgs_ctx* h, int* active);
/* Fractions of bin's x-pass column groups that are transformed and of its
 * gridding tile columns that are written (1 and 1 without pruning); valid
 * while the bins stay packed (after cgs_cg_prepare). */
int cgs_get_prune_fractions(const cgs_ctx* h, int bin, double* groups, double* tiles);
/* Per-tile gridding candidate lists (on = 1, the default): the samples whose
 * KB window meets each gridding tile are listed once when the bins are packed
 * (the trajectory is fixed across CG iterations) instead of being found by a
 * scan of the 27 neighbouring tiles in every gridding launch (on = 0).  Same
 * candidates in the same order, so the same images bit for bit.  Takes effect
 * when the bins are next packed.  No reference counterpart. */
int cgs_set_grid_lists(cgs_ctx* h, int on);
/* Inverse z pass of batched launches (the fused IFFT along z, deapodization
 * and coil combination of E^H): bins per workgroup, 1..64 (default 64).
 * With 2 or more, one workgroup takes one cell line for that many bins of
 * the launch, the line's coil maps held in registers (up to 16 coils) and
 * read once for them all; with 1, one workgroup per bin and cell line.  The
 * same images bit for bit.  No reference counterpart. */
int cgs_set_zpass_bins(cgs_ctx* h, int bins);
/* Gridding fused into the inverse x pass (on = 1, the default): for the CG
 * iterations and right-hand sides of the packed bins, the inverse x pass
 * forms each 16-coil chunk of a (y, z) cell column from the samples whose KB
 * window covers the cell (lists sorted by window origin, built when the bins
 * are packed) instead of reading a gridded grid -- the same taps and weights
 * as gridding, summed in a different order.  Used where the x pass takes
 * 16-coil chunks (coils a multiple of 16, x planes of 1 MB or more) and the
 * densest bin stages at most a quarter of the bytes of the coil grids' write
 * and read (each sample once per (y, z) of its window: sparse respiratory
 * bins, not a fully sampled frame); takes effect when the bins are next
 * packed.  No reference counterpart. */
int cgs_set_fused_gridding(cgs_ctx* h, int on);
/* *on = 1 when the packed bins use the fused gridding. */
int cgs_fused_gridding_active(cgs_ctx* h, int* on);
/* Compact form: m acquired samples with original indices idx[m] (any order),
 * values y[c][m] complex and optional weights dcf[m]. */
int cgs_set_bin_samples(cgs_ctx* h, int bin, long long m, const int32_t* idx, const float* y, const float* dcf);
/* Right-hand sides b = E^H W y of every bin and CG state reset. */
int cgs_cg_prepare(cgs_ctx* h, double thresh, int max_iter);
/* Run CG iterations on all bins (mode CGS_SOLVE_PLAIN or CGS_SOLVE_BATCHED)
 * until every bin stopped or n_iter iterations were issued; returns after the
 * device finished. */
int cgs_cg_iterate(cgs_ctx* h, int mode, int n_iter);
/* Results: x_out nbins*Ng complex (may be NULL), resid_out nbins*n doubles
 * with n from cgs_cg_history_length (max_iter + 1 of the last prepare or
 * solve), iters_out nbins. */
int cgs_cg_results(cgs_ctx* h, float* x_out, double* resid_out, int* iters_out);
/* Residual history length per bin (Reco::cgResiduals, Reco.cpp:299) of the
 * current CG state: the doubles per bin that cgs_cg_results writes. */
int cgs_cg_history_length(const cgs_ctx* h, int* n);
/* prepare + iterate(max_iter + 1) + results. */
int cgs_cg_sense_bins(cgs_ctx* h, int mode, double thresh, int max_iter, float* x_out, double* resid_out,
                      int* iters_out);

/* ---- setup introspection (bit-exactness tests) ---- */
/* Active samples of a bin and their tile-sorted order. */
int cgs_bin_count(cgs_ctx* h, int bin, long long* m);
/* nearest[i][3] of every sample, perm[m] original index per sorted position,
 * tile_off[ntiles+1], base[m][3] window origins, vals[m][3*kw] per-axis KB
 * factors (FP64); any pointer may be NULL. */
int cgs_bin_tables(cgs_ctx* h, int bin, int32_t* nearest, int32_t* perm, int32_t* tile_off, int32_t* base,
                   double* vals);
int cgs_kernel_info(const cgs_ctx* h, int* kw, int* ntiles, double* thr);

/* ---- Voronoi density compensation (estimateDensityV3D, functions.h:161-258) ---- */
/* Rules for the cells whose Voronoi region reaches far out:
 *   CGS_VORONOI_HULL     the reference's (functions.h:148-150): volume 0 only
 *                        for unbounded cells, i.e. points on the convex hull;
 *                        every other point gets its finite volume however
 *                        large (the normalisation by the maximum then follows
 *                        the largest one).
 *   CGS_VORONOI_CLIPPED  extension for the per-bin weights of respiratory
 *                        bins: cells that reach more than one hash cell
 *                        (about the mean point spacing) beyond the points'
 *                        bounding box get 0 too, so that the near-degenerate
 *                        boundary cells of sparse line subsets (up to 1e26
 *                        times the median volume) do not crush every other
 *                        weight in the normalisation.  No reference
 *                        counterpart: the reference computes one set of
 *                        weights over all trajectory points. */
#define CGS_VORONOI_HULL 0
#define CGS_VORONOI_CLIPPED 1
/* Cell volumes of np points (np x 3 doubles) by the given rule. */
int cgs_voronoi_volumes_rule(const double* pts, long long np, int rule, double* vol);
/* Weights = volume / max volume (functions.h:234-245) by the given rule; with
 * average != 0 the per-index averaging of functions.h:212-232 (index i mod
 * sx, cut-off) is applied first. */
int cgs_voronoi_density_rule(const double* pts, long long np, int rule, int average, int sx, double cutoff,
                             float* dcf);
/* The same with the reference's rule (CGS_VORONOI_HULL). */
int cgs_voronoi_volumes(const double* pts, long long np, double* vol);
int cgs_voronoi_density(const double* pts, long long np, int average, int sx, double cutoff, float* dcf);
/* Voronoi weights (CGS_VORONOI_CLIPPED) over the active samples of one bin
 * become that bin's density weights (per-bin DCF of the 4D workloads). */
int cgs_set_bin_voronoi(cgs_ctx* h, int bin);

/* ---- differentiable operators on device memory ----
 * For a binding that differentiates through E and E^H (unrolled CG).  All
 * pointers below are device pointers; calls are ordered on the context's
 * stream (see cgs_set_stream) and return without synchronising.  Layouts:
 *   x     image / CG vector, Ng complex ([x][y][z])
 *   y     samples [m][C] complex (coil fastest), in the point set's order
 *   k     coordinates [m][3] float32 in [0,1] per axis (as Trajectory.h:32)
 *   gk    [m][3] float32
 * E and E^H are exactly Reco::nufftFWD / nufftBWD with the context's maps,
 * kernel and deapodization (Reco.cpp:468-553), so <E x, y> * 1e-3/sqrt(Ng)
 * = <x, E^H y>.  No counterpart in the reference, which has no gradients. */
typedef struct cgs_pts cgs_pts;
/* Run every later call of the context on `stream` (a hipStream_t of the
 * caller, default priority; NULL = the context's own stream). */
int cgs_set_stream(cgs_ctx* h, void* stream);
/* A point set on the context's grid; destroy it before the context. */
int cgs_pts_create(cgs_ctx* h, cgs_pts** out);
int cgs_pts_destroy(cgs_pts* p);
/* Coordinates of m samples: binning, tile sort, KB value and KB derivative
 * tables (the sample-to-grid rules of ConvKernel::calculateWeights,
 * Reco.cpp:698-787).  Every sample is active. */
int cgs_pts_set_coords(cgs_pts* p, long long m, const float* k_dev);
int cgs_pts_count(const cgs_pts* p, long long* m);
/* y = E x (Reco::nufftFWD of the coil images x * S_c). */
int cgs_pts_forward(cgs_pts* p, const float* x_dev, float* y_dev);
/* x = E^H y (Reco::nufftBWD + conj(S_c) coil sum). */
int cgs_pts_adjoint(cgs_pts* p, const float* y_dev, float* x_dev);
/* gk[i] = d/dk_i Re<gy, E x> = sum_c sum_taps Re(conj(gy[i][c]) G_c[tap]) dW/dk_i,
 * G_c = FFT(kernelFT * S_c * x), with the analytic derivative of the KB
 * kernel (Reco.h:218-242) over the kept taps. */
int cgs_pts_grad_coords_fwd(cgs_pts* p, const float* x_dev, const float* gy_dev, float* gk_dev);
/* gk[i] = d/dk_i Re<gx, E^H y>. */
int cgs_pts_grad_coords_adj(cgs_pts* p, const float* y_dev, const float* gx_dev, float* gk_dev);

/* ---- Fourier (Fourier.h:12-54) on one complex grid ---- */
int cgs_fft_create(const int dims[3], cgs_fft** out);
int cgs_fft_destroy(cgs_fft* f);
/* Fourier::transformFWD (Fourier.cpp:19-24), in place on a host array. */
int cgs_fft_forward(cgs_fft* f, float* a);
/* Fourier::transformBWD (Fourier.cpp:26-42), in place on a host array. */
int cgs_fft_backward(cgs_fft* f, float* a, int rescale);

/* Centred 3D FFT of one complex grid held in device memory, in place, on
 * `stream` (a hipStream_t; NULL = the default stream): dir -1 is
 * Fourier::transformFWD, dir +1 Fourier::transformBWD with rescale as there
 * (Fourier.cpp:19-42).  Allocates nothing: the twiddle tables live in a fixed
 * 12 KiB device array that each GPU holds for the life of the process, so
 * callers that cannot allocate (code under HIP graph capture, serving loops,
 * harnesses that forbid allocation) can use it.  Every dimension must be a
 * supported length of at most 512 (else CGS_ERR_ARG).  Calls that overlap on
 * one GPU share that array and race.  No counterpart in the reference. */
int cgs_fft_device_noalloc(const int dims[3], float* a_dev, int dir, int rescale, void* stream);

/* ---- timing ---- */
/* Per-kernel HIP-event timing on the context's stream (off by default). */
int cgs_profile_enable(cgs_ctx* h, int on);
/* Writes "name count total_ms\n" lines; returns the needed length. */
int cgs_profile_report(cgs_ctx* h, char* buf, int len);
int cgs_profile_reset(cgs_ctx* h);
/* Device bytes held by the context. */
long long cgs_device_bytes(const cgs_ctx* h);

#ifdef __cplusplus
}
#endif

#endif /* CGS_H */
