// cgs_kernels.h -- kernel argument blocks and host-side launchers of
// libcgs_hip.so.  Everything here is internal: the public surface is the
// C-ABI in include/cgs.h.
#pragma once
#include <hip/hip_runtime.h>
#include "cgs_internal.h"

namespace cgs {

// ------------------------------------------------------------------ FFT
// Strided line pass: element (outer o, line index l, column c) lives at
// g[o*os + l*ls + c], c < ncols.
struct PlainArgs {
    float2* g;
    long long os, ls, ncols;
    const float2* tw;      // exp(-2 pi i n / L), n < L
    long long outer_per_bin;
    const int* done;       // per-bin "solve finished" flags of the launch's bins, or null
    // Column groups (the columns one block transforms) of a bin's x pass that
    // its samples never reach can be skipped (outer = bin, x passes only):
    const int* glist;      // forward: [bin][gstride] numbers of the groups to transform
    const int* gcnt;       // forward: [bin] their count; the launch has nouter * gstride blocks
    long long gstride;
    const uint8_t* gact;   // inverse: [bin][groups] 1 = transform, 0 = write zeros (no reads)
};

// Column group of the strided passes: W columns per chunk, cpb chunks per block.
void plain_geometry(int L, long long ls, int* W, int* cpb);

// z pass (lines along z, columns = coils) with fused prologue/epilogue.
struct ZArgs {
    float2* g;             // [bin][x][y][z][C], x planes xp apart, bins bp apart
    const float2* tw;
    const float2* p;       // [bin][Ng] image vector (zpre expand; zpost <p,q>)
    const float2* maps;    // [Ng][C]
    const float* kft;      // [Ng] deapodization map (ConvKernel::kernelFT)
    float2* q;             // [bin][Ng] (zpost combine)
    double* part;          // [nblocks][2] partial <p, q> (zpost combine), or null
    float scale;           // zpost: 1e-3 / sqrt(Ng)
    int C;
    long long ng;
    long long xp, bp;      // coil-grid x-plane and bin pitch (elements)
    long long nxy;         // Nx*Ny
    int ny;
    int crop;              // reduceFOV active
    int lo[3], hi[3];      // kept box, inclusive on both ends
    const int* done;       // per-bin flags of the launch's bins, or null
    int nb;                // bins of the launch (workgroup order, zpass_line)
    int bpw;               // bins per workgroup of the inverse z pass (k_fft_zpost_mb; < 2: k_fft_zpost)
};

// Inverse x pass with gridding fused in (k_fft_xgrid): the x pass of the
// launch's bins (x.outer = bin), each chunk formed from its cell's listed
// samples instead of read.
struct XGridArgs {
    PlainArgs x;
    const float2* Y;        // [Mtot][C] sample values
    const float* gdcf;      // multiply Y by these weights first (or null), as gridding does
    const int* rec;         // sorted per-cell lists: [entries][xrec_stride(KW)] records
    const long long* off;   // [nbins_total][ny * nz + 1] list positions
    int C, ny, nz;
    float thrf;             // KbParams::thr_up
    int b0;                 // first bin of the launch
};
bool xgrid_supported(int L, int C, long long ls);
hipError_t launch_fft_xgrid(int L, int kw, const XGridArgs& xa, long long nouter, hipStream_t s);

bool fft_len_supported(int L);
int zpass_w(int L, int C);
hipError_t launch_fft_plain(int L, int dir, const PlainArgs& a, long long nouter, hipStream_t s);
hipError_t launch_fft_z1(int L, int dir, float2* g, long long nlines, const float2* tw, hipStream_t s);
hipError_t launch_fft_zpre(int L, int W, const ZArgs& a, long long nouter, hipStream_t s);
hipError_t launch_fft_zpost(int L, int W, const ZArgs& a, long long nouter, hipStream_t s);

// ------------------------------------------------------ KB setup / binning
// Per-axis constants of Reco::ConvKernel::convFunction (Reco.h:221-226), all
// computed in float exactly as the reference does.
struct AxisConst {
    float dkf;     // float(1.0 / float(N-1))
    float cut;     // float(w/2) * dkf
    float wdk;     // w * dkf
    float denom;   // (100 * w) * dkf
};

struct KbParams {
    AxisConst ax[3];
    double b;      // double(float b)
    double thr;    // double(float(convFunction(0))) * 1e-7 (Reco.cpp:716-717)
    float thr_up;  // smallest float above thr: for a float w, w >= thr_up <=> double(w) > thr
    int h;         // window half width (taps at nearest-h .. nearest+h)
    int kw;        // 2h+1
    int order;     // reference search order (Reco.cpp:380)
};

hipError_t launch_bin_keys(const double* traj, long long ns, const float2* data0, const uint8_t* mask,
                           Geom g, Tiles t, int* key, int* nearest, int* counts, hipStream_t s);
hipError_t launch_scan_offsets(const int* counts, int n, int* off, hipStream_t s);
hipError_t launch_scatter(const int* key, long long ns, const int* off, int* cursor, int* tmp, hipStream_t s);
hipError_t launch_segment_rank(const int* tmp, const int* off, int ntiles, int* perm, hipStream_t s);
hipError_t launch_sample_tables(const double* traj, const int* perm, long long m, Geom g, KbParams kb,
                                int4* base, double* vals, hipStream_t s);
hipError_t launch_kernel_ft(const double* axr, const double* axi, Geom g, float* kft, hipStream_t s);
hipError_t launch_sample_dtables(const double* traj, const int* perm, long long m, Geom g, KbParams kb,
                                 double* dvals, hipStream_t s);
hipError_t launch_widen_coords(const float* k32, long long n, double* k64, hipStream_t s);
hipError_t launch_permute_rows(const float2* in, const int4* base, long long m, int C, int to_caller, float2* out,
                               hipStream_t s);
hipError_t launch_gather_data(const float2* data, long long ns, const int4* base, long long m, int C,
                              const float* dcf_full, const int* perm, float* dcf_sorted, float2* y, hipStream_t s);
hipError_t launch_scatter_data(const float2* y, const int4* base, long long m, int C, long long ns,
                               float2* data, hipStream_t s);
hipError_t launch_transpose_coils(const float2* in, float2* out, long long ng, int C, long long nyz, long long xp,
                                  int to_interleaved, hipStream_t s);

// ---------------------------------------------------------- gridding
// Samples of all bins are concatenated; bin b owns [bin_off[b], bin_off[b+1]).
// A launch covers the bins b0 .. b0+nb-1, whose samples are contiguous.
struct SampleArgs {
    const int4* base;      // [Mtot] (bx, by, bz, original sample index)
    const double* vals;    // [Mtot][val_stride(kw)]
    const float* dcf;      // [Mtot] or null (density weighting applied by degrid)
    const int* sbin;       // [Mtot] global bin of each sample
    long long first;       // first sample of the launch
    long long m;           // samples in the launch
};

struct GridArgs {
    float2* G;             // [nb][Ng][C] coil grids of the launch's bins
    float2* Y;             // [Mtot][C] sample values (global indexing)
    SampleArgs s;
    Geom g;
    Tiles t;
    const int* tile_off;   // [nbins_total][ntiles+1] absolute sample offsets
    const float* gdcf;     // gridding only: multiply Y by these weights first (or null)
    double thr;
    float thrf;            // KbParams::thr_up
    int b0, nb;
    const int* done;       // [nbins_total] flags, or null
    const uint8_t* tneed;  // gridding: [nbins_total][nty*ntz] tile columns to write (others are
                           // never read: their x-pass groups are zero-filled), or null = all
    // gridding: per tile the samples whose window meets it, in the order the
    // kernel's own scan of the 27 neighbouring tiles finds them (built once
    // per packed bin set by launch_grid_lists), or null = scan every launch
    const int* tlist;      // sample indices (global)
    const long long* tlist_off;   // [nbins_total * ntiles + 1]
    int c0;                // gridding: first coil of this launch (coils in groups of at most 32)
};

hipError_t launch_degrid(int kw, const GridArgs& a, hipStream_t s);
hipError_t launch_grid(int kw, int tcfg, const GridArgs& a, hipStream_t s);
// Candidate lists of every tile of bins b0..b0+nb-1 (a.b0, a.nb): with
// counts != null the number per tile goes to counts[bin * ntiles + tile];
// otherwise the lists are written to a.tlist at a.tlist_off.
hipError_t launch_grid_lists(int kw, int tcfg, const GridArgs& a, int* counts, hipStream_t s);

// Per-cell sample lists of the fused gridding (cgs_xlists.hip): keys
// (y * nz + z) << kXListBits | (bx + kXListOff), sorted, and per sorted
// entry a record of xrec_stride(kw) ints: bx, the kw weights along x of the
// entry's cell (float bits, w[ox] = float((vx[ox] * vy) * vz)) and the
// sample's global index; off[bin][cell] (absolute positions) of ny * nz + 1
// cells.
constexpr int kXListBits = 12;
constexpr int kXListOff = 16;
__host__ __device__ constexpr int xrec_stride(int kw) { return (kw + 2 + 3) / 4 * 4; }
bool xlist_supported(int nx, int ny, int nz);
// hipCUB radix-sort workspace for n keys
size_t xlist_sort_bytes(long long n);
// The lists of one bin: its m samples from global index first; keys_tmp /
// idx_tmp / keys_out / idx_out of m * kw * kw entries (work space); the
// records go to rec (same number of records) and the cell offsets, plus
// off0, to off.
hipError_t build_xlists_bin(int kw, const int4* base, const double* vals, long long first, long long m, int ny, int nz,
                            unsigned* keys_tmp, int* idx_tmp, void* sort_tmp, size_t sort_bytes, unsigned* keys_out,
                            int* idx_out, int* rec, long long* off, long long off0, hipStream_t s);

// Coordinate gradient of the sample values: for every sample i (sorted
// position s, original index base[s].w)
//   gk[i][axis] = scale * sum_c sum_taps Re(conj(B[s][c]) * A[tap][c]) * dW_tap/dk_axis
// over the kept taps (W > thr), dW/dk from the per-axis value and derivative
// tables.  Serves both d<gy, degrid(A)>/dk (B = gy) and d<A, grid(B)>/dk.
struct GradKArgs {
    const float2* A;       // [Ng][C] grid
    const float2* B;       // [m][C] sample values, sorted order
    const int4* base;
    const double* vals;
    const double* dvals;
    long long m;
    Geom g;
    double thr;
    float thrf;            // KbParams::thr_up
    float scale;
    float* gk;             // [m][3], original order
};
hipError_t launch_grad_k(int kw, const GradKArgs& a, hipStream_t s);
int grid_tile_configs();
void grid_tile_dims(int tcfg, int* tx, int* ty, int* tz);

// ---------------------------------------------------------------- CG
struct CgScalars {
    double rr[2];      // <r, r>
    double pq[2];      // <p, q>
    double alpha[2];
    double beta[2];
    double resid;      // ||r||_2 after the last update
    int iter;          // completed iterations
    int done;          // no further iterations for this bin
    int nonfinite;     // q[0] was NaN/Inf (Reco.cpp:264-268)
    int pad;
};

hipError_t launch_dot_partials(const float2* a, const float2* b, long long n, int nbins, int nblk, double* part,
                               hipStream_t s);
hipError_t launch_cg_init(CgScalars* sc, const double* part, int nblk, int nbins, double thresh, int* done,
                          hipStream_t s);
hipError_t launch_cg_alpha(CgScalars* sc, const double* part, int nblk, int nbins, const float2* q,
                           long long ng, int* done, hipStream_t s);
hipError_t launch_cg_update_xr(float2* x, float2* r, const float2* p, const float2* q, long long n, int nbins,
                               int nblk, const CgScalars* sc, double* part, hipStream_t s);
hipError_t launch_cg_beta(CgScalars* sc, const double* part, int nblk, int nbins, double thresh, int max_iter,
                          double* resid_hist, int hist_stride, int* done, hipStream_t s);
hipError_t launch_cg_update_p(float2* p, const float2* r, long long n, int nbins, const CgScalars* sc,
                              hipStream_t s);
hipError_t launch_copy_rhs(float2* x, float2* r, float2* p, const float2* b, long long n, int nbins, hipStream_t s);
hipError_t launch_fill_zero(float2* a, long long n, hipStream_t s);
hipError_t launch_coil_combine(const float2* G, const float2* maps, float2* q, long long ng, int C, long long nyz,
                               long long xp, hipStream_t s);
hipError_t launch_scale_d(float2* a, long long n, double f, hipStream_t s);

// ---------------------------------------------------------------- Voronoi
// Uniform hash of the points: cell (x, y, z) holds idx[start[c] .. start[c+1])
// in ascending point order, c = (x*n1 + y)*n2 + z.
struct VorHash {
    double lo[3];
    double blo[3], bhi[3];   // bounding box of the points widened by one hash cell
    double wlo[3], whi[3];   // bounding box widened by kVorWideBox times the largest extent
    double h;
    int n[3];
    const int* start;
    const int* idx;
};

// Hull rule (the reference's: 0 only for unbounded cells): cells that keep a
// face of the first box are built again from a box wider than the points by
// this factor times their largest extent, visiting every hash cell.
constexpr double kVorWideBox = 65536.0;

// Voronoi cells, one wave per cell (cgs_voronoi_cell.h).  First tier (small
// limits, LDS): volumes of the listed points (list[0..n), or every point when
// list is null); points whose cell outgrew the limits are appended to
// redo[*nredo].  wide: build from the wide box (hull rule, second stage).
hipError_t launch_voronoi(const double* pts, long long n, const int* list, bool wide, const VorHash& H, double* vol,
                          int* redo, int* nredo, hipStream_t s);
// Second tier (medium limits, LDS), the same over list[0..n) (or every point).
hipError_t launch_voronoi_mid(const double* pts, long long n, const int* list, bool wide, const VorHash& H,
                              double* vol, int* redo, int* nredo, hipStream_t s);
// Last tier (large limits) over redo[0..nredo) (every point when redo is
// null) with nwaves waves, each
// building its cells in its own slot of the workspace `cells`
// (voronoi_large_cell_bytes() per wave); closed cells that still overflow
// are counted in *overflow.
size_t voronoi_large_cell_bytes();
hipError_t launch_voronoi_large(const double* pts, bool wide, const VorHash& H, const int* redo, long long nredo,
                                int nwaves, void* cells, double* vol, int* overflow, hipStream_t s);

}  // namespace cgs
