// cgs_setup.hip -- per-trajectory and per-dataset setup kernels: the integer
// sample-to-tile binning pass, the stable tile sort, the per-sample KB factor
// tables, the deapodization map and the data gather/scatter.
//
// Built with -ffp-contract=off: every multiply and add below is rounded on
// its own, so the tables are bit-identical to the reference's arithmetic
// (Reco.cpp:698-787, Reco.h:218-242), which MSVC /fp:precise evaluates
// without contraction.
#include "cgs_internal.h"
#include "cgs_kernels.h"

namespace cgs {

// Modified Bessel I0 by its power series (Abramowitz & Stegun 9.6.12),
// standing in for boost::math::cyl_bessel_i(0, x) at Reco.h:238-240.  The
// operation sequence and the stopping rule are fixed so the result does not
// depend on the device.
__device__ double i0_series(double x) {
    const double q = (0.25 * x) * x;
    double term = 1.0;
    double sum = 1.0;
    for (int k = 1; k < 500; ++k) {
        const double kk = (double)k * (double)k;
        term = (term * q) / kk;
        sum = sum + term;
        if (term < sum * 1.0e-17) break;
    }
    return sum;
}

// One axis factor of convFunction (Reco.h:228-240) for a float distance.
__device__ double kb_axis(float dxf, const AxisConst& c, double b) {
    if (dxf > c.cut) return 0.0;
    const float q = (2.0f * dxf) / c.wdk;
    const double q2 = (double)q * (double)q;
    const double arg = b * sqrt(1.0 - q2);
    return i0_series(arg) / (double)c.denom;
}

// Nearest grid point (int)((N-1)*k + 0.5) in double (Reco.cpp:751-755).
__device__ __forceinline__ int nearest_pt(int n, double k) {
    return (int)((double)(n - 1) * k + 0.5);
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Integer binning pass: active flag (coil-0 value not exactly 0+0i,
// Reco.cpp:681-696, or an explicit mask), nearest grid point, tile key.
__global__ void k_bin_keys(const double* __restrict__ traj, long long ns, const float2* __restrict__ data0,
                           const uint8_t* __restrict__ mask, Geom g, Tiles t, int* __restrict__ key,
                           int* __restrict__ nearest, int* __restrict__ counts) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    const int nx = nearest_pt(g.nx, traj[3 * i + 0]);
    const int ny = nearest_pt(g.ny, traj[3 * i + 1]);
    const int nz = nearest_pt(g.nz, traj[3 * i + 2]);
    if (nearest) {
        nearest[3 * i + 0] = nx;
        nearest[3 * i + 1] = ny;
        nearest[3 * i + 2] = nz;
    }
    bool active = true;
    if (data0) {
        const float2 d = data0[i];
        active = !(d.x == 0.0f && d.y == 0.0f);
    } else if (mask) {
        active = mask[i] != 0;
    }
    int k = -1;
    if (active) {
        const int tx = clampi(nx, 0, g.nx - 1) / t.tx;
        const int ty = clampi(ny, 0, g.ny - 1) / t.ty;
        const int tz = clampi(nz, 0, g.nz - 1) / t.tz;
        k = (tx * t.nty + ty) * t.ntz + tz;
        atomicAdd(&counts[k], 1);
    }
    key[i] = k;
}

// Exclusive scan of n counts into off[0..n] (single block).
__global__ void k_scan_offsets(const int* __restrict__ counts, int n, int* __restrict__ off) {
    __shared__ int part[1024];
    const int tid = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int lo = tid * per;
    int s = 0;
    for (int i = lo; i < lo + per && i < n; ++i) s += counts[i];
    part[tid] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const int v = tid >= d ? part[tid - d] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = tid ? part[tid - 1] : 0;
    for (int i = lo; i < lo + per && i < n; ++i) {
        off[i] = run;
        run += counts[i];
    }
    if (tid == 1023) off[n] = part[1023];
}

__global__ void k_scatter(const int* __restrict__ key, long long ns, const int* __restrict__ off,
                          int* __restrict__ cursor, int* __restrict__ tmp) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    const int k = key[i];
    if (k < 0) return;
    const int pos = off[k] + atomicAdd(&cursor[k], 1);
    tmp[pos] = (int)i;
}

// Stable order inside each tile: an element's rank is the number of
// elements of its tile with a smaller original index, so the final order is
// the original sample order (what a stable counting sort gives).
__global__ void k_segment_rank(const int* __restrict__ tmp, const int* __restrict__ off, int* __restrict__ perm) {
    const int t = blockIdx.x;
    const int b = off[t], n = off[t + 1] - b;
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
        const int v = tmp[b + e];
        int r = 0;
        for (int j = 0; j < n; ++j) r += tmp[b + j] < v;
        perm[b + r] = v;
    }
}

// Per-sample KB factor tables in sorted order.  For axis a the window holds
// the grid indices base_a .. base_a + kw - 1 (base = nearest - h); entries
// outside the grid (collectNeighborIndices clips, Reco.cpp:789-806) are 0.
// The distance is |k - float(idx)*dkf| in double, which is what
// sqrt((k - idx*dk)^2) of Reco.cpp:767-771 returns (sqrt of a correctly
// rounded square is exact here: no underflow for k in [0,1]).
__global__ void k_sample_tables(const double* __restrict__ traj, const int* __restrict__ perm, long long m,
                                Geom g, KbParams kb, int4* __restrict__ base, double* __restrict__ vals) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m) return;
    const int i = perm[s];
    const int n3[3] = {g.nx, g.ny, g.nz};
    int b3[3];
    const int vs = val_stride(kb.kw);
    for (int a = 0; a < 3; ++a) {
        const double k = traj[3LL * i + a];
        const int near = nearest_pt(n3[a], k);
        b3[a] = near - kb.h;
        for (int o = 0; o < kb.kw; ++o) {
            const int idx = b3[a] + o;
            double v = 0.0;
            if (idx >= 0 && idx <= n3[a] - 1) {
                const double e = k - (double)((float)idx * kb.ax[a].dkf);
                v = kb_axis((float)fabs(e), kb.ax[a], kb.b);
            }
            vals[s * vs + a * kb.kw + o] = v;
        }
    }
    for (int o = 3 * kb.kw; o < vs; ++o) vals[s * vs + o] = 0.0;
    base[s] = make_int4(b3[0], b3[1], b3[2], i);
}

// I1(x) / x by its power series (Abramowitz & Stegun 9.6.10 divided by x):
// (1/2) sum_k (x^2/4)^k / (k! (k+1)!), finite at x = 0.
__device__ double i1_over_x(double x) {
    const double q = (0.25 * x) * x;
    double term = 0.5;
    double sum = 0.5;
    for (int k = 1; k < 500; ++k) {
        term = (term * q) / ((double)k * (double)(k + 1));
        sum = sum + term;
        if (term < sum * 1.0e-17) break;
    }
    return sum;
}

// d/dk of one axis factor of convFunction (Reco.h:228-240) at offset
// e = k - idx*dk, taken analytically: with q = 2|e|/(w dk) and
// beta = b sqrt(1 - q^2), f = I0(beta)/denom and
//   df/dk = -sign(e) (2 b^2 q / (w dk)) (I1(beta)/beta) / denom,
// which stays finite at the window edge (beta -> 0).  0 beyond the cut-off,
// like the factor itself.
__device__ double kb_axis_dk(double e, const AxisConst& c, double b) {
    const float dxf = (float)fabs(e);
    if (dxf > c.cut) return 0.0;
    const double q = (double)((2.0f * dxf) / c.wdk);
    const double beta = b * sqrt(1.0 - q * q);
    const double g = -((2.0 * b * b * q) / (double)c.wdk) * i1_over_x(beta) / (double)c.denom;
    return e < 0.0 ? -g : g;
}

// Per-axis KB derivative table of the sorted samples, laid out like vals.
__global__ void k_sample_dtables(const double* __restrict__ traj, const int* __restrict__ perm, long long m,
                                 Geom g, KbParams kb, double* __restrict__ dvals) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m) return;
    const int i = perm[s];
    const int n3[3] = {g.nx, g.ny, g.nz};
    const int vs = val_stride(kb.kw);
    for (int a = 0; a < 3; ++a) {
        const double k = traj[3LL * i + a];
        const int b0 = nearest_pt(n3[a], k) - kb.h;
        for (int o = 0; o < kb.kw; ++o) {
            const int idx = b0 + o;
            double v = 0.0;
            if (idx >= 0 && idx <= n3[a] - 1) {
                const double e = k - (double)((float)idx * kb.ax[a].dkf);
                v = kb_axis_dk(e, kb.ax[a], kb.b);
            }
            dvals[s * vs + a * kb.kw + o] = v;
        }
    }
    for (int o = 3 * kb.kw; o < vs; ++o) dvals[s * vs + o] = 0.0;
}

__global__ void k_widen_coords(const float* __restrict__ k32, long long n, double* __restrict__ k64) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) k64[i] = (double)k32[i];
}

// Rows of C complex values between caller order (row = original index) and
// sorted order (row = sorted position s, original index base[s].w).
__global__ void k_permute_rows(const float2* __restrict__ in, const int4* __restrict__ base, long long m, int C,
                               int to_caller, float2* __restrict__ out) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * C) return;
    const long long s = e / C;
    const int c = (int)(e % C);
    const long long i = base[s].w;
    if (to_caller) out[i * C + c] = in[e];
    else out[e] = in[i * C + c];
}

// kernelFT = 1 / Re(Xx Xy Xz) with the per-axis centred inverse DFTs of the
// sampled kernel (Reco.cpp:560-601), rounded to float.
__global__ void k_kernel_ft(const double* __restrict__ axr, const double* __restrict__ axi, Geom g,
                            float* __restrict__ kft) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.ng) return;
    const int z = (int)(c % g.nz);
    const int y = (int)((c / g.nz) % g.ny);
    const int x = (int)(c / ((long long)g.nz * g.ny));
    const double* xr = axr;
    const double* yr = axr + g.nx;
    const double* zr = axr + g.nx + g.ny;
    const double* xi = axi;
    const double* yi = axi + g.nx;
    const double* zi = axi + g.nx + g.ny;
    const double xyr = xr[x] * yr[y] - xi[x] * yi[y];
    const double xyi = xr[x] * yi[y] + xi[x] * yr[y];
    const double re = xyr * zr[z] - xyi * zi[z];
    kft[c] = (float)(re / (re * re));
}

// Y[m][c] = dcf[i] * data[c][i] for the sorted active samples (i = original
// index); the density weighting is performDensityCorrection (functions.h:319-335).
__global__ void k_gather_data(const float2* __restrict__ data, long long ns, const int4* __restrict__ base,
                              long long m, int C, const float* __restrict__ dcf_full, float* __restrict__ dcf_sorted,
                              float2* __restrict__ y) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * C) return;
    const long long s = e / C;
    const int c = (int)(e % C);
    const int i = base[s].w;
    float2 v = data ? data[(long long)c * ns + i] : make_float2(0.0f, 0.0f);
    if (dcf_full) {
        const float d = dcf_full[i];
        v.x *= d;
        v.y *= d;
        if (c == 0 && dcf_sorted) dcf_sorted[s] = d;
    }
    if (data) y[e] = v;
}

// data[c][i] = Y[m][c] back in the reference's sample order (inactive samples
// stay as the caller zeroed them).
__global__ void k_scatter_data(const float2* __restrict__ y, const int4* __restrict__ base, long long m, int C,
                               long long ns, float2* __restrict__ data) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * C) return;
    const long long s = e / C;
    const int c = (int)(e % C);
    data[(long long)c * ns + base[s].w] = y[e];
}

// [C][Ng] <-> [Ng][C]; the interleaved side's x planes are xp elements apart
// (nyz = ny*nz cells per plane).
__global__ void k_transpose_coils(const float2* __restrict__ in, float2* __restrict__ out, long long ng, int C,
                                  long long nyz, long long xp, int to_interleaved) {
    const long long cell = (long long)blockIdx.x * 64 + (threadIdx.x % 64);
    const int c0 = threadIdx.x / 64;
    const long long il = (cell / nyz) * xp + (cell % nyz) * C;
    for (int c = c0; c < C; c += blockDim.x / 64) {
        if (cell >= ng) continue;
        if (to_interleaved) out[il + c] = in[(long long)c * ng + cell];
        else out[(long long)c * ng + cell] = in[il + c];
    }
}

static inline unsigned nblocks(long long n, int t) { return (unsigned)((n + t - 1) / t); }

hipError_t launch_bin_keys(const double* traj, long long ns, const float2* data0, const uint8_t* mask, Geom g,
                           Tiles t, int* key, int* nearest, int* counts, hipStream_t s) {
    k_bin_keys<<<nblocks(ns, 256), 256, 0, s>>>(traj, ns, data0, mask, g, t, key, nearest, counts);
    return hipGetLastError();
}

hipError_t launch_scan_offsets(const int* counts, int n, int* off, hipStream_t s) {
    k_scan_offsets<<<1, 1024, 0, s>>>(counts, n, off);
    return hipGetLastError();
}

hipError_t launch_scatter(const int* key, long long ns, const int* off, int* cursor, int* tmp, hipStream_t s) {
    k_scatter<<<nblocks(ns, 256), 256, 0, s>>>(key, ns, off, cursor, tmp);
    return hipGetLastError();
}

hipError_t launch_segment_rank(const int* tmp, const int* off, int ntiles, int* perm, hipStream_t s) {
    k_segment_rank<<<ntiles, 256, 0, s>>>(tmp, off, perm);
    return hipGetLastError();
}

hipError_t launch_sample_tables(const double* traj, const int* perm, long long m, Geom g, KbParams kb, int4* base,
                                double* vals, hipStream_t s) {
    if (m == 0) return hipSuccess;
    k_sample_tables<<<nblocks(m, 128), 128, 0, s>>>(traj, perm, m, g, kb, base, vals);
    return hipGetLastError();
}

hipError_t launch_sample_dtables(const double* traj, const int* perm, long long m, Geom g, KbParams kb,
                                 double* dvals, hipStream_t s) {
    if (m == 0) return hipSuccess;
    k_sample_dtables<<<nblocks(m, 128), 128, 0, s>>>(traj, perm, m, g, kb, dvals);
    return hipGetLastError();
}

hipError_t launch_widen_coords(const float* k32, long long n, double* k64, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_widen_coords<<<nblocks(n, 256), 256, 0, s>>>(k32, n, k64);
    return hipGetLastError();
}

hipError_t launch_permute_rows(const float2* in, const int4* base, long long m, int C, int to_caller, float2* out,
                               hipStream_t s) {
    if (m == 0) return hipSuccess;
    k_permute_rows<<<nblocks(m * C, 256), 256, 0, s>>>(in, base, m, C, to_caller, out);
    return hipGetLastError();
}

hipError_t launch_kernel_ft(const double* axr, const double* axi, Geom g, float* kft, hipStream_t s) {
    k_kernel_ft<<<nblocks(g.ng, 256), 256, 0, s>>>(axr, axi, g, kft);
    return hipGetLastError();
}

hipError_t launch_gather_data(const float2* data, long long ns, const int4* base, long long m, int C,
                              const float* dcf_full, const int* perm, float* dcf_sorted, float2* y, hipStream_t s) {
    (void)perm;
    if (m == 0) return hipSuccess;
    k_gather_data<<<nblocks(m * C, 256), 256, 0, s>>>(data, ns, base, m, C, dcf_full, dcf_sorted, y);
    return hipGetLastError();
}

hipError_t launch_scatter_data(const float2* y, const int4* base, long long m, int C, long long ns, float2* data,
                               hipStream_t s) {
    if (m == 0) return hipSuccess;
    k_scatter_data<<<nblocks(m * C, 256), 256, 0, s>>>(y, base, m, C, ns, data);
    return hipGetLastError();
}

hipError_t launch_transpose_coils(const float2* in, float2* out, long long ng, int C, long long nyz, long long xp,
                                  int to_interleaved, hipStream_t s) {
    k_transpose_coils<<<nblocks(ng, 64), 256, 0, s>>>(in, out, ng, C, nyz, xp, to_interleaved);
    return hipGetLastError();
}

}  // namespace cgs
