// cgs_cg.hip -- CG-SENSE vector updates and FP64 device-side scalars.
//
// Follows Reco::cgSenseDynamic (Reco.cpp:175-321): complex CG on
// A = E^H W E with FP32 vectors and FP64 scalars.  Inner products are
// Reco::conjScalarProduct (sum conj(x) y, Reco.cpp:131-143) accumulated in
// FP64 from per-block partial sums in a fixed order; alpha = rr / pq and
// beta = rr_new / rr are Reco::complexDivision (Reco.cpp:145-151).  The
// scalars never leave the device, so a whole iteration can be replayed from a
// captured graph; a per-bin "done" flag reproduces the loop condition
// (||r|| > threshold, at most max_iter + 1 passes, Reco.cpp:228-230, 308-309)
// and every kernel of a finished bin returns at once.
#include "cgs_internal.h"
#include "cgs_kernels.h"

namespace cgs {

constexpr int kRedThreads = 256;
constexpr int kScalarThreads = 1024;   // blocks that reduce partial lists

__device__ __forceinline__ void block_sum2(double& a, double& b, double* red) {
    for (int off = 32; off >= 1; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    const int wid = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) { red[2 * wid] = a; red[2 * wid + 1] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, sb = 0.0;
        for (int w = 0; w < (int)blockDim.x / 64; ++w) { sa += red[2 * w]; sb += red[2 * w + 1]; }
        a = sa;
        b = sb;
    }
}

// Partial sums of conj(a) b over blocks: grid (nblk, nbins); part[bin][blk][2].
__global__ __launch_bounds__(kRedThreads) void k_dot_partials(const float2* __restrict__ a, const float2* __restrict__ b,
                                                              long long n, int nblk, double* __restrict__ part) {
    __shared__ double red[2 * (kRedThreads / 64)];
    const int bin = blockIdx.y;
    const float2* A = a + (long long)bin * n;
    const float2* B = b + (long long)bin * n;
    double sr = 0.0, si = 0.0;
    for (long long i = (long long)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (long long)nblk * kRedThreads) {
        const float2 x = A[i], y = B[i];
        sr += (double)x.x * (double)y.x + (double)x.y * (double)y.y;
        si += (double)x.x * (double)y.y - (double)x.y * (double)y.x;
    }
    block_sum2(sr, si, red);
    if (threadIdx.x == 0) {
        part[((long long)bin * nblk + blockIdx.x) * 2] = sr;
        part[((long long)bin * nblk + blockIdx.x) * 2 + 1] = si;
    }
}

// Sum of nblk (re, im) partials of one bin by one block.  Each thread keeps
// eight independent accumulators (eight loads in flight; the partial lists
// reach 10^5 entries), combined in a fixed order, so the result does not
// depend on timing.
__device__ void reduce_parts(const double* part, int nblk, int bin, double* out, double* red) {
    constexpr int U = 8;
    double a[U], b[U];
#pragma unroll
    for (int u = 0; u < U; ++u) a[u] = b[u] = 0.0;
    const double2* pp = (const double2*)part + (long long)bin * nblk;
    const int step = blockDim.x * U;
    int i0 = threadIdx.x;
    for (; i0 + (U - 1) * (int)blockDim.x < nblk; i0 += step) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double2 v = pp[i0 + u * blockDim.x];
            a[u] += v.x;
            b[u] += v.y;
        }
    }
    for (int i = i0; i < nblk; i += blockDim.x) {
        const double2 v = pp[i];
        a[0] += v.x;
        b[0] += v.y;
    }
    double sa = 0.0, sb = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) { sa += a[u]; sb += b[u]; }
    block_sum2(sa, sb, red);
    if (threadIdx.x == 0) { out[0] = sa; out[1] = sb; }
}

__device__ __forceinline__ void cdiv(const double* x, const double* y, double* r) {
    const double den = y[0] * y[0] + y[1] * y[1];
    r[0] = (x[0] * y[0] + x[1] * y[1]) / den;
    r[1] = (x[1] * y[0] - x[0] * y[1]) / den;
}

// rr = <r0, r0> of the right-hand side; reset the per-bin state.
__global__ __launch_bounds__(kScalarThreads) void k_cg_init(CgScalars* sc, const double* part, int nblk, double thresh,
                                                          int* done) {
    __shared__ double red[2 * (kScalarThreads / 64)];
    const int bin = blockIdx.x;
    double rr[2];
    reduce_parts(part, nblk, bin, rr, red);
    if (threadIdx.x == 0) {
        CgScalars& s = sc[bin];
        s.rr[0] = rr[0];
        s.rr[1] = rr[1];
        s.iter = 0;
        s.resid = 1000.0;                // Reco.cpp:229
        s.done = !(1000.0 > thresh);
        s.nonfinite = 0;
        if (done) done[bin] = s.done;
    }
}

// pq = <p, q>; NaN/Inf check of q[0] (Reco.cpp:264-268); alpha = rr / pq.
__global__ __launch_bounds__(kScalarThreads) void k_cg_alpha(CgScalars* sc, const double* part, int nblk, const float2* q,
                                                           long long ng, int* done) {
    __shared__ double red[2 * (kScalarThreads / 64)];
    const int bin = blockIdx.x;
    CgScalars& s = sc[bin];
    if (s.done) return;
    double pq[2];
    reduce_parts(part, nblk, bin, pq, red);
    if (threadIdx.x == 0) {
        const float2 q0 = q[(long long)bin * ng];
        if (!isfinite(q0.x) || !isfinite(q0.y)) {
            s.nonfinite = 1;
            s.done = 1;
            if (done) done[bin] = 1;
            return;
        }
        s.pq[0] = pq[0];
        s.pq[1] = pq[1];
        cdiv(s.rr, pq, s.alpha);
    }
}

// x += alpha p; r -= alpha q (Reco.cpp:284-290) and partial sums of |r|^2.
__global__ __launch_bounds__(kRedThreads) void k_cg_update_xr(float2* __restrict__ x, float2* __restrict__ r,
                                                              const float2* __restrict__ p, const float2* __restrict__ q,
                                                              long long n, int nblk, const CgScalars* sc,
                                                              double* __restrict__ part) {
    __shared__ double red[2 * (kRedThreads / 64)];
    const int bin = blockIdx.y;
    const CgScalars& s = sc[bin];
    if (s.done) return;
    const double a0 = s.alpha[0], a1 = s.alpha[1];
    const long long o = (long long)bin * n;
    double rr = 0.0, dummy = 0.0;
    for (long long i = (long long)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (long long)nblk * kRedThreads) {
        const float2 pv = p[o + i], qv = q[o + i];
        float2 xv = x[o + i], rv = r[o + i];
        xv.x = (float)((double)xv.x + (a0 * (double)pv.x - a1 * (double)pv.y));
        xv.y = (float)((double)xv.y + (a0 * (double)pv.y + a1 * (double)pv.x));
        rv.x = (float)((double)rv.x + (-a0 * (double)qv.x + a1 * (double)qv.y));
        rv.y = (float)((double)rv.y + (-a0 * (double)qv.y - a1 * (double)qv.x));
        x[o + i] = xv;
        r[o + i] = rv;
        rr += (double)rv.x * (double)rv.x + (double)rv.y * (double)rv.y;
    }
    block_sum2(rr, dummy, red);
    if (threadIdx.x == 0) {
        part[((long long)bin * nblk + blockIdx.x) * 2] = rr;
        part[((long long)bin * nblk + blockIdx.x) * 2 + 1] = 0.0;
    }
}

// rr_new = <r, r>; residual ||r|| (Reco.cpp:292-299); beta = rr_new / rr
// (Reco.cpp:303-304); loop control (Reco.cpp:230, 308-309).
__global__ void k_cg_beta(CgScalars* sc, const double* part, int nblk, double thresh, int max_iter,
                          double* resid_hist, int hist_stride, int* done) {
    __shared__ double red[2 * (kRedThreads / 64)];
    const int bin = blockIdx.x;
    CgScalars& s = sc[bin];
    if (s.done) return;
    double rn[2];
    reduce_parts(part, nblk, bin, rn, red);
    if (threadIdx.x == 0) {
        const double resid = sqrt(rn[0]);
        if (resid_hist) resid_hist[(long long)bin * hist_stride + s.iter] = resid;
        s.resid = resid;
        const double rnc[2] = {rn[0], 0.0};
        cdiv(rnc, s.rr, s.beta);
        s.rr[0] = rn[0];
        s.rr[1] = 0.0;
        s.iter += 1;
        if (!(resid > thresh) || s.iter > max_iter) {
            s.done = 1;     // p is not needed any more: the update below is skipped
            if (done) done[bin] = 1;
        }
    }
}

// p = r + beta p (Reco.cpp:163-173, 306).
__global__ __launch_bounds__(kRedThreads) void k_cg_update_p(float2* __restrict__ p, const float2* __restrict__ r,
                                                             long long n, CgScalars* sc) {
    const int bin = blockIdx.y;
    const CgScalars& s = sc[bin];
    if (s.done) return;
    const double b0 = s.beta[0], b1 = s.beta[1];
    const long long o = (long long)bin * n;
    for (long long i = (long long)blockIdx.x * kRedThreads + threadIdx.x; i < n;
         i += (long long)gridDim.x * kRedThreads) {
        const float2 rv = r[o + i];
        const float2 pv = p[o + i];
        const double pr = pv.x, pi = pv.y;
        p[o + i] = make_float2((float)((double)rv.x + b0 * pr - b1 * pi), (float)((double)rv.y + b0 * pi + b1 * pr));
    }
}

// r = p = b, x = 0 (Reco.cpp:211-221).
__global__ void k_copy_rhs(float2* x, float2* r, float2* p, const float2* b, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 v = b[i];
    r[i] = v;
    p[i] = v;
    x[i] = make_float2(0.0f, 0.0f);
}

__global__ void k_fill_zero(float2* a, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = make_float2(0.0f, 0.0f);
}

// q[cell] = sum_c conj(S[cell][c]) G[cell][c] -- the initial image of
// Reco::cgSenseDynamic (Reco.cpp:183-190) from coil images (G's x planes xp
// elements apart, nyz = ny*nz cells per plane).
__global__ void k_coil_combine(const float2* __restrict__ G, const float2* __restrict__ maps, float2* __restrict__ q,
                               long long ng, int C, long long nyz, long long xp) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ng) return;
    const long long gi = (i / nyz) * xp + (i % nyz) * C;
    float2 a = make_float2(0.0f, 0.0f);
    for (int c = 0; c < C; ++c) {
        const float2 s = maps[i * C + c], v = G[gi + c];
        a.x += s.x * v.x + s.y * v.y;
        a.y += s.x * v.y - s.y * v.x;
    }
    q[i] = a;
}

// a *= f with the product formed in double (Fourier.cpp:34-41).
__global__ void k_scale_d(float2* a, long long n, double f) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 v = a[i];
    a[i] = make_float2((float)((double)v.x * f), (float)((double)v.y * f));
}

hipError_t launch_coil_combine(const float2* G, const float2* maps, float2* q, long long ng, int C, long long nyz,
                               long long xp, hipStream_t s) {
    k_coil_combine<<<(unsigned)((ng + 255) / 256), 256, 0, s>>>(G, maps, q, ng, C, nyz, xp);
    return hipGetLastError();
}

hipError_t launch_scale_d(float2* a, long long n, double f, hipStream_t s) {
    k_scale_d<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(a, n, f);
    return hipGetLastError();
}

hipError_t launch_dot_partials(const float2* a, const float2* b, long long n, int nbins, int nblk, double* part,
                               hipStream_t s) {
    k_dot_partials<<<dim3(nblk, nbins), kRedThreads, 0, s>>>(a, b, n, nblk, part);
    return hipGetLastError();
}

hipError_t launch_cg_init(CgScalars* sc, const double* part, int nblk, int nbins, double thresh, int* done,
                          hipStream_t s) {
    k_cg_init<<<nbins, kScalarThreads, 0, s>>>(sc, part, nblk, thresh, done);
    return hipGetLastError();
}

hipError_t launch_cg_alpha(CgScalars* sc, const double* part, int nblk, int nbins, const float2* q, long long ng,
                           int* done, hipStream_t s) {
    k_cg_alpha<<<nbins, kScalarThreads, 0, s>>>(sc, part, nblk, q, ng, done);
    return hipGetLastError();
}

hipError_t launch_cg_update_xr(float2* x, float2* r, const float2* p, const float2* q, long long n, int nbins,
                               int nblk, const CgScalars* sc, double* part, hipStream_t s) {
    k_cg_update_xr<<<dim3(nblk, nbins), kRedThreads, 0, s>>>(x, r, p, q, n, nblk, sc, part);
    return hipGetLastError();
}

hipError_t launch_cg_beta(CgScalars* sc, const double* part, int nblk, int nbins, double thresh, int max_iter,
                          double* resid_hist, int hist_stride, int* done, hipStream_t s) {
    k_cg_beta<<<nbins, kRedThreads, 0, s>>>(sc, part, nblk, thresh, max_iter, resid_hist, hist_stride, done);
    return hipGetLastError();
}

hipError_t launch_cg_update_p(float2* p, const float2* r, long long n, int nbins, const CgScalars* sc,
                              hipStream_t s) {
    const int nblk = (int)((n + kRedThreads * 8 - 1) / (kRedThreads * 8));
    k_cg_update_p<<<dim3(nblk, nbins), kRedThreads, 0, s>>>(p, r, n, (CgScalars*)sc);
    return hipGetLastError();
}

hipError_t launch_copy_rhs(float2* x, float2* r, float2* p, const float2* b, long long n, int nbins, hipStream_t s) {
    const long long tot = n * nbins;
    k_copy_rhs<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(x, r, p, b, tot);
    return hipGetLastError();
}

hipError_t launch_fill_zero(float2* a, long long n, hipStream_t s) {
    k_fill_zero<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(a, n);
    return hipGetLastError();
}

}  // namespace cgs
