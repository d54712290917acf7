// cgs_fft.hip -- centred 3D FFT as three LDS-staged Stockham line passes.
//
// Replaces Fourier::transformFWD / transformBWD (Fourier.cpp:19-42): the
// reference applies signAlternate3D (negate where x+y+z is even,
// Fourier.cpp:45-72) before and after an FFTW plan.  The two negations cancel,
// leaving diag((-1)^(x+y+z)) * DFT * diag((-1)^(x+y+z)); because that diagonal
// and the DFT both factor over the axes, every line pass multiplies its input
// and its output by (-1)^(index along its own axis).  Forward = exp(-2 pi i),
// unnormalized; inverse = exp(+2 pi i) (scaling is applied by the caller's
// fused epilogue, as Fourier.cpp:34-41 does after the transform).
//
// A pass loads a [lines x columns] tile of the coil-interleaved grid into LDS
// (columns are contiguous in HBM, so every row of the tile is one coalesced
// run), runs mixed-radix Stockham stages (radix 8/4/2 then 3/5/7) with all
// butterfly inputs held in registers between two barriers, and writes back in
// place.  The z pass (lines along the contiguous z axis, columns = coils)
// carries the fused prologue of E (p * S_c * deapodization) and the fused
// epilogue of E^H (deapodization * scale * crop, conj(S_c) multiply, sum over
// coils, partial <p, q>), so the coil images never make an extra HBM round trip.
#include "cgs_internal.h"
#include "cgs_kernels.h"

namespace cgs {

// ---------------------------------------------------------------- radix plan
__host__ __device__ constexpr int plan_nstages(int L) {
    int n = 0;
    while (L % 8 == 0) { L /= 8; ++n; }
    while (L % 4 == 0) { L /= 4; ++n; }
    while (L % 2 == 0) { L /= 2; ++n; }
    while (L % 3 == 0) { L /= 3; ++n; }
    while (L % 5 == 0) { L /= 5; ++n; }
    while (L % 7 == 0) { L /= 7; ++n; }
    return L == 1 ? n : -1;
}

__host__ __device__ constexpr int plan_radix(int L, int s) {
    int r = 0;
    for (int i = 0; i <= s; ++i) {
        if (L % 8 == 0) r = 8;
        else if (L % 4 == 0) r = 4;
        else if (L % 2 == 0) r = 2;
        else if (L % 3 == 0) r = 3;
        else if (L % 5 == 0) r = 5;
        else if (L % 7 == 0) r = 7;
        else return 0;
        L /= r;
    }
    return r;
}

// ------------------------------------------------------------ small DFTs
// DIR < 0: y_u = sum_i x_i exp(-2 pi i iu/R);  DIR > 0: exp(+...).
template <int DIR>
__device__ __forceinline__ float2 mul_mi(float2 d) {  // d * (-i) for DIR<0, d * (+i) for DIR>0
    return DIR < 0 ? make_float2(d.y, -d.x) : make_float2(-d.y, d.x);
}

template <int R, int DIR>
struct Dft;

template <int DIR>
struct Dft<2, DIR> {
    __device__ static __forceinline__ void run(float2* x) {
        const float2 a = x[0];
        x[0] = cadd(a, x[1]);
        x[1] = csub(a, x[1]);
    }
};

template <int DIR>
struct Dft<4, DIR> {
    __device__ static __forceinline__ void run(float2* x) {
        const float2 a0 = cadd(x[0], x[2]), a1 = csub(x[0], x[2]);
        const float2 a2 = cadd(x[1], x[3]), a3 = mul_mi<DIR>(csub(x[1], x[3]));
        x[0] = cadd(a0, a2);
        x[2] = csub(a0, a2);
        x[1] = cadd(a1, a3);
        x[3] = csub(a1, a3);
    }
};

template <int DIR>
struct Dft<8, DIR> {
    __device__ static __forceinline__ void run(float2* x) {
        float2 e[4] = {x[0], x[2], x[4], x[6]};
        float2 o[4] = {x[1], x[3], x[5], x[7]};
        Dft<4, DIR>::run(e);
        Dft<4, DIR>::run(o);
        const float c = 0.70710678118654752440f;
        // o[u] *= exp(DIR * 2 pi i u / 8)
        o[1] = DIR < 0 ? make_float2((o[1].x + o[1].y) * c, (o[1].y - o[1].x) * c)
                       : make_float2((o[1].x - o[1].y) * c, (o[1].y + o[1].x) * c);
        o[2] = mul_mi<DIR>(o[2]);
        o[3] = DIR < 0 ? make_float2((o[3].y - o[3].x) * c, (-o[3].y - o[3].x) * c)
                       : make_float2((-o[3].x - o[3].y) * c, (o[3].x - o[3].y) * c);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            x[u + 4] = csub(e[u], o[u]);
            x[u] = cadd(e[u], o[u]);
        }
    }
};

// Odd radix R in {3,5,7}: pairs (k, R-k) share cos(2 pi k u/R) and
// +-sin(2 pi k u/R).  COSR[m] = cos(2 pi m / R), SINR[m] = sin(2 pi m / R).
template <int R>
struct OddTab;
template <>
struct OddTab<3> {
    __device__ static __forceinline__ float c(int m) { return m == 0 ? 1.0f : -0.5f; }
    __device__ static __forceinline__ float s(int m) {
        return m == 0 ? 0.0f : (m == 1 ? 0.86602540378443864676f : -0.86602540378443864676f);
    }
};
template <>
struct OddTab<5> {
    __device__ static __forceinline__ float c(int m) {
        return m == 0 ? 1.0f : (m == 1 || m == 4) ? 0.30901699437494742410f : -0.80901699437494742410f;
    }
    __device__ static __forceinline__ float s(int m) {
        switch (m) {
            case 1: return 0.95105651629515357212f;
            case 2: return 0.58778525229247312917f;
            case 3: return -0.58778525229247312917f;
            case 4: return -0.95105651629515357212f;
            default: return 0.0f;
        }
    }
};
template <>
struct OddTab<7> {
    __device__ static __forceinline__ float c(int m) {
        switch (m) {
            case 0: return 1.0f;
            case 1: case 6: return 0.62348980185873353053f;
            case 2: case 5: return -0.22252093395631440429f;
            default: return -0.90096886790241912624f;
        }
    }
    __device__ static __forceinline__ float s(int m) {
        switch (m) {
            case 1: return 0.78183148246802980871f;
            case 2: return 0.97492791218182360702f;
            case 3: return 0.43388373911755812048f;
            case 4: return -0.43388373911755812048f;
            case 5: return -0.97492791218182360702f;
            case 6: return -0.78183148246802980871f;
            default: return 0.0f;
        }
    }
};

template <int R, int DIR>
struct DftOdd {
    __device__ static __forceinline__ void run(float2* x) {
        constexpr int H = (R - 1) / 2;
        float2 a[H + 1], b[H + 1];
        float2 y0 = x[0];
#pragma unroll
        for (int k = 1; k <= H; ++k) {
            a[k] = cadd(x[k], x[R - k]);
            b[k] = csub(x[k], x[R - k]);
            y0 = cadd(y0, a[k]);
        }
        float2 out[R];
        out[0] = y0;
#pragma unroll
        for (int u = 1; u <= H; ++u) {
            float2 re = x[0];
            float2 im = make_float2(0.0f, 0.0f);
#pragma unroll
            for (int k = 1; k <= H; ++k) {
                const int m = (k * u) % R;
                re = cadd(re, cscale(a[k], OddTab<R>::c(m)));
                im = cadd(im, cscale(b[k], OddTab<R>::s(m)));
            }
            // forward: y_u = re - i*im, y_{R-u} = re + i*im
            const float2 t = mul_mi<DIR>(im);   // -i*im (fwd) / +i*im (inv)
            out[u] = cadd(re, t);
            out[R - u] = csub(re, t);
        }
#pragma unroll
        for (int u = 0; u < R; ++u) x[u] = out[u];
    }
};
template <int DIR>
struct Dft<3, DIR> : DftOdd<3, DIR> {};
template <int DIR>
struct Dft<5, DIR> : DftOdd<5, DIR> {};
template <int DIR>
struct Dft<7, DIR> : DftOdd<7, DIR> {};

// ------------------------------------------------------- Stockham stages
// LDS tile layout: element (line index l, column col) at lds[l * STRIDE + col]
// (STRIDE = NCOL, or NCOL + 1 where rows are later reduced by one thread).
// Stage with radix R on sub-transforms of size LS (Stockham autosort):
//   butterfly j reads l = j + i*L/R, twiddles by exp(-+2 pi i k i/(LS R)) with
//   k = j mod LS, and writes l = (j/LS)*LS*R + k + i*LS.
template <int L, int R, int LS, int DIR, int NCOL, int STRIDE>
__device__ __forceinline__ void fft_stage(float2* lds, const float2* __restrict__ tw) {
    constexpr int NB = NCOL * (L / R);
    constexpr int PER = (NB + CGS_THREADS - 1) / CGS_THREADS;
    constexpr int TS = L / (LS * R);
    const int tid = threadIdx.x;
    float2 v[PER][R];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int b = tid + p * CGS_THREADS;
        if ((NB % CGS_THREADS) == 0 || b < NB) {
            const int col = b % NCOL, j = b / NCOL, k = j % LS;
#pragma unroll
            for (int i = 0; i < R; ++i) {
                float2 x = lds[(j + i * (L / R)) * STRIDE + col];
                if (LS > 1 && i > 0) {
                    float2 t = tw[k * i * TS];
                    if (DIR > 0) t.y = -t.y;
                    x = cmul(x, t);
                }
                v[p][i] = x;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int b = tid + p * CGS_THREADS;
        if ((NB % CGS_THREADS) == 0 || b < NB) {
            const int col = b % NCOL, j = b / NCOL, k = j % LS;
            Dft<R, DIR>::run(v[p]);
            const int base = (j / LS) * LS * R + k;
#pragma unroll
            for (int i = 0; i < R; ++i) lds[(base + i * LS) * STRIDE + col] = v[p][i];
        }
    }
    __syncthreads();
}

template <int L, int S, int LS, int DIR, int NCOL, int STRIDE = NCOL>
__device__ __forceinline__ void fft_stages(float2* lds, const float2* __restrict__ tw) {
    if constexpr (S < plan_nstages(L)) {
        constexpr int R = plan_radix(L, S);
        fft_stage<L, R, LS, DIR, NCOL, STRIDE>(lds, tw);
        fft_stages<L, S + 1, LS * R, DIR, NCOL, STRIDE>(lds, tw);
    }
}

__device__ __forceinline__ float2 alt_sign(float2 v, int l) {
    return (l & 1) ? make_float2(-v.x, -v.y) : v;
}

// ------------------------------------------------------------ pass kernels
// Strided pass (y or x axis): block = one outer index x one chunk of W
// consecutive columns; lines of length L with stride a.ls.
template <int L, int W, int DIR>
__global__ __launch_bounds__(CGS_THREADS) void k_fft_plain(PlainArgs a) {
    __shared__ float2 lds[L * W];
    const long long nchunk = (a.ncols + W - 1) / W;
    const long long bid = blockIdx.x;
    const long long chunk = bid % nchunk, outer = bid / nchunk;
    if (a.done && a.done[outer / a.outer_per_bin]) return;
    float2* g = a.g + outer * a.os + chunk * W;
    const long long ncol_here = a.ncols - chunk * W;
    for (int e = threadIdx.x; e < L * W; e += CGS_THREADS) {
        const int c = e % W, l = e / W;
        float2 v = make_float2(0.0f, 0.0f);
        if (c < ncol_here) v = g[(long long)l * a.ls + c];
        lds[e] = alt_sign(v, l);
    }
    __syncthreads();
    fft_stages<L, 0, 1, DIR, W>(lds, a.tw);
    for (int e = threadIdx.x; e < L * W; e += CGS_THREADS) {
        const int c = e % W, l = e / W;
        if (c < ncol_here) g[(long long)l * a.ls + c] = alt_sign(lds[e], l);
    }
}

// z pass with a single column (one complex grid, C = 1): block = OB lines.
template <int L, int OB, int DIR>
__global__ __launch_bounds__(CGS_THREADS) void k_fft_z1(float2* g, long long nlines, const float2* __restrict__ tw) {
    __shared__ float2 lds[L * OB];
    const long long line0 = (long long)blockIdx.x * OB;
    for (int e = threadIdx.x; e < L * OB; e += CGS_THREADS) {
        const int l = e % L, ob = e / L;
        float2 v = make_float2(0.0f, 0.0f);
        if (line0 + ob < nlines) v = g[(line0 + ob) * L + l];
        lds[l * OB + ob] = alt_sign(v, l);
    }
    __syncthreads();
    fft_stages<L, 0, 1, DIR, OB>(lds, tw);
    for (int e = threadIdx.x; e < L * OB; e += CGS_THREADS) {
        const int l = e % L, ob = e / L;
        if (line0 + ob < nlines) g[(line0 + ob) * L + l] = alt_sign(lds[l * OB + ob], l);
    }
}

// Forward z pass of E with the fused prologue.  Block = one (bin, x, y) line
// group; the block walks the coils in chunks of W.
//   a.p != null : value = p[bin][cell] * S[cell][c]      (elementProduct, Reco.cpp:85-93)
//   a.p == null : value = G[bin][cell][c]                 (coil image given)
//   a.kft != null: value *= kft[cell]                     (deapodization, Reco.cpp:529-535)
template <int L, int W>
__global__ __launch_bounds__(CGS_THREADS) void k_fft_zpre(ZArgs a) {
    __shared__ float2 lds[L * W];
    const long long outer = blockIdx.x;             // bin * nxy + xy
    const long long bin = outer / a.nxy, xy = outer % a.nxy;
    if (a.done && a.done[bin]) return;
    const long long cell0 = xy * L;                  // cell index of (x, y, z=0)
    float2* g = a.g + (bin * a.ng + cell0) * a.C;
    for (int c0 = 0; c0 < a.C; c0 += W) {
        for (int e = threadIdx.x; e < L * W; e += CGS_THREADS) {
            const int c = e % W, l = e / W;
            const int coil = c0 + c;
            float2 v = make_float2(0.0f, 0.0f);
            if (coil < a.C) {
                const long long cell = cell0 + l;
                if (a.p) v = cmul(a.p[bin * a.ng + cell], a.maps[cell * a.C + coil]);
                else v = g[(long long)l * a.C + coil];
                if (a.kft) v = cscale(v, a.kft[cell]);
            }
            lds[e] = alt_sign(v, l);
        }
        __syncthreads();
        fft_stages<L, 0, 1, -1, W>(lds, a.tw);
        for (int e = threadIdx.x; e < L * W; e += CGS_THREADS) {
            const int c = e % W, l = e / W;
            const int coil = c0 + c;
            if (coil < a.C) g[(long long)l * a.C + coil] = alt_sign(lds[e], l);
        }
        __syncthreads();
    }
}

// Inverse z pass of E^H with the fused epilogue.  Block = one (bin, x, y).
// After the inverse transform each coil value is scaled by kft[cell]*a.scale
// (IFFT 1/sqrt(Ng) and deapodization *1e-3, Reco.cpp:506-515) and cropped to
// the reduceFOV box (Reco.cpp:626-671, inclusive upper bound).
//   a.maps != null : q[bin][cell] = sum_c conj(S[cell][c]) * value
//                    (mutliplyByConjSenseMap + addToSumOfImages, Reco.cpp:107-129):
//                    the products replace the tile in LDS (rows padded by one
//                    element) and one thread per z sums its row over the coils;
//                    the block's partial <p, q> goes to a.part[blockIdx.x]
//   a.maps == null : G[bin][cell][c] = value (coil images, nufftBWD)
template <int L, int W>
__global__ __launch_bounds__(CGS_THREADS) void k_fft_zpost(ZArgs a) {
    constexpr int LDW = W + 1;
    constexpr int RPT = (L + CGS_THREADS - 1) / CGS_THREADS;
    __shared__ float2 lds[L * LDW];
    __shared__ double red[2 * (CGS_THREADS / 64)];
    const long long outer = blockIdx.x;
    const long long bin = outer / a.nxy, xy = outer % a.nxy;
    if (a.done && a.done[bin]) return;
    const long long cell0 = xy * L;
    const int gx = (int)(xy / a.ny), gy = (int)(xy % a.ny);
    const bool xy_in = !a.crop || (gx >= a.lo[0] && gx <= a.hi[0] && gy >= a.lo[1] && gy <= a.hi[1]);
    float2* g = a.g + (bin * a.ng + cell0) * a.C;
    const bool combine = a.maps != nullptr;
    float2 acc[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) acc[r] = make_float2(0.0f, 0.0f);
    for (int c0 = 0; c0 < a.C; c0 += W) {
        for (int e = threadIdx.x; e < L * W; e += CGS_THREADS) {
            const int c = e % W, l = e / W;
            const int coil = c0 + c;
            float2 v = make_float2(0.0f, 0.0f);
            if (coil < a.C) v = g[(long long)l * a.C + coil];
            lds[l * LDW + c] = alt_sign(v, l);
        }
        __syncthreads();
        fft_stages<L, 0, 1, +1, W, LDW>(lds, a.tw);
        for (int e = threadIdx.x; e < L * W; e += CGS_THREADS) {
            const int c = e % W, l = e / W;
            const int coil = c0 + c;
            const long long cell = cell0 + l;
            float2 v = make_float2(0.0f, 0.0f);
            if (coil < a.C) {
                v = alt_sign(lds[l * LDW + c], l);
                const float sc = a.kft ? a.kft[cell] * a.scale : a.scale;
                v = cscale(v, sc);
                if (!xy_in || (a.crop && (l < a.lo[2] || l > a.hi[2]))) v = make_float2(0.0f, 0.0f);
                if (combine) v = cmulc(a.maps[cell * a.C + coil], v);
                else g[(long long)l * a.C + coil] = v;
            }
            if (combine) lds[l * LDW + c] = v;
        }
        if (combine) {
            __syncthreads();
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const int l = threadIdx.x + r * CGS_THREADS;
                if (l < L) {
                    float2 sum = make_float2(0.0f, 0.0f);
#pragma unroll
                    for (int c = 0; c < W; ++c) sum = cadd(sum, lds[l * LDW + c]);
                    acc[r] = cadd(acc[r], sum);
                }
            }
        }
        __syncthreads();
    }
    if (!combine) return;
    double pr = 0.0, pi = 0.0;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int l = threadIdx.x + r * CGS_THREADS;
        if (l < L) {
            const long long idx = bin * a.ng + cell0 + l;
            a.q[idx] = acc[r];
            if (a.p) {
                const float2 pv = a.p[idx];
                pr += (double)pv.x * (double)acc[r].x + (double)pv.y * (double)acc[r].y;
                pi += (double)pv.x * (double)acc[r].y - (double)pv.y * (double)acc[r].x;
            }
        }
    }
    if (a.part) {
        for (int off = 32; off >= 1; off >>= 1) {
            pr += __shfl_xor(pr, off, 64);
            pi += __shfl_xor(pi, off, 64);
        }
        const int wid = threadIdx.x / 64, lane = threadIdx.x % 64;
        if (lane == 0) { red[2 * wid] = pr; red[2 * wid + 1] = pi; }
        __syncthreads();
        if (threadIdx.x == 0) {
            double sr = 0.0, si = 0.0;
            for (int w = 0; w < CGS_THREADS / 64; ++w) { sr += red[2 * w]; si += red[2 * w + 1]; }
            a.part[2 * blockIdx.x] = sr;
            a.part[2 * blockIdx.x + 1] = si;
        }
    }
}

// ------------------------------------------------------------- dispatch
#define CGS_FFT_LENGTHS(X) \
    X(8) X(12) X(16) X(20) X(24) X(28) X(32) X(40) X(48) X(56) X(64) X(80) X(96) X(112) X(128) \
    X(160) X(192) X(224) X(256) X(320) X(384) X(448) X(512)

bool fft_len_supported(int L) {
    switch (L) {
#define CGS_CASE(n) case n: return true;
        CGS_FFT_LENGTHS(CGS_CASE)
#undef CGS_CASE
        default: return false;
    }
}

// Column chunk of the strided passes: keep the LDS tile at or below 32K complex.
static inline int plain_w(int L) { return L <= 256 ? 32 : 16; }

// Coil chunk of the z pass: the whole coil set when the LDS tile stays at or
// below 8K complex (64 KB), else the largest power of two that fits.
int zpass_w(int L, int C) {
    int w = C <= 8 ? 8 : (C <= 16 ? 16 : 32);
    while (w > 8 && L * (w + 1) > 8192) w /= 2;
    return w;
}

template <int L, int DIR>
static void plain_launch(const PlainArgs& a, dim3 grid, hipStream_t s) {
    if constexpr (L <= 256) k_fft_plain<L, 32, DIR><<<grid, CGS_THREADS, 0, s>>>(a);
    else k_fft_plain<L, 16, DIR><<<grid, CGS_THREADS, 0, s>>>(a);
}

hipError_t launch_fft_plain(int L, int dir, const PlainArgs& a, long long nouter, hipStream_t s) {
    const int W = plain_w(L);
    const long long nchunk = (a.ncols + W - 1) / W;
    const dim3 grid((unsigned)(nouter * nchunk));
    switch (L) {
#define CGS_CASE(n)                                            \
    case n:                                                    \
        if (dir < 0) plain_launch<n, -1>(a, grid, s);          \
        else plain_launch<n, +1>(a, grid, s);                  \
        break;
        CGS_FFT_LENGTHS(CGS_CASE)
#undef CGS_CASE
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fft_z1(int L, int dir, float2* g, long long nlines, const float2* tw, hipStream_t s) {
    constexpr int OB = 16;
    const dim3 grid((unsigned)((nlines + OB - 1) / OB));
    switch (L) {
#define CGS_CASE(n)                                                                   \
    case n:                                                                           \
        if (dir < 0) k_fft_z1<n, OB, -1><<<grid, CGS_THREADS, 0, s>>>(g, nlines, tw); \
        else k_fft_z1<n, OB, +1><<<grid, CGS_THREADS, 0, s>>>(g, nlines, tw);         \
        break;
        CGS_FFT_LENGTHS(CGS_CASE)
#undef CGS_CASE
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int L>
static hipError_t zpre_w(int W, const ZArgs& a, long long nouter, hipStream_t s) {
    const dim3 grid((unsigned)nouter);
    if (W == 8) {
        k_fft_zpre<L, 8><<<grid, CGS_THREADS, 0, s>>>(a);
    } else if (W == 16 && L * 17 <= 8192) {
        if constexpr (L * 17 <= 8192) k_fft_zpre<L, 16><<<grid, CGS_THREADS, 0, s>>>(a);
    } else if (W == 32 && L * 33 <= 8192) {
        if constexpr (L * 33 <= 8192) k_fft_zpre<L, 32><<<grid, CGS_THREADS, 0, s>>>(a);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int L>
static hipError_t zpost_w(int W, const ZArgs& a, long long nouter, hipStream_t s) {
    const dim3 grid((unsigned)nouter);
    if (W == 8) {
        k_fft_zpost<L, 8><<<grid, CGS_THREADS, 0, s>>>(a);
    } else if (W == 16 && L * 17 <= 8192) {
        if constexpr (L * 17 <= 8192) k_fft_zpost<L, 16><<<grid, CGS_THREADS, 0, s>>>(a);
    } else if (W == 32 && L * 33 <= 8192) {
        if constexpr (L * 33 <= 8192) k_fft_zpost<L, 32><<<grid, CGS_THREADS, 0, s>>>(a);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fft_zpre(int L, int W, const ZArgs& a, long long nouter, hipStream_t s) {
    switch (L) {
#define CGS_CASE(n) case n: return zpre_w<n>(W, a, nouter, s);
        CGS_FFT_LENGTHS(CGS_CASE)
#undef CGS_CASE
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fft_zpost(int L, int W, const ZArgs& a, long long nouter, hipStream_t s) {
    switch (L) {
#define CGS_CASE(n) case n: return zpost_w<n>(W, a, nouter, s);
        CGS_FFT_LENGTHS(CGS_CASE)
#undef CGS_CASE
        default: return hipErrorInvalidValue;
    }
}

}  // namespace cgs
