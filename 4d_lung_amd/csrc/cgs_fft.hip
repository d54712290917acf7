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
#include <type_traits>

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
template <int L, int R, int LS, int DIR, int NCOL, int STRIDE, int NT = CGS_THREADS>
__device__ __forceinline__ void fft_stage(float2* lds, const float2* __restrict__ tw) {
    constexpr int NB = NCOL * (L / R);
    constexpr int PER = (NB + NT - 1) / NT;
    constexpr int TS = L / (LS * R);
    const int tid = threadIdx.x;
    float2 v[PER][R];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int b = tid + p * NT;
        if ((NB % NT) == 0 || b < NB) {
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
        const int b = tid + p * NT;
        if ((NB % NT) == 0 || b < NB) {
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

// Stages S .. SEND-1 through LDS (each ends with a barrier).
template <int L, int S, int SEND, int LS, int DIR, int NCOL, int STRIDE, int NT>
__device__ __forceinline__ void fft_mid_stages(float2* lds, const float2* tw) {
    if constexpr (S < SEND) {
        constexpr int R = plan_radix(L, S);
        fft_stage<L, R, LS, DIR, NCOL, STRIDE, NT>(lds, tw);
        fft_mid_stages<L, S + 1, SEND, LS * R, DIR, NCOL, STRIDE, NT>(lds, tw);
    }
}

// Product of the radices of stages 0 .. S-1 (sub-transform size before stage S).
__host__ __device__ constexpr int plan_ls(int L, int S) {
    int ls = 1;
    for (int i = 0; i < S; ++i) ls *= plan_radix(L, i);
    return ls;
}

// Last stage (LS * R = L) of a tile held in LDS: butterfly j reads lines
// j + i * LS, and its outputs -- lines j + i * LS again -- go to st(l, col, v)
// instead of back to LDS.  With LDS_STORE a barrier separates the LDS reads
// from the stores, so that st may write the tile.  Also serves L with a single
// stage (LS = 1).
template <int L, int DIR, int NCOL, int STRIDE, bool LDS_STORE, int NT, class Store>
__device__ __forceinline__ void fft_last_stage(const float2* lds, const float2* tw, Store st) {
    constexpr int NS = plan_nstages(L);
    constexpr int R = plan_radix(L, NS - 1);
    constexpr int LS = plan_ls(L, NS - 1);
    constexpr int NB = NCOL * (L / R);
    constexpr int PER = (NB + NT - 1) / NT;
    const int tid = threadIdx.x;
    float2 v[PER][R];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int b = tid + p * NT;
        if ((NB % NT) == 0 || b < NB) {
            const int col = b % NCOL, j = b / NCOL;
#pragma unroll
            for (int i = 0; i < R; ++i) {
                float2 x = lds[(j + i * LS) * STRIDE + col];
                if (LS > 1 && i > 0) {
                    float2 t = tw[j * i];
                    if (DIR > 0) t.y = -t.y;
                    x = cmul(x, t);
                }
                v[p][i] = x;
            }
        }
    }
    if constexpr (LDS_STORE) __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int b = tid + p * NT;
        if ((NB % NT) == 0 || b < NB) {
            const int col = b % NCOL, j = b / NCOL;
            Dft<R, DIR>::run(v[p]);
#pragma unroll
            for (int i = 0; i < R; ++i) st(j + i * LS, col, v[p][i], p, i);
        }
    }
}

// Output slot (p, i) of the last stage: line and column, as fft_last_stage
// hands them to its store functor (for epilogue operands fetched early).
template <int L, int NCOL, int NT>
struct LastStageMap {
    static constexpr int NS = plan_nstages(L);
    static constexpr int R = plan_radix(L, NS - 1);
    static constexpr int LS = plan_ls(L, NS - 1);
    static constexpr int NB = NCOL * (L / R);
    static constexpr int PER = (NB + NT - 1) / NT;
    __device__ static __forceinline__ bool valid(int p) { return (NB % NT) == 0 || (int)threadIdx.x + p * NT < NB; }
    __device__ static __forceinline__ int col(int p) { return ((int)threadIdx.x + p * NT) % NCOL; }
    __device__ static __forceinline__ int line(int p, int i) { return ((int)threadIdx.x + p * NT) / NCOL + i * LS; }
};

// Last stage into registers: out[p][i] is the output of slot (p, i) (line
// and column from LastStageMap), left to the caller to store.
template <int L, int DIR, int NCOL, int STRIDE, int NT, int PER, int R>
__device__ __forceinline__ void fft_last_stage_regs(const float2* lds, const float2* tw, float2 (&out)[PER][R]) {
    constexpr int NS = plan_nstages(L);
    constexpr int LS = plan_ls(L, NS - 1);
    constexpr int NB = NCOL * (L / R);
    static_assert(R == plan_radix(L, NS - 1) && PER == (NB + NT - 1) / NT, "register tile shape");
    const int tid = threadIdx.x;
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int b = tid + p * NT;
        if ((NB % NT) == 0 || b < NB) {
            const int col = b % NCOL, j = b / NCOL;
#pragma unroll
            for (int i = 0; i < R; ++i) {
                float2 x = lds[(j + i * LS) * STRIDE + col];
                if (LS > 1 && i > 0) {
                    float2 t = tw[j * i];
                    if (DIR > 0) t.y = -t.y;
                    x = cmul(x, t);
                }
                out[p][i] = x;
            }
            Dft<R, DIR>::run(out[p]);
        }
    }
}

// Line transform of an [L x NCOL] tile already in LDS (and visible to the
// block): every stage but the last through LDS, the last into st.
template <int L, int DIR, int NCOL, int STRIDE, bool LDS_STORE, int NT, class Store>
__device__ __forceinline__ void fft_lds_tile(float2* lds, const float2* tw, Store st) {
    constexpr int NS = plan_nstages(L);
    fft_mid_stages<L, 0, NS - 1, 1, DIR, NCOL, STRIDE, NT>(lds, tw);
    fft_last_stage<L, DIR, NCOL, STRIDE, LDS_STORE, NT>(lds, tw, st);
}

// Whole line transform of an [L x NCOL] tile whose first stage takes its
// inputs straight from ld(l, col) (HBM -> registers) and whose last stage
// hands its outputs straight to st(l, col, v) (registers -> HBM or an
// epilogue); only the stages in between go through LDS.  tw: L twiddles
// exp(-2 pi i m / L), visible to the block after the first barrier.  Ends
// without a barrier.
template <int L, int DIR, int NCOL, int STRIDE, bool LDS_STORE, int NT, class Load, class Store>
__device__ __forceinline__ void fft_line_tile(float2* lds, const float2* tw, Load ld, Store st) {
    constexpr int NS = plan_nstages(L);
    const int tid = threadIdx.x;
    if constexpr (NS == 1) {
        constexpr int R = L;
        constexpr int PER = (NCOL + NT - 1) / NT;
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int col = tid + p * NT;
            if (col < NCOL) {
                float2 v[R];
#pragma unroll
                for (int i = 0; i < R; ++i) v[i] = ld(i, col);
                Dft<R, DIR>::run(v);
#pragma unroll
                for (int i = 0; i < R; ++i) st(i, col, v[i], p, i);
            }
        }
    } else {
        {   // first stage: LS = 1, no twiddles
            constexpr int R = plan_radix(L, 0);
            constexpr int NB = NCOL * (L / R);
            constexpr int PER = (NB + NT - 1) / NT;
            float2 v[PER][R];
#pragma unroll
            for (int p = 0; p < PER; ++p) {
                const int b = tid + p * NT;
                if ((NB % NT) == 0 || b < NB) {
                    const int col = b % NCOL, j = b / NCOL;
#pragma unroll
                    for (int i = 0; i < R; ++i) v[p][i] = ld(j + i * (L / R), col);
                }
            }
#pragma unroll
            for (int p = 0; p < PER; ++p) {
                const int b = tid + p * NT;
                if ((NB % NT) == 0 || b < NB) {
                    const int col = b % NCOL, j = b / NCOL;
                    Dft<R, DIR>::run(v[p]);
#pragma unroll
                    for (int i = 0; i < R; ++i) lds[(j * R + i) * STRIDE + col] = v[p][i];
                }
            }
            __syncthreads();
        }
        fft_mid_stages<L, 1, NS - 1, plan_radix(L, 0), DIR, NCOL, STRIDE, NT>(lds, tw);
        fft_last_stage<L, DIR, NCOL, STRIDE, LDS_STORE, NT>(lds, tw, st);
    }
}

// A thread's share of an [L x NCOL] tile in registers, element e = tid + k*NT
// at line e / NCOL, column e % NCOL (consecutive threads on consecutive
// columns, so each line is one coalesced run).  Holds the raw loads (type T)
// of the next tile while the current one is transformed; cv(l, col, raw)
// turns them into tile values when they go to LDS, after the loads landed.
template <int L, int NCOL, int NT, class T>
struct TileRegs {
    static constexpr int EPT = (L * NCOL + NT - 1) / NT;
    T v[EPT];
    template <class Load>
    __device__ __forceinline__ void load(Load ld) {
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int e = threadIdx.x + k * NT;
            if ((L * NCOL) % NT == 0 || e < L * NCOL) v[k] = ld(e / NCOL, e % NCOL);
        }
    }
    template <int STRIDE, class Conv>
    __device__ __forceinline__ void to_lds(float2* lds, Conv cv) const {
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int e = threadIdx.x + k * NT;
            if ((L * NCOL) % NT == 0 || e < L * NCOL) {
                const int l = e / NCOL, c = e % NCOL;
                lds[l * STRIDE + c] = cv(l, c, v[k]);
            }
        }
    }
};

// The first stage's inputs of a thread (butterfly b = tid + p*NT reads lines
// j + i*L/R0 of column b % NCOL, j = b / NCOL) held in registers, so that the
// next tile can be fetched while the current one is transformed; run() does
// the radix-R0 butterflies and writes the stage output to LDS.
template <int L, int NCOL, int NT>
struct FirstRegs {
    static constexpr int R = plan_radix(L, 0);
    static constexpr int NB = NCOL * (L / R);
    static constexpr int PER = (NB + NT - 1) / NT;
    float2 v[PER][R];
    template <class Load>
    __device__ __forceinline__ void load(Load ld) {
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int b = threadIdx.x + p * NT;
            if ((NB % NT) == 0 || b < NB) {
                const int col = b % NCOL, j = b / NCOL;
#pragma unroll
                for (int i = 0; i < R; ++i) v[p][i] = ld(i, j, col);   // line j + i * L / R
            }
        }
    }
    template <int DIR, int STRIDE, class Conv>
    __device__ __forceinline__ void run(float2* lds, Conv cv) {
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int b = threadIdx.x + p * NT;
            if ((NB % NT) == 0 || b < NB) {
                const int col = b % NCOL, j = b / NCOL;
#pragma unroll
                for (int i = 0; i < R; ++i) v[p][i] = cv(j + i * (L / R), col, v[p][i]);
                Dft<R, DIR>::run(v[p]);
#pragma unroll
                for (int i = 0; i < R; ++i) lds[(j * R + i) * STRIDE + col] = v[p][i];
            }
        }
    }
};

// Raw inputs of one forward z-pass element: u * w * k (w = 1, k = 1 where
// the product has fewer factors).
struct ZRaw {
    float2 u, w;
    float k;
};

__device__ __forceinline__ float2 alt_sign(float2 v, int l) {
    return (l & 1) ? make_float2(-v.x, -v.y) : v;
}

// ------------------------------------------------------------ pass kernels

// Strided pass (y or x axis): block = one outer index x CPB consecutive
// chunks of W columns; lines of length L with stride a.ls.  The first-stage
// inputs of the next chunk are fetched while the current chunk is transformed.
// Passes whose lines stride a megabyte or more (x) take chunks of 16 columns,
// 4 per block; the others (y) 32 columns up to 256-point lines (16 beyond),
// 6 per block (measured best on the 256x256x192 grids of 16 coils; neutral
// on the 320x320x224 ones).
#ifndef CGS_PLAIN_CPB_FAR
#define CGS_PLAIN_CPB_FAR 4
#endif
#ifndef CGS_PLAIN_CPB_NEAR
#define CGS_PLAIN_CPB_NEAR 6
#endif
constexpr int plain_cpb(bool far) { return far ? CGS_PLAIN_CPB_FAR : CGS_PLAIN_CPB_NEAR; }
// Element at a 32-bit byte offset (or a 64-bit element offset) from a
// block-uniform base: the 32-bit form lets the load use a scalar base address
// plus one offset register.
template <class Off>
__device__ __forceinline__ float2* elem(float2* base, Off off) {
    if constexpr (sizeof(Off) == 4) return (float2*)((char*)base + (uint32_t)(off * (Off)sizeof(float2)));
    else return base + off;
}

// Addressing of the strided passes: line l = j + i * S of a column chunk
// (i, S: radix index and stride of the first or last stage) at
//   kOff32: one 32-bit byte offset from the chunk (column span below 4 GB);
//   kOffSeg: the uniform base of line group i plus a 32-bit offset of
//            j * ls + c (a group of S lines spans less than 4 GB, the whole
//            column does not: x passes of the largest grids);
//   kOff64: 64-bit element offsets.
enum { kOff32 = 0, kOffSeg = 1, kOff64 = 2 };

template <int AM>
__device__ __forceinline__ float2* line_at(float2* g, long long ls, int i, int S, int j, int c) {
    if constexpr (AM == kOff32) return elem<uint32_t>(g, (uint32_t)(j + i * S) * (uint32_t)ls + (uint32_t)c);
    else if constexpr (AM == kOffSeg) return elem<uint32_t>(g + (long long)(i * S) * ls, (uint32_t)j * (uint32_t)ls + (uint32_t)c);
    else return g + (long long)(j + i * S) * ls + c;
}

// Zeros in columns [col_lo, col_lo + ncol) of the L lines (stride ls) of a
// column group whose chunks no sample reaches.  The group's NCOL columns are
// written row by row with 16-byte stores where the rows allow (even stride
// and first column, full group, 16-byte aligned base), else element by
// element.
template <int L, int NCOL>
__device__ __forceinline__ void zero_group(float2* g, long long ls, long long col_lo, int ncol) {
    float2* const b = g + col_lo;
    if (NCOL % 2 == 0 && ncol == NCOL && ((ls | col_lo) & 1) == 0 && (reinterpret_cast<uintptr_t>(g) & 15) == 0) {
        constexpr int N2 = NCOL / 2 > 0 ? NCOL / 2 : 1;
        for (int e = threadIdx.x; e < L * N2; e += CGS_THREADS) {
            const int l = e / N2, c2 = e - (e / N2) * N2;
            *reinterpret_cast<float4*>(b + (long long)l * ls + 2 * c2) = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    } else {
        for (int e = threadIdx.x; e < L * ncol; e += CGS_THREADS) {
            const int l = e / ncol, c = e - (e / ncol) * ncol;
            b[(long long)l * ls + c] = make_float2(0.0f, 0.0f);
        }
    }
}

template <int L, int W, int DIR, int CPB, int AM>
__global__ __launch_bounds__(CGS_THREADS) void k_fft_plain(PlainArgs a) {
    __shared__ float2 lds[L * W];
    __shared__ float2 twl[L];
    const long long nchunk = (a.ncols + W - 1) / W;
    const long long ngrp = (nchunk + CPB - 1) / CPB;
    const long long bid = blockIdx.x;
    long long grp, outer;
    if (a.glist) {   // only the listed column groups
        outer = bid / a.gstride;
        const long long k = bid % a.gstride;
        if (k >= a.gcnt[outer]) return;
        grp = a.glist[outer * a.gstride + k];
    } else {
        grp = bid % ngrp;
        outer = bid / ngrp;
    }
    if (a.done && a.done[outer / a.outer_per_bin]) return;
    float2* gout = a.g + outer * a.os;
    const long long ls = a.ls;
    const long long ncols = a.ncols;
    constexpr int NS = plan_nstages(L);
    const long long c_first = grp * CPB;
    const long long c_end = min(c_first + CPB, nchunk);
    if (a.gact && !a.gact[outer * ngrp + grp]) {   // no sample reaches these columns: zeros
        zero_group<L, W * CPB>(gout, ls, c_first * W, (int)(min(c_end * W, ncols) - c_first * W));
        return;
    }
    for (int e = threadIdx.x; e < L; e += CGS_THREADS) twl[e] = a.tw[e];
    if constexpr (NS == 1) {   // one radix stage: registers only (line l = group l of one line)
        for (long long chunk = c_first; chunk < c_end; ++chunk) {
            float2* g = gout + chunk * W;
            const long long nh = ncols - chunk * W;
            fft_line_tile<L, DIR, W, W, false, CGS_THREADS>(
                lds, twl,
                [&](int l, int c) {
                    float2 v = make_float2(0.0f, 0.0f);
                    if (W <= nh || c < nh) v = *line_at<AM>(g, ls, l, 1, 0, c);
                    return alt_sign(v, l);
                },
                [&](int l, int c, float2 v, int, int) {
                    if (W <= nh || c < nh) *line_at<AM>(g, ls, l, 1, 0, c) = alt_sign(v, l);
                });
        }
        return;
    }
    constexpr int S0 = L / plan_radix(L, 0);        // first-stage line stride
    constexpr int SL = plan_ls(L, NS - 1);          // last-stage line stride
    // columns past the end of the last chunk: loads clamped to its last
    // column (unconditional), values zeroed in registers, stores skipped
    int cmax = 0;
    auto cv = [&](int l, int c, float2 v) {
        if (c > cmax) v = make_float2(0.0f, 0.0f);
        return alt_sign(v, l);
    };
    auto load_chunk = [&](FirstRegs<L, W, CGS_THREADS>& r, long long chunk) {
        float2* g = gout + chunk * W;
        const int cm = (int)min((long long)W, ncols - chunk * W) - 1;
        r.load([&](int i, int j, int c) { return *line_at<AM>(g, ls, i, S0, j, min(c, cm)); });
    };
    FirstRegs<L, W, CGS_THREADS> next;
    load_chunk(next, c_first);
    for (long long chunk = c_first; chunk < c_end; ++chunk) {
        cmax = (int)min((long long)W, ncols - chunk * W) - 1;
        next.template run<DIR, W>(lds, cv);
        __syncthreads();
        if (chunk + 1 < c_end) load_chunk(next, chunk + 1);
        fft_mid_stages<L, 1, NS - 1, plan_radix(L, 0), DIR, W, W, CGS_THREADS>(lds, twl);
        float2* g = gout + chunk * W;
        fft_last_stage<L, DIR, W, W, false, CGS_THREADS>(lds, twl, [&](int l, int c, float2 v, int, int i) {
            if (c <= cmax) *line_at<AM>(g, ls, i, SL, l - i * SL, c) = alt_sign(v, l);
        });
        if (chunk + 1 < c_end) __syncthreads();
    }
}

// Column groups of the fused gridding pass that one XCD takes in a row
// (neighbouring cells share samples; 32 measured best, 0 keeps the plain
// workgroup order).
#ifndef CGS_XGRID_XG
#define CGS_XGRID_XG 32
#endif

// Inverse x pass with the gridding of E^H fused in (columns = coil groups of
// W = 16 coils of one (y, z) cell, C a multiple of 16).  Instead of reading a
// gridded chunk, the block forms it: the value of line x, coil c of cell
// (y, z) is the sum over the cell's listed samples with bx <= x < bx + KW
// (cgs_xlists.hip, sorted by (bx, sample)) of w * Y[s][c] (times gdcf[s]
// when given), w = float((vx * vy) * vz), added when w >= thrf -- the
// weights and the tap set of gridding (Reco.cpp:492-504; Reco.h:218-242).
// The cell's listed samples are staged in LDS, in the memory of the FFT tile
// before the tile is filled: per sample its bx and its KW weights along x
// for this cell (read from the entry's record, computed when the lists were
// built) and its 16 coil values; a table of the entry run reaching each line
// x (found by binary searches on bx) follows, and each first-stage input of a
// thread sums its run from LDS.  The chunk runs through the same stages and
// stores as the plain inverse x pass.
// Column groups that no sample reaches (a.gact) are written as zeros.
template <int L, int W, int CPB, int AM, int KW>
__global__ __launch_bounds__(CGS_THREADS) void k_fft_xgrid(XGridArgs xa) {
    const PlainArgs& a = xa.x;
    constexpr int RS = xrec_stride(KW);
    constexpr int NS = plan_nstages(L);
    static_assert(NS > 1, "fused gridding needs a multi-stage line transform");
    static_assert(W == 16, "fused gridding stages 16 coil values per sample");
    using FR = FirstRegs<L, W, CGS_THREADS>;
    constexpr int R0 = FR::R;
    constexpr int SL = plan_ls(L, NS - 1);
    // staged samples per batch: W values (float2) and bx (int), needed for
    // the current chunk only; per entry its KW weights in a row of EWS slots
    // (weight ox at slot (bx + ox) mod EWS: line x of the entry's run finds
    // its weight at slot x mod EWS without reading bx), sample index (int),
    // density weight (float); per line x the range of staged entries that
    // reach it.  All in the tile's memory, or -- where a side region of LDS
    // costs no block per CU (kSide) -- the per-entry rows and the line ranges
    // in a region of their own, which outlives the chunk's transform: the
    // next chunk of the same cell (C > W) then stages only its values.
    constexpr int EWS = KW <= 4 ? 4 : 8;
    constexpr int kMeta = 4 * EWS + 8;
    constexpr int kLdsCu = 160 * 1024, kBase = L * W * 8 + L * 8;
    constexpr int kBudget = kLdsCu / (kLdsCu / kBase);   // per block, keeping the blocks per CU
    constexpr int kCap = L < CGS_THREADS ? (L / 16) * 16 : CGS_THREADS;
    constexpr int NES0 = ((kBudget - kBase - 4 * L - 64) / kMeta / 16) * 16;
    constexpr int NES = NES0 < kCap ? NES0 : kCap;
    constexpr bool kSide = NES >= 128;
    constexpr int kSideBytes = kSide ? NES * kMeta + 4 * L : 4;
    constexpr int NEB = (L * W * 8 - 4 * L) / (8 * W + 4 + kMeta);
    constexpr int NE = kSide ? NES : ((NEB / 16) * 16 < kCap ? (NEB / 16) * 16 : kCap);
    static_assert(NE >= 16, "tile too small to stage samples");
    static_assert(NE * (8 * W + 4) + (kSide ? 0 : NE * kMeta + 4 * L) <= L * W * 8, "staging exceeds the tile");
    __shared__ float2 lds[L * W];
    __shared__ float2 twl[L];
    __shared__ int side[kSideBytes / 4];
    float2* const ey = lds;                                        // [NE][W]
    int* const ebx = reinterpret_cast<int*>(lds + NE * W);         // [NE]
    float* const ew = kSide ? reinterpret_cast<float*>(side)       // [NE][EWS]
                            : reinterpret_cast<float*>(ebx + NE);
    int* const esid = reinterpret_cast<int*>(ew + NE * EWS);       // [NE]
    float* const edcf = reinterpret_cast<float*>(esid + NE);       // [NE]
    int* const xr = reinterpret_cast<int*>(edcf + NE);             // [L] first | end << 16
    const long long nchunk = (a.ncols + W - 1) / W;
    const long long ngrp = (nchunk + CPB - 1) / CPB;
#if CGS_XGRID_XG > 0
    // consecutive workgroups go round-robin to the 8 XCDs: runs of XG
    // consecutive column groups (neighbouring cells, which read the same
    // samples) are given to one XCD, so its L2 serves their shared samples
    unsigned lin = blockIdx.x;
    if (gridDim.x % (8u * CGS_XGRID_XG) == 0) {
        const unsigned s8 = blockIdx.x >> 3;
        lin = (s8 / CGS_XGRID_XG) * (8u * CGS_XGRID_XG) + (blockIdx.x & 7u) * CGS_XGRID_XG + s8 % CGS_XGRID_XG;
    }
    const long long grp = lin % ngrp, outer = lin / ngrp;   // outer: bin of the launch
#else
    const long long grp = blockIdx.x % ngrp, outer = blockIdx.x / ngrp;   // outer: bin of the launch
#endif
    if (a.done && a.done[outer]) return;
    float2* gout = a.g + outer * a.os;
    const long long ls = a.ls;
    const long long c_first = grp * CPB;
    const long long c_end = min(c_first + CPB, nchunk);
    if (a.gact && !a.gact[outer * ngrp + grp]) {   // no sample reaches these columns: zeros
        zero_group<L, W * CPB>(gout, ls, c_first * W, (int)(min(c_end * W, a.ncols) - c_first * W));
        return;
    }
    for (int e = threadIdx.x; e < L; e += CGS_THREADS) twl[e] = a.tw[e];
    const int ncell = xa.ny * xa.nz;
    const long long* __restrict__ offb = xa.off + (long long)(xa.b0 + outer) * (ncell + 1);
    const int tid = threadIdx.x;
    // list offsets of the group's cells (W <= C: at most CPB cells), loaded
    // before any store of the block (a later wait for them would also wait
    // for the stores issued before it)
    const int cell_lo = (int)((c_first * W) / xa.C);
    long long offs[CPB + 1];
#pragma unroll
    for (int k = 0; k <= CPB; ++k) {   // uniform: kept in scalar registers
        const long long v = cell_lo + k <= ncell ? offb[cell_lo + k] : 0;
        offs[k] = (long long)(((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(v >> 32)) << 32) |
                              (unsigned)__builtin_amdgcn_readfirstlane((int)v));
    }
    int kept = -1;   // cell whose (single) batch of records and run table the side region holds
    for (long long chunk = c_first; chunk < c_end; ++chunk) {
        const long long col0 = chunk * W;
        const int cell = (int)(col0 / xa.C), c0 = (int)(col0 % xa.C);
        long long e0 = offs[0], e1 = offs[1];
#pragma unroll
        for (int k = 1; k < CPB; ++k)
            if (cell - cell_lo == k) { e0 = offs[k]; e1 = offs[k + 1]; }
        const bool reuse = kSide && cell == kept;
        FR next;
#pragma unroll
        for (int p = 0; p < FR::PER; ++p)
#pragma unroll
            for (int i = 0; i < R0; ++i) next.v[p][i] = make_float2(0.0f, 0.0f);
        for (long long eb = e0; eb < e1; eb += NE) {
            const int ne = (int)min((long long)NE, e1 - eb);
            // stage: the entries' records (one thread per entry), then the
            // coil values, straight from memory into LDS (16 bytes per lane, a
            // wave writes 8 entries' 16 values contiguously: no registers,
            // every load in flight while the run table is built)
            if (!reuse && tid < ne) {
                const int4* r = reinterpret_cast<const int4*>(xa.rec + (eb + tid) * RS);
                int rr[RS];
#pragma unroll
                for (int q = 0; q < RS / 4; ++q) {
                    const int4 v = r[q];
                    rr[4 * q] = v.x; rr[4 * q + 1] = v.y; rr[4 * q + 2] = v.z; rr[4 * q + 3] = v.w;
                }
                ebx[tid] = rr[0];
#pragma unroll
                for (int ox = 0; ox < KW; ++ox) ew[tid * EWS + ((rr[0] + ox) & (EWS - 1))] = __int_as_float(rr[1 + ox]);
                esid[tid] = rr[KW + 1];
                if (xa.gdcf) edcf[tid] = xa.gdcf[rr[KW + 1]];
            }
            __syncthreads();
            {
                static_assert(W == 16, "a wave stages 8 entries of 16 values");
                const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
                for (int k = wave; k < (ne + 7) / 8; k += CGS_THREADS / 64) {
                    const int e = k * 8 + (lane >> 3);
                    if (e < ne) {
                        const float2* src = xa.Y + (long long)esid[e] * xa.C + c0 + 2 * (lane & 7);
                        __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)(ey + k * 128),
                                                         16, 0, 0);
                    }
                }
            }
            // entries reaching line x (x - KW < bx <= x) are the sorted run
            // [first bx >= x - KW + 1, first bx > x): two searches per line,
            // shared by the line's W columns
            for (int x = reuse ? L : tid; x < L; x += CGS_THREADS) {
                int lo = 0, hi = ne;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (ebx[mid] < x - KW + 1) lo = mid + 1; else hi = mid;
                }
                int en = lo;
                hi = ne;
                while (en < hi) {
                    const int mid = (en + hi) >> 1;
                    if (ebx[mid] <= x) en = mid + 1; else hi = mid;
                }
                xr[x] = lo | (en << 16);
            }
            __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): the direct loads have reached LDS
            __syncthreads();
            // each line adds its run in ascending entry order
#pragma unroll
            for (int p = 0; p < FR::PER; ++p) {
                const int b = tid + p * CGS_THREADS;
                if ((FR::NB % CGS_THREADS) == 0 || b < FR::NB) {
                    const int col = b % W, j = b / W;
#pragma unroll
                    for (int i = 0; i < R0; ++i) {
                        const int x = j + i * (L / R0);
                        const int r = xr[x];
                        float2 acc = next.v[p][i];
                        for (int e = r & 0xffff; e < (r >> 16); ++e) {
                            const float w = ew[e * EWS + (x & (EWS - 1))];   // weight x - bx of entry e
                            float2 yv = ey[e * W + col];
                            if (w >= xa.thrf) {
                                if (xa.gdcf) yv = cscale(yv, edcf[e]);   // as gridding weights the values
                                acc.x = fmaf(w, yv.x, acc.x);
                                acc.y = fmaf(w, yv.y, acc.y);
                            }
                        }
                        next.v[p][i] = acc;
                    }
                }
            }
            __syncthreads();   // staged samples read (the memory is the tile's)
        }
        kept = kSide && e1 > e0 && e1 - e0 <= NE ? cell : -1;
        const int cmax = (int)min((long long)W, a.ncols - col0) - 1;
        next.template run<+1, W>(lds, [&](int l, int c, float2 v) {
            if (c > cmax) v = make_float2(0.0f, 0.0f);
            return alt_sign(v, l);
        });
        __syncthreads();
        fft_mid_stages<L, 1, NS - 1, R0, +1, W, W, CGS_THREADS>(lds, twl);
        float2* g = gout + chunk * W;
        fft_last_stage<L, +1, W, W, false, CGS_THREADS>(lds, twl, [&](int l, int c, float2 v, int, int i) {
            if (c <= cmax) *line_at<AM>(g, ls, i, SL, l - i * SL, c) = alt_sign(v, l);
        });
        __syncthreads();   // the tile read before the next chunk's samples are staged in it
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

// Workgroup order of the z passes.  The bins of one launch share the coil
// maps and the deapodization line of every (x, y) cell line, and the device
// hands consecutive workgroups round-robin to its 8 XCDs (each with its own
// L2).  Workgroup k therefore takes cell line xy = (s / nb) * 8 + k % 8 of
// bin s % nb, s = k / 8: the nb workgroups of one cell line have ids 8 apart,
// run on one XCD close together in time, and all but the first find that
// line's maps in L2.  A bijection onto bin * nxy + xy whenever nxy is a
// multiple of 8; otherwise (and for one bin) the bin-major order.  32-bit
// arithmetic (nb * nxy < 2^32): a workgroup computes its indices before its
// first load, and 64-bit divisions there cost a long scalar sequence.
__device__ __forceinline__ unsigned zpass_line(unsigned k, unsigned nxy, unsigned nb) {
    if (nb <= 1 || (nxy & 7)) return k;
    const unsigned s = k >> 3;
    return (s % nb) * nxy + (s / nb) * 8 + (k & 7);
}

// Forward z pass of E with the fused prologue.  Block = one (bin, x, y) line;
// the block walks the coils in chunks of W, fetching the next chunk into
// registers while the current one is transformed.
//   a.p != null : value = p[bin][cell] * S[cell][c]      (elementProduct, Reco.cpp:85-93)
//   a.p == null : value = G[bin][cell][c]                 (coil image given)
//   a.kft != null: value *= kft[cell]                     (deapodization, Reco.cpp:529-535)
template <int L, int W, int NT, bool HAS_P, bool HAS_K>
__global__ __launch_bounds__(NT) void k_fft_zpre(ZArgs a) {
    __shared__ float2 lds[L * W];
    __shared__ float2 twl[L];
    __shared__ float2 pline[HAS_P ? L : 1];   // p along the line (shared by all coils)
    __shared__ float kline[HAS_K ? L : 1];    // kernelFT along the line
    const unsigned nxy = (unsigned)a.nxy, ny = (unsigned)a.ny;
    const unsigned outer = zpass_line(blockIdx.x, nxy, (unsigned)a.nb);   // bin * nxy + xy
    const unsigned bin = outer / nxy, xy = outer - bin * nxy;
    const unsigned gx = xy / ny, gy = xy - gx * ny;
    const long long cell0 = (long long)xy * L;       // cell index of (x, y, z=0)
    for (int e = threadIdx.x; e < L; e += NT) {
        twl[e] = a.tw[e];
        if constexpr (HAS_P) pline[e] = a.p[bin * a.ng + cell0 + e];
        if constexpr (HAS_K) kline[e] = a.kft[cell0 + e];
    }
    // block-uniform bases and 32-bit offsets within the line; coil indices
    // are clamped so that every load is unconditional (columns past the last
    // coil are zeroed when the tile is written to LDS)
    float2* g = a.g + bin * a.bp + gx * a.xp + gy * (long long)(L * a.C);
    const int C = a.C;
    const float2* ml = HAS_P ? a.maps + cell0 * C : nullptr;
    // raw element: S (with p) or the coil image value (without)
    auto ld = [&](int c0) {
        return [&, c0](int l, int c) {
            const int coil = min(c0 + c, C - 1);
            if constexpr (HAS_P) return ml[l * C + coil];
            else return g[l * C + coil];
        };
    };
    int cur = 0;
    auto cv = [&](int l, int c, float2 r) {
        float2 v = r;
        if constexpr (HAS_P) v = cmul(pline[l], r);
        if constexpr (HAS_K) v = cscale(v, kline[l]);
        if (cur + c >= C) v = make_float2(0.0f, 0.0f);
        return alt_sign(v, l);
    };
    TileRegs<L, W, NT, float2> next;
    next.load(ld(0));
    // a finished bin's workgroup leaves after issuing its loads (reading a
    // finished bin is harmless; checking first would put the flag's latency
    // in front of every load)
    if (a.done && a.done[bin]) return;
    __syncthreads();   // pline / kline / twl visible
    // The outputs of an even chunk are held in registers and stored together
    // with the next chunk's, so that both halves of a cell's 16-coil line are
    // written back to back.
    using LM = LastStageMap<L, W, NT>;
    constexpr int NS = plan_nstages(L);
    float2 hold[LM::PER][LM::R];
    for (int c0 = 0; c0 < C; c0 += W) {
        cur = c0;
        next.template to_lds<W>(lds, cv);
        __syncthreads();
        if (c0 + W < C) next.load(ld(c0 + W));
        fft_mid_stages<L, 0, NS - 1, 1, -1, W, W, NT>(lds, twl);
        float2 out[LM::PER][LM::R];
        fft_last_stage_regs<L, -1, W, W, NT>(lds, twl, out);
        const bool keep = ((c0 / W) % 2 == 0) && (c0 + W < C);
        const bool pair = (c0 / W) % 2 == 1;
#pragma unroll
        for (int p = 0; p < LM::PER; ++p) {
            if (!LM::valid(p)) continue;
            const int c = LM::col(p);
#pragma unroll
            for (int i = 0; i < LM::R; ++i) {
                const int l = LM::line(p, i);
                const float2 o = alt_sign(out[p][i], l);
                if (keep) {
                    hold[p][i] = o;
                } else {
                    if (pair) g[l * C + c0 - W + c] = hold[p][i];
                    if (c0 + c < C) g[l * C + c0 + c] = o;
                }
            }
        }
        if (c0 + W < C) __syncthreads();
    }
}

// Inverse z pass of E^H with the fused epilogue.  Block = one (bin, x, y);
// coil chunks of W, the next one fetched while the current one is transformed.
// After the inverse transform each coil value is scaled by kft[cell]*a.scale
// (IFFT 1/sqrt(Ng) and deapodization *1e-3, Reco.cpp:506-515) and cropped to
// the reduceFOV box (Reco.cpp:626-671, inclusive upper bound).
//   a.maps != null : q[bin][cell] = sum_c conj(S[cell][c]) * value
//                    (mutliplyByConjSenseMap + addToSumOfImages, Reco.cpp:107-129):
//                    the products replace the tile in LDS (rows padded by one
//                    element) and one thread per z sums its row over the coils;
//                    the block's partial <p, q> goes to a.part[blockIdx.x]
//   a.maps == null : G[bin][cell][c] = value (coil images, nufftBWD)
template <int L, int W, int NT>
__global__ __launch_bounds__(NT) void k_fft_zpost(ZArgs a) {
    constexpr int LDW = W + 1;
    constexpr int RPT = (L + NT - 1) / NT;
    __shared__ float2 lds[L * LDW];
    __shared__ float2 twl[L];
    __shared__ float kline[L];      // kernelFT * scale along the line
    __shared__ double red[2 * (NT / 64)];
    const unsigned nxy = (unsigned)a.nxy, ny = (unsigned)a.ny;
    const unsigned outer = zpass_line(blockIdx.x, nxy, (unsigned)a.nb);   // bin * nxy + xy
    const unsigned bin = outer / nxy, xy = outer - bin * nxy;
    const long long cell0 = (long long)xy * L;
    for (int e = threadIdx.x; e < L; e += NT) {
        twl[e] = a.tw[e];
        kline[e] = a.kft ? a.kft[cell0 + e] * a.scale : a.scale;
    }
    const int gx = (int)(xy / ny), gy = (int)(xy - (unsigned)gx * ny);
    const bool xy_in = !a.crop || (gx >= a.lo[0] && gx <= a.hi[0] && gy >= a.lo[1] && gy <= a.hi[1]);
    float2* g = a.g + bin * a.bp + gx * a.xp + gy * (long long)(L * a.C);
    const bool combine = a.maps != nullptr;
    const int C = a.C;
    const float2* ml = a.maps ? a.maps + cell0 * C : nullptr;
    float2 acc[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) acc[r] = make_float2(0.0f, 0.0f);
    auto ld = [&](int c0) {
        return [&, c0](int l, int c) { return g[l * C + min(c0 + c, C - 1)]; };
    };
    int cur = 0;
    auto cv = [&](int l, int c, float2 v) {
        if (cur + c >= C) v = make_float2(0.0f, 0.0f);
        return alt_sign(v, l);
    };
    TileRegs<L, W, NT, float2> next;
    next.load(ld(0));
    if (a.done && a.done[bin]) return;   // after the loads are issued (see k_fft_zpre)
    for (int c0 = 0; c0 < C; c0 += W) {
        cur = c0;
        next.template to_lds<LDW>(lds, cv);
        __syncthreads();
        if (c0 + W < C) next.load(ld(c0 + W));
        // the epilogue's S_c of every output of this thread's last-stage
        // butterflies is fetched now, while the stages run
        using LM = LastStageMap<L, W, NT>;
        float2 spre[LM::PER][LM::R];
#pragma unroll
        for (int p = 0; p < LM::PER; ++p) {
            if (!LM::valid(p)) continue;
            const int coil = min(c0 + LM::col(p), C - 1);
#pragma unroll
            for (int i = 0; i < LM::R; ++i)
                spre[p][i] = ml ? ml[LM::line(p, i) * C + coil] : make_float2(0.0f, 0.0f);
        }
        // inverse transform; the epilogue runs on the last stage's outputs:
        // products with conj(S) go back into the tile for the coil sum, coil
        // images go straight to the grid
        fft_lds_tile<L, +1, W, LDW, true, NT>(lds, twl, [&](int l, int c, float2 v, int p, int i) {
            const int coil = c0 + c;
            float2 r = make_float2(0.0f, 0.0f);
            if (coil < C) {
                r = alt_sign(v, l);
                r = cscale(r, kline[l]);
                if (!xy_in || (a.crop && (l < a.lo[2] || l > a.hi[2]))) r = make_float2(0.0f, 0.0f);
                if (combine) r = cmulc(spre[p][i], r);
                else g[l * C + coil] = r;
            }
            if (combine) lds[l * LDW + c] = r;
        });
        if (combine) {
            __syncthreads();
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const int l = threadIdx.x + r * NT;
                if (l < L) {
                    float2 sum = make_float2(0.0f, 0.0f);
#pragma unroll
                    for (int c = 0; c < W; ++c) sum = cadd(sum, lds[l * LDW + c]);
                    acc[r] = cadd(acc[r], sum);
                }
            }
        }
        if (c0 + W < C) __syncthreads();
    }
    if (!combine) return;
    double pr = 0.0, pi = 0.0;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int l = threadIdx.x + r * NT;
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
            for (int w = 0; w < NT / 64; ++w) { sr += red[2 * w]; si += red[2 * w + 1]; }
            a.part[2 * (long long)outer] = sr;
            a.part[2 * (long long)outer + 1] = si;
        }
    }
}

// ------------------------------------------- inverse z pass over all bins
// The bins of one launch share every cell line's coil maps and deapodization
// line.  k_fft_zpost_mb takes one (x, y) cell line for a group of bins of the
// launch: the line's maps at this thread's epilogue positions (two coil
// chunks, at most 2W coils) are loaded into registers once and used for
// every bin of the group, and the next bin's first grid chunk is fetched
// while the current bin is transformed.  Element for element it performs the
// same operations in the same order as k_fft_zpost, so the results are
// bit-identical.  (A forward counterpart holding the maps for the prologue
// p * S_c * kft was slower than k_fft_zpre with the maps shared through L2:
// 5.26 against 4.72 ms per launch of eight 256x256x192 16-coil bins.)

// The launch's bins g0 <= bin < g1 that are not finished, as a bit mask
// (nb <= 64; one load per lane, the same mask in every wave of the block).
__device__ __forceinline__ uint64_t active_bins(const int* done, int nb, int g0, int g1) {
    const int lane = threadIdx.x % 64;
    return __ballot(lane >= g0 && lane < g1 && lane < nb && !(done && done[lane]));
}

// Workgroup k of the multi-bin z pass with bpw bins per workgroup: its cell
// line and its bins [*g0, *g1).  With ng = ceil(nb / bpw) groups, the ng
// workgroups of one line get ids 8 apart (one XCD, close in time, so the
// line's maps are read from L2 after the first) whenever nxy is a multiple
// of 8, as zpass_line does for single bins.
__device__ __forceinline__ unsigned zpass_group_line(unsigned k, unsigned nxy, int nb, int bpw, int* g0,
                                                     int* g1) {
    const unsigned ng = (unsigned)((nb + bpw - 1) / bpw);
    unsigned xy, g;
    if (ng <= 1 || (nxy & 7)) {
        g = k / nxy;
        xy = k - g * nxy;
    } else {
        const unsigned s = k >> 3;
        g = s % ng;
        xy = (s / ng) * 8 + (k & 7);
    }
    *g0 = (int)g * bpw;
    *g1 = *g0 + bpw < nb ? *g0 + bpw : nb;
    return xy;
}

// Lowest bin left in the mask (nb when none), removed from the mask.
__device__ __forceinline__ int pop_bin(uint64_t& m, int nb) {
    if (!m) return nb;
    const int b = __ffsll((unsigned long long)m) - 1;
    m &= m - 1;
    return b;
}

// Inverse z pass of E^H with the fused epilogue and coil combination for
// every bin: q[bin][cell] = sum_c conj(S_c) * (IFFT * kft * scale, cropped),
// and the block's <p, q> partial of each bin in a.part[bin * nxy + xy].
template <int L, int W, int NT>
__global__ __launch_bounds__(NT) void k_fft_zpost_mb(ZArgs a) {
    constexpr int LDW = W + 1;
    constexpr int RPT = (L + NT - 1) / NT;
    __shared__ float2 lds[L * LDW];
    __shared__ float2 twl[L];
    __shared__ float kline[L];      // kernelFT * scale along the line
    __shared__ double red[2 * (NT / 64)];
    int g0, g1;
    const unsigned xy = zpass_group_line(blockIdx.x, (unsigned)a.nxy, a.nb, a.bpw, &g0, &g1);
    const long long cell0 = (long long)xy * L;
    for (int e = threadIdx.x; e < L; e += NT) {
        twl[e] = a.tw[e];
        kline[e] = a.kft ? a.kft[cell0 + e] * a.scale : a.scale;
    }
    const int gx = (int)(xy / (unsigned)a.ny), gy = (int)(xy - (unsigned)gx * (unsigned)a.ny);
    const bool xy_in = !a.crop || (gx >= a.lo[0] && gx <= a.hi[0] && gy >= a.lo[1] && gy <= a.hi[1]);
    const long long gline = gx * a.xp + gy * (long long)(L * a.C);
    const int C = a.C;
    const float2* ml = a.maps + cell0 * C;
    using LM = LastStageMap<L, W, NT>;
    // S_c at every output of this thread's last-stage butterflies, both chunks
    float2 sm[2][LM::PER][LM::R];
#pragma unroll
    for (int ci = 0; ci < 2; ++ci)
#pragma unroll
        for (int p = 0; p < LM::PER; ++p) {
            if (!LM::valid(p)) continue;
            const int coil = min(ci * W + LM::col(p), C - 1);
#pragma unroll
            for (int i = 0; i < LM::R; ++i)
                sm[ci][p][i] = (ci == 0 || C > W) ? ml[LM::line(p, i) * C + coil] : make_float2(0.0f, 0.0f);
        }
    TileRegs<L, W, NT, float2> next;
    auto load_grid = [&](int b, int c0) {
        const float2* g = a.g + b * a.bp + gline;
        next.load([&](int l, int c) { return g[l * C + min(c0 + c, C - 1)]; });
    };
    uint64_t left = active_bins(a.done, a.nb, g0, g1);
    int b = pop_bin(left, a.nb);
    if (b < a.nb) load_grid(b, 0);
    while (b < a.nb) {
        const int bn = pop_bin(left, a.nb);
        float2 acc[RPT];
#pragma unroll
        for (int r = 0; r < RPT; ++r) acc[r] = make_float2(0.0f, 0.0f);
        auto chunk = [&](auto ci_) {
            constexpr int ci = decltype(ci_)::value;
            constexpr int c0 = ci * W;
            __syncthreads();   // the previous tile's reads done
            next.template to_lds<LDW>(lds, [&](int l, int c, float2 v) {
                if (c0 + c >= C) v = make_float2(0.0f, 0.0f);
                return alt_sign(v, l);
            });
            __syncthreads();
            // the next tile: this bin's second chunk, or the next bin's first
            if (ci == 0 && C > W) load_grid(b, W);
            else if (bn < a.nb) load_grid(bn, 0);
            fft_lds_tile<L, +1, W, LDW, true, NT>(lds, twl, [&](int l, int c, float2 v, int p, int i) {
                float2 r = make_float2(0.0f, 0.0f);
                if (c0 + c < C) {
                    r = alt_sign(v, l);
                    r = cscale(r, kline[l]);
                    if (!xy_in || (a.crop && (l < a.lo[2] || l > a.hi[2]))) r = make_float2(0.0f, 0.0f);
                    r = cmulc(sm[ci][p][i], r);
                }
                lds[l * LDW + c] = r;
            });
            __syncthreads();
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const int l = threadIdx.x + r * NT;
                if (l < L) {
                    float2 sum = make_float2(0.0f, 0.0f);
#pragma unroll
                    for (int c = 0; c < W; ++c) sum = cadd(sum, lds[l * LDW + c]);
                    acc[r] = cadd(acc[r], sum);
                }
            }
        };
        chunk(std::integral_constant<int, 0>{});
        if (C > W) chunk(std::integral_constant<int, 1>{});
        double pr = 0.0, pi = 0.0;
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int l = threadIdx.x + r * NT;
            if (l < L) {
                const long long idx = b * a.ng + cell0 + l;
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
                for (int w = 0; w < NT / 64; ++w) { sr += red[2 * w]; si += red[2 * w + 1]; }
                const long long outer = (long long)b * a.nxy + xy;
                a.part[2 * outer] = sr;
                a.part[2 * (long long)outer + 1] = si;
            }
        }
        b = bn;
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

// Column chunk of the strided passes.
#ifndef CGS_PLAIN_WS
#define CGS_PLAIN_WS 32
#endif
#ifndef CGS_PLAIN_WB
#define CGS_PLAIN_WB 16
#endif
#ifndef CGS_PLAIN_WF
#define CGS_PLAIN_WF 16
#endif
static constexpr int plain_w(int L, bool far) { return far ? CGS_PLAIN_WF : (L <= 256 ? CGS_PLAIN_WS : CGS_PLAIN_WB); }

// Coil chunk of the z pass: 8 coils per LDS tile (small tiles keep four or
// more blocks resident per CU, which hides the HBM latency better than wider
// tiles with fewer blocks).
#ifndef CGS_ZPASS_W
#define CGS_ZPASS_W 8
#endif
#ifndef CGS_ZPASS_NT
#define CGS_ZPASS_NT 256
#endif
int zpass_w(int L, int C) {
    (void)L;
    (void)C;
    return CGS_ZPASS_W;
}

static inline bool plain_far(const PlainArgs& a) { return a.ls * (long long)sizeof(float2) >= (1LL << 20); }

template <int L, int DIR, bool FAR>
static void plain_launch_far(const PlainArgs& a, dim3 grid, hipStream_t s) {
    constexpr int W = plain_w(L, FAR);
    constexpr int CPB = plain_cpb(FAR);
    constexpr int NS = plan_nstages(L);
    // largest byte offset from a chunk's first column, and from the base of
    // a line group of the first or last stage
    const long long span = ((long long)(L - 1) * a.ls + W) * (long long)sizeof(float2);
    constexpr int SG = NS == 1 ? 1 : (L / plan_radix(L, 0) > plan_ls(L, NS - 1) ? L / plan_radix(L, 0)
                                                                               : plan_ls(L, NS - 1));
    const long long seg = ((long long)(SG - 1) * a.ls + W) * (long long)sizeof(float2);
    if (span < (1LL << 32)) k_fft_plain<L, W, DIR, CPB, kOff32><<<grid, CGS_THREADS, 0, s>>>(a);
    else if (seg < (1LL << 32)) k_fft_plain<L, W, DIR, CPB, kOffSeg><<<grid, CGS_THREADS, 0, s>>>(a);
    else k_fft_plain<L, W, DIR, CPB, kOff64><<<grid, CGS_THREADS, 0, s>>>(a);
}

template <int L, int DIR>
static void plain_launch(const PlainArgs& a, dim3 grid, hipStream_t s) {
    if (plain_far(a)) plain_launch_far<L, DIR, true>(a, grid, s);
    else plain_launch_far<L, DIR, false>(a, grid, s);
}

void plain_geometry(int L, long long ls, int* W, int* cpb) {
    const bool far = ls * (long long)sizeof(float2) >= (1LL << 20);
    *W = plain_w(L, far);
    *cpb = plain_cpb(far);
}

hipError_t launch_fft_plain(int L, int dir, const PlainArgs& a, long long nouter, hipStream_t s) {
    const int W = plain_w(L, plain_far(a));
    const long long nchunk = (a.ncols + W - 1) / W;
    const int cpb = plain_cpb(plain_far(a));
    const long long ngrp = (nchunk + cpb - 1) / cpb;
    const dim3 grid((unsigned)(nouter * (a.glist ? a.gstride : ngrp)));
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

template <int L, int KW>
static void xgrid_launch(const XGridArgs& xa, dim3 grid, hipStream_t s) {
    constexpr int W = 16, CPB = plain_cpb(true);
    constexpr int NS = plan_nstages(L);
    const PlainArgs& a = xa.x;
    const long long span = ((long long)(L - 1) * a.ls + W) * (long long)sizeof(float2);
    constexpr int SG = L / plan_radix(L, 0) > plan_ls(L, NS - 1) ? L / plan_radix(L, 0) : plan_ls(L, NS - 1);
    const long long seg = ((long long)(SG - 1) * a.ls + W) * (long long)sizeof(float2);
    (void)span;
    if (seg < (1LL << 32)) k_fft_xgrid<L, W, CPB, kOffSeg, KW><<<grid, CGS_THREADS, 0, s>>>(xa);
    else k_fft_xgrid<L, W, CPB, kOff64, KW><<<grid, CGS_THREADS, 0, s>>>(xa);
}

bool xgrid_supported(int L, int C, long long ls) {
    return (L == 64 || L == 96 || L == 128 || L == 160 || L == 192 || L == 224 || L == 256 || L == 320 ||
            L == 384 || L == 448 || L == 512) && C % 16 == 0 && plain_far(PlainArgs{nullptr, 0, ls}) &&
           plain_w(L, true) == 16;
}

hipError_t launch_fft_xgrid(int L, int kw, const XGridArgs& xa, long long nouter, hipStream_t s) {
    if (!xgrid_supported(L, xa.C, xa.x.ls)) return hipErrorInvalidValue;
    const long long nchunk = (xa.x.ncols + 15) / 16;
    const long long ngrp = (nchunk + plain_cpb(true) - 1) / plain_cpb(true);
    const dim3 grid((unsigned)(nouter * ngrp));
#define CGS_XG(n, k) \
    case n: xgrid_launch<n, k>(xa, grid, s); break;
#define CGS_XG_KW(k)                                                                            \
    switch (L) {                                                                                \
        CGS_XG(64, k) CGS_XG(96, k) CGS_XG(128, k) CGS_XG(160, k) CGS_XG(192, k) CGS_XG(224, k)  \
        CGS_XG(256, k) CGS_XG(320, k) CGS_XG(384, k) CGS_XG(448, k) CGS_XG(512, k)              \
        default: return hipErrorInvalidValue;                                                   \
    }
    switch (kw) {
        case 3: CGS_XG_KW(3); break;
        case 5: CGS_XG_KW(5); break;
        case 7: CGS_XG_KW(7); break;
        default: return hipErrorInvalidValue;
    }
#undef CGS_XG_KW
#undef CGS_XG
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


// The multi-bin inverse z pass: launches of 2 to 64 bins whose coils fit two
// chunks (the line's maps stay in registers) that combine the coils (E^H of
// the CG iteration), when the context allows two or more bins per workgroup.
static bool zpass_multibin(const ZArgs& a) {
    return a.nb > 1 && a.nb <= 64 && a.maps != nullptr && a.C <= 2 * CGS_ZPASS_W && a.bpw > 1;
}

// z-pass block shape: 8 coils per LDS tile, 256 threads.
template <int L>
static hipError_t zpre_w(int W, const ZArgs& a, long long nouter, hipStream_t s) {
    constexpr int ZW = CGS_ZPASS_W, NT = CGS_ZPASS_NT;
    const dim3 grid((unsigned)nouter);
    if (W != ZW) return hipErrorInvalidValue;
    if (a.p && a.kft) k_fft_zpre<L, ZW, NT, true, true><<<grid, NT, 0, s>>>(a);
    else if (a.p) k_fft_zpre<L, ZW, NT, true, false><<<grid, NT, 0, s>>>(a);
    else if (a.kft) k_fft_zpre<L, ZW, NT, false, true><<<grid, NT, 0, s>>>(a);
    else k_fft_zpre<L, ZW, NT, false, false><<<grid, NT, 0, s>>>(a);
    return hipGetLastError();
}

template <int L>
static hipError_t zpost_w(int W, const ZArgs& a, long long nouter, hipStream_t s) {
    constexpr int ZW = CGS_ZPASS_W, NT = CGS_ZPASS_NT;
    const dim3 grid((unsigned)nouter);
    if (W != ZW) return hipErrorInvalidValue;
    if (zpass_multibin(a) && a.q)
        k_fft_zpost_mb<L, ZW, NT><<<dim3((unsigned)(a.nxy * ((a.nb + a.bpw - 1) / a.bpw))), NT, 0, s>>>(a);
    else k_fft_zpost<L, ZW, NT><<<grid, NT, 0, s>>>(a);
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
