// cgs_nufft.hip -- Kaiser-Bessel degridding (gather) and gridding (tile
// accumulation) over the coil-interleaved grid.
//
// The tap set and weights are the reference's (Reco.cpp:698-787): for a sample
// with per-axis factors vx, vy, vz (FP64), the weight of grid cell
// (bx+ox, by+oy, bz+oz) is w = float((vx[ox]*vy[oy])*vz[oz]) and the tap is kept
// when double(w) > thr; taps are visited x-major, then y, then z, like the
// candidate list of Reco.cpp:719-729.
//
// Degridding (nufftFWD, Reco.cpp:539-550; DCF of Reco.cpp:243-254 fused):
// one lane per (sample, coil); the lanes of a sample read the C consecutive
// coil values of each tap cell, so every tap is one coalesced C*8-byte load.
// Samples are tile-sorted, so neighbouring samples hit the same L2 lines.
//
// Gridding (nufftBWD, Reco.cpp:484-504): one workgroup owns one TX*TY*TZ
// tile of cells for all coils and keeps it in registers (no atomics, no
// halo write-back).  It scans the samples of the 27 neighbouring tiles (9
// contiguous ranges of the tile-sorted array), compacts those whose KB window
// meets the tile into an LDS list in a fixed order, and every thread adds the
// contributions to the (column, coil) pairs it owns.  The result is
// deterministic and every cell is written exactly once.
#include "cgs_internal.h"
#include "cgs_kernels.h"

namespace cgs {

static inline int pow2_at_least(int c, int cap) {
    int p = 1;
    while (p < c && p < cap) p <<= 1;
    return p;
}

// ------------------------------------------------------------ degridding
// The loads of one x-plane of the window (KW*KW cells) are issued without
// conditions -- indices are clamped into the grid, and a tap outside the
// support or below the threshold gets weight 0 -- so each lane keeps KW*KW
// independent loads in flight instead of waiting on one tap at a time.
template <int KW, int CW>
__global__ __launch_bounds__(CGS_THREADS) void k_degrid(GridArgs a) {
    constexpr int VS = val_stride(KW);
    const long long e = (long long)blockIdx.x * CGS_THREADS + threadIdx.x;
    const long long s = a.s.first + e / CW;
    const int lane_c = (int)(e % CW);
    if (e / CW >= a.s.m) return;
    const int bin = a.s.sbin[s];
    if (a.done && a.done[bin]) return;
    const int4 bs = a.s.base[s];
    const double* __restrict__ v = a.s.vals + s * VS;
    double vx[KW], vy[KW], vz[KW];
    int iy[KW], iz[KW];
#pragma unroll
    for (int o = 0; o < KW; ++o) {
        vx[o] = v[o];
        vy[o] = v[KW + o];
        vz[o] = v[2 * KW + o];
        iy[o] = min(max(bs.y + o, 0), a.g.ny - 1);
        iz[o] = min(max(bs.z + o, 0), a.g.nz - 1);
    }
    // Rows (columns) whose y (z) factor is 0 give weight 0 to every tap: their
    // loads are pointed at the centre row (column), whose lines the lanes load
    // anyway, so they cost no L2 request (a finite value times 0 adds nothing).
#pragma unroll
    for (int o = 0; o < KW; ++o) {
        if (vy[o] == 0.0) iy[o] = iy[KW / 2];
        if (vz[o] == 0.0) iz[o] = iz[KW / 2];
    }
    const int C = a.g.C;
    const float2* __restrict__ G = a.G + (long long)(bin - a.b0) * a.g.bp;
    const long long ys = (long long)a.g.nz * C;
    const float d = a.s.dcf ? a.s.dcf[s] : 1.0f;
    const float thrf = a.thrf;
    for (int c = lane_c; c < C; c += CW) {
        float2 acc = make_float2(0.0f, 0.0f);
        // one x plane in flight at a time: 25 loads per lane, about 110
        // VGPRs, four waves per SIMD (unrolling the planes doubles the VGPRs
        // and halves the waves, which measured 5 % slower)
#pragma unroll 1
        for (int ox = 0; ox < KW; ++ox) {
            const double wx = vx[ox];
            if (wx == 0.0) continue;
            const long long xo = (long long)min(max(bs.x + ox, 0), a.g.nx - 1) * a.g.xp;
            float2 gv[KW][KW];
#pragma unroll
            for (int oy = 0; oy < KW; ++oy) {
                const long long row = xo + iy[oy] * ys;
#pragma unroll
                for (int oz = 0; oz < KW; ++oz) gv[oy][oz] = G[row + iz[oz] * C + c];
            }
#pragma unroll
            for (int oy = 0; oy < KW; ++oy) {
                const double vxy = wx * vy[oy];
#pragma unroll
                for (int oz = 0; oz < KW; ++oz) {
                    float w = (float)(vxy * vz[oz]);
                    w = w >= thrf ? w : 0.0f;
                    acc.x = fmaf(w, gv[oy][oz].x, acc.x);
                    acc.y = fmaf(w, gv[oy][oz].y, acc.y);
                }
            }
        }
        a.Y[s * C + c] = make_float2(acc.x * d, acc.y * d);
    }
}

// -------------------------------------------------------------- gridding
// Candidate samples of tile (itx, ity, itz) of one bin: the samples of the
// 3 x 3 tile columns around it (z tiles itz-1 .. itz+1) -- nine CSR ranges,
// walked as one flat sequence in (dx, dy, sample) order -- whose window meets
// the tile.  The block compacts them 256 at a time with a ballot and prefix
// (so the order is fixed) into the LDS array `list` and calls
// chunk(total) after each compaction, all threads together.
template <int KW, int TX, int TY, int TZ, class Chunk>
__device__ __forceinline__ void tile_candidates(const GridArgs& a, int bin, int itx, int ity, int itz, int* list,
                                                int* wcount, int* rbeg, int* rlen, Chunk chunk) {
    constexpr int NWAVE = CGS_THREADS / 64;
    const int ntiles = a.t.ntiles;
    const int* __restrict__ off = a.tile_off + (long long)bin * (ntiles + 1);
    const int x0 = itx * TX, y0 = ity * TY, z0 = itz * TZ;
    const int tid = threadIdx.x;
    const int lane = tid % 64, wid = tid / 64;
    const int tz_lo = itz > 0 ? itz - 1 : 0;
    const int tz_hi = itz + 1 < a.t.ntz ? itz + 1 : a.t.ntz - 1;
    if (tid < 9) {
        const int jx = itx + tid / 3 - 1, jy = ity + tid % 3 - 1;
        int rb = 0, rl = 0;
        if (jx >= 0 && jx < a.t.ntx && jy >= 0 && jy < a.t.nty) {
            const int tcol = (jx * a.t.nty + jy) * a.t.ntz;
            rb = off[tcol + tz_lo];
            rl = off[tcol + tz_hi + 1] - rb;
        }
        rbeg[tid] = rb;
        rlen[tid] = rl;
    }
    __syncthreads();
    int rpre[10];
    rpre[0] = 0;
#pragma unroll
    for (int r = 0; r < 9; ++r) rpre[r + 1] = rpre[r] + rlen[r];
    const int nflat = rpre[9];
    for (int f0 = 0; f0 < nflat; f0 += CGS_THREADS) {
        const int f = f0 + tid;
        bool hit = false;
        int s = 0;
        if (f < nflat) {
            int r = 0;
#pragma unroll
            for (int q = 1; q < 9; ++q) r += f >= rpre[q] ? 1 : 0;
            s = rbeg[r] + (f - rpre[r]);
            const int4 bs = a.s.base[s];
            hit = bs.x < x0 + TX && bs.x + KW > x0 && bs.y < y0 + TY && bs.y + KW > y0 &&
                  bs.z < z0 + TZ && bs.z + KW > z0;
        }
        const unsigned long long bal = __ballot(hit);
        const int wpre = __popcll(bal & ((1ULL << lane) - 1ULL));
        if (lane == 0) wcount[wid] = __popcll(bal);
        __syncthreads();
        int wbase = 0, total = 0;
#pragma unroll
        for (int w = 0; w < NWAVE; ++w) {
            if (w < wid) wbase += wcount[w];
            total += wcount[w];
        }
        if (hit) list[wbase + wpre] = s;
        __syncthreads();
        chunk(total);
    }
}

// A thread owns CG coils of one column (x, y) of the tile and its TZ cells
// along z: each tap weight is formed once per CG coils.  The candidates come
// from the tile's precomputed list (a.tlist) or, without one, from the scan
// of tile_candidates; either way in the same order, so the sums are the same.
template <int KW, int TX, int TY, int TZ, int CP, int CG>
__global__ __launch_bounds__(CGS_THREADS) void k_grid(GridArgs a) {
    constexpr int VS = val_stride(KW);
    constexpr int NCOL = TX * TY;
    constexpr int NG = CP / CG;              // coil groups
    constexpr int NPAIR = NCOL * NG;
    constexpr int NPT = (NPAIR + CGS_THREADS - 1) / CGS_THREADS;
    constexpr int NWAVE = CGS_THREADS / 64;
    constexpr int SUB = 64;                  // candidates staged in LDS at a time
    __shared__ int list[CGS_THREADS];
    __shared__ int wcount[NWAVE];
    __shared__ int rbeg[9], rlen[9];
    __shared__ int4 sbase[SUB];
    __shared__ double svals[SUB * VS];
    __shared__ float2 sy[SUB * CP];

    const int ntiles = a.t.ntiles;
    const int lbin = blockIdx.x / ntiles;
    const int tile = blockIdx.x % ntiles;
    const int bin = a.b0 + lbin;
    if (a.done && a.done[bin]) return;
    const int itz = tile % a.t.ntz;
    const int ity = (tile / a.t.ntz) % a.t.nty;
    const int itx = tile / (a.t.ntz * a.t.nty);
    if (a.tneed && !a.tneed[(long long)bin * (a.t.nty * a.t.ntz) + ity * a.t.ntz + itz]) return;
    const int x0 = itx * TX, y0 = ity * TY, z0 = itz * TZ;
    const int C = a.g.C;

    float2 acc[NPT][CG][TZ];
#pragma unroll
    for (int k = 0; k < NPT; ++k)
#pragma unroll
        for (int q = 0; q < CG; ++q)
#pragma unroll
            for (int z = 0; z < TZ; ++z) acc[k][q][z] = make_float2(0.0f, 0.0f);

    const int tid = threadIdx.x;
    const int lane = tid % 64, wid = tid / 64;
    // grid rows x (of this tile) that this wave's threads own, conservatively
    int wxlo = TX, wxhi = -1;
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
        const int e0 = wid * 64 + k * CGS_THREADS;
        if (e0 < NPAIR) {
            const int e1 = min(e0 + 63, NPAIR - 1);
            wxlo = min(wxlo, (e0 / NG) / TY);
            wxhi = max(wxhi, (e1 / NG) / TY);
        }
    }
    const int gxlo = x0 + wxlo, gxhi = x0 + wxhi;

    // stage and accumulate `total` candidates, the j-th being entry(j)
    auto accumulate = [&](int total, auto entry) {
        for (int j0 = 0; j0 < total; j0 += SUB) {
            const int nsub = min(SUB, total - j0);
            // --- stage window origins, KB factors and (weighted) values in LDS
            for (int t = tid; t < nsub * VS; t += CGS_THREADS) {
                const int j = t / VS, o = t % VS;
                svals[t] = a.s.vals[(long long)entry(j0 + j) * VS + o];
            }
            for (int t = tid; t < nsub * CP; t += CGS_THREADS) {
                const int j = t / CP, c = t % CP;
                const int sj = entry(j0 + j);
                float2 yv = make_float2(0.0f, 0.0f);
                if (a.c0 + c < C) {
                    yv = a.Y[(long long)sj * C + a.c0 + c];
                    if (a.gdcf) yv = cscale(yv, a.gdcf[sj]);
                }
                sy[t] = yv;
            }
            if (tid < nsub) sbase[tid] = a.s.base[entry(j0 + tid)];
            __syncthreads();
            // --- accumulate, walking only the candidates whose window covers
            // one of the wave's rows (the others add nothing to its cells),
            // in ascending order
            unsigned long long todo = __ballot(lane < nsub && sbase[lane].x <= gxhi && sbase[lane].x + KW > gxlo);
            while (todo) {
                const int j = __builtin_ctzll(todo);
                todo &= todo - 1ULL;
                const int4 bs = sbase[j];
                const double* v = svals + j * VS;
#pragma unroll
                for (int k = 0; k < NPT; ++k) {
                    const int e = tid + k * CGS_THREADS;
                    const int g = e % NG, col = e / NG;
                    if ((NPAIR % CGS_THREADS == 0 || e < NPAIR) && a.c0 + g * CG < C) {
                        const int ox = x0 + col / TY - bs.x;
                        const int oy = y0 + col % TY - bs.y;
                        if (ox >= 0 && ox < KW && oy >= 0 && oy < KW) {
                            const double vxy = v[ox] * v[KW + oy];
                            if (vxy != 0.0) {
                                // staged values of coils >= C are 0
                                float2 yv[CG];
#pragma unroll
                                for (int q = 0; q < CG; ++q) yv[q] = sy[j * CP + g * CG + q];
#pragma unroll
                                for (int z = 0; z < TZ; ++z) {
                                    const int oz = z0 + z - bs.z;
                                    if (oz >= 0 && oz < KW) {
                                        const float w = (float)(vxy * v[2 * KW + oz]);
                                        if (w >= a.thrf) {
#pragma unroll
                                            for (int q = 0; q < CG; ++q) {
                                                acc[k][q][z].x = fmaf(w, yv[q].x, acc[k][q][z].x);
                                                acc[k][q][z].y = fmaf(w, yv[q].y, acc[k][q][z].y);
                                            }
                                        }
                                    }
                                }
                            }
                        }
                    }
                }
            }
            __syncthreads();
        }
    };
    if (a.tlist) {
        // the tile's list, 256 entries at a time through LDS (one coalesced load)
        const long long g0 = a.tlist_off[(long long)bin * ntiles + tile];
        const int total = (int)(a.tlist_off[(long long)bin * ntiles + tile + 1] - g0);
        const int* __restrict__ tl = a.tlist + g0;
        for (int c0 = 0; c0 < total; c0 += CGS_THREADS) {
            const int n = min(CGS_THREADS, total - c0);
            if (tid < n) list[tid] = tl[c0 + tid];
            __syncthreads();
            accumulate(n, [&](int j) { return list[j]; });   // ends with a barrier (n > 0)
        }
    } else {
        tile_candidates<KW, TX, TY, TZ>(a, bin, itx, ity, itz, list, wcount, rbeg, rlen, [&](int total) {
            accumulate(total, [&](int j) { return list[j]; });
        });
    }
    // --- write the tile (every cell of the grid belongs to exactly one tile)
    float2* __restrict__ G = a.G + (long long)lbin * a.g.bp;
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
        const int e = tid + k * CGS_THREADS;
        const int g = e % NG, col = e / NG;
        if ((NPAIR % CGS_THREADS == 0 || e < NPAIR) && a.c0 + g * CG < C) {
            const int x = x0 + col / TY, y = y0 + col % TY;
            if (x < a.g.nx && y < a.g.ny) {
                const long long row = grid_at(a.g, x, y, z0);
#pragma unroll
                for (int z = 0; z < TZ; ++z)
                    if (z0 + z < a.g.nz) {
#pragma unroll
                        for (int q = 0; q < CG; ++q)
                            if (a.c0 + g * CG + q < C) G[row + z * C + a.c0 + g * CG + q] = acc[k][q][z];
                    }
            }
        }
    }
}

// The candidate lists of the tiles of bins b0.. (one block per tile, the
// scan of k_grid): counts per tile (counts != null), else the lists.  Tiles
// that gridding skips (tneed) get none.
template <int KW, int TX, int TY, int TZ>
__global__ __launch_bounds__(CGS_THREADS) void k_grid_lists(GridArgs a, int* counts) {
    constexpr int NWAVE = CGS_THREADS / 64;
    __shared__ int list[CGS_THREADS];
    __shared__ int wcount[NWAVE];
    __shared__ int rbeg[9], rlen[9];
    const int ntiles = a.t.ntiles;
    const int bin = a.b0 + (int)(blockIdx.x / ntiles);
    const int tile = blockIdx.x % ntiles;
    const int itz = tile % a.t.ntz;
    const int ity = (tile / a.t.ntz) % a.t.nty;
    const int itx = tile / (a.t.ntz * a.t.nty);
    const long long gt = (long long)bin * ntiles + tile;
    if (a.tneed && !a.tneed[(long long)bin * (a.t.nty * a.t.ntz) + ity * a.t.ntz + itz]) {
        if (counts && threadIdx.x == 0) counts[gt] = 0;
        return;
    }
    long long pos = counts ? 0 : a.tlist_off[gt];
    int n = 0;
    tile_candidates<KW, TX, TY, TZ>(a, bin, itx, ity, itz, list, wcount, rbeg, rlen, [&](int total) {
        if (!counts)
            for (int t = threadIdx.x; t < total; t += CGS_THREADS) ((int*)a.tlist)[pos + n + t] = list[t];
        n += total;
        __syncthreads();   // list is rewritten by the next compaction
    });
    if (counts && threadIdx.x == 0) counts[gt] = n;
}

// ------------------------------------------------------ coordinate gradient
// Lane per (sample, coil) as in k_degrid; the three partial sums of a sample
// are reduced over its CW lanes with shuffles.  Taps are the forward's kept
// taps (float((vx*vy)*vz) > thr); the derivative of each is
// (dvx*vy*vz, vx*dvy*vz, vx*vy*dvz) in FP64, accumulated in FP32.
template <int KW, int CW>
__global__ __launch_bounds__(CGS_THREADS) void k_grad_k(GradKArgs a) {
    constexpr int VS = val_stride(KW);
    const long long e = (long long)blockIdx.x * CGS_THREADS + threadIdx.x;
    const long long s = e / CW;
    const int lane_c = (int)(e % CW);
    const bool valid = s < a.m;
    float g0 = 0.0f, g1 = 0.0f, g2 = 0.0f;
    if (valid) {
        const int4 bs = a.base[s];
        const double* __restrict__ v = a.vals + s * VS;
        const double* __restrict__ dv = a.dvals + s * VS;
        const int C = a.g.C;
        for (int c = lane_c; c < C; c += CW) {
            const float2 bv = a.B[s * C + c];
            for (int ox = 0; ox < KW; ++ox) {
                const int xg = bs.x + ox;
                if (xg < 0 || xg >= a.g.nx) continue;
                for (int oy = 0; oy < KW; ++oy) {
                    const int yg = bs.y + oy;
                    if (yg < 0 || yg >= a.g.ny) continue;
                    const double vxy = v[ox] * v[KW + oy];
                    const double dxy = dv[ox] * v[KW + oy];
                    const double xdy = v[ox] * dv[KW + oy];
                    const long long row = grid_at(a.g, xg, yg, 0);
#pragma unroll
                    for (int oz = 0; oz < KW; ++oz) {
                        const int zg = bs.z + oz;
                        if (zg < 0 || zg >= a.g.nz) continue;
                        const float w = (float)(vxy * v[2 * KW + oz]);
                        if (!(w >= a.thrf)) continue;
                        const float2 av = a.A[row + (long long)zg * C + c];
                        const float t = bv.x * av.x + bv.y * av.y;
                        g0 = fmaf(t, (float)(dxy * v[2 * KW + oz]), g0);
                        g1 = fmaf(t, (float)(xdy * v[2 * KW + oz]), g1);
                        g2 = fmaf(t, (float)(vxy * dv[2 * KW + oz]), g2);
                    }
                }
            }
        }
    }
#pragma unroll
    for (int off = CW / 2; off >= 1; off >>= 1) {
        g0 += __shfl_xor(g0, off, 64);
        g1 += __shfl_xor(g1, off, 64);
        g2 += __shfl_xor(g2, off, 64);
    }
    if (valid && lane_c == 0) {
        const long long i = a.base[s].w;
        a.gk[3 * i + 0] = a.scale * g0;
        a.gk[3 * i + 1] = a.scale * g1;
        a.gk[3 * i + 2] = a.scale * g2;
    }
}

// -------------------------------------------------------------- dispatch
hipError_t launch_degrid(int kw, const GridArgs& a, hipStream_t s) {
    const int C = a.g.C;
    const int cw = pow2_at_least(C, 64);
    const long long nthr = a.s.m * cw;
    if (a.s.m == 0) return hipSuccess;
    const dim3 grid((unsigned)((nthr + CGS_THREADS - 1) / CGS_THREADS));
#define CGS_DEG(KWv, CWv) k_degrid<KWv, CWv><<<grid, CGS_THREADS, 0, s>>>(a)
#define CGS_DEG_CW(KWv)                        \
    switch (cw) {                              \
        case 1: CGS_DEG(KWv, 1); break;        \
        case 2: CGS_DEG(KWv, 2); break;        \
        case 4: CGS_DEG(KWv, 4); break;        \
        case 8: CGS_DEG(KWv, 8); break;        \
        case 16: CGS_DEG(KWv, 16); break;      \
        case 32: CGS_DEG(KWv, 32); break;      \
        default: CGS_DEG(KWv, 64); break;      \
    }
    switch (kw) {
        case 3: CGS_DEG_CW(3); break;
        case 5: CGS_DEG_CW(5); break;
        case 7: CGS_DEG_CW(7); break;
        default: return hipErrorInvalidValue;
    }
#undef CGS_DEG_CW
#undef CGS_DEG
    return hipGetLastError();
}

// Gridding tile shapes: cells per tile along x, y, z and coils per thread
// (at most; fewer when the padded coil count is smaller).
static const int kTileCfg[][4] = {{8, 8, 8, 1}, {8, 8, 4, 1}, {4, 8, 8, 1}, {4, 4, 8, 1},
                                  {8, 8, 8, 4}, {8, 8, 4, 4}, {4, 8, 8, 2}, {8, 8, 8, 2}};

int grid_tile_configs() { return (int)(sizeof(kTileCfg) / sizeof(kTileCfg[0])); }

void grid_tile_dims(int tcfg, int* tx, int* ty, int* tz) {
    *tx = kTileCfg[tcfg][0];
    *ty = kTileCfg[tcfg][1];
    *tz = kTileCfg[tcfg][2];
}

template <int KW, int TX, int TY, int TZ, int CG>
static hipError_t grid_cp(const GridArgs& a, hipStream_t s) {
    // more than 32 coils: launches of 32 coils each (launch_grid)
    const int cp = a.g.C > 32 ? 32 : pow2_at_least(a.g.C, 64);
    const dim3 grid((unsigned)((long long)a.nb * a.t.ntiles));
    switch (cp) {
        case 1: case 2: case 4: case 8:
            k_grid<KW, TX, TY, TZ, 8, (CG < 8 ? CG : 8)><<<grid, CGS_THREADS, 0, s>>>(a);
            break;
        case 16: k_grid<KW, TX, TY, TZ, 16, CG><<<grid, CGS_THREADS, 0, s>>>(a); break;
        case 32: k_grid<KW, TX, TY, TZ, 32, CG><<<grid, CGS_THREADS, 0, s>>>(a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int KW>
static hipError_t grid_tile(int tcfg, const GridArgs& a, hipStream_t s) {
    switch (tcfg) {
        case 0: return grid_cp<KW, 8, 8, 8, 1>(a, s);
        case 1: return grid_cp<KW, 8, 8, 4, 1>(a, s);
        case 2: return grid_cp<KW, 4, 8, 8, 1>(a, s);
        case 3: return grid_cp<KW, 4, 4, 8, 1>(a, s);
        case 4: return grid_cp<KW, 8, 8, 8, 4>(a, s);
        case 5: return grid_cp<KW, 8, 8, 4, 4>(a, s);
        case 6: return grid_cp<KW, 4, 8, 8, 2>(a, s);
        case 7: return grid_cp<KW, 8, 8, 8, 2>(a, s);
        default: return hipErrorInvalidValue;
    }
}

static hipError_t grid_coils(int kw, int tcfg, const GridArgs& a, hipStream_t s) {
    switch (kw) {
        case 3: return grid_tile<3>(tcfg, a, s);
        case 5: return grid_tile<5>(tcfg, a, s);
        case 7: return grid_tile<7>(tcfg, a, s);
        default: return hipErrorInvalidValue;
    }
}

// Coil grids of more than 32 coils are gridded 32 coils per launch (every
// launch walks the same candidates; a cell's coils are independent sums).
hipError_t launch_grid(int kw, int tcfg, const GridArgs& a, hipStream_t s) {
    GridArgs b = a;
    for (b.c0 = 0; b.c0 < a.g.C; b.c0 += 32) {
        const hipError_t e = grid_coils(kw, tcfg, b, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <int KW, int TX, int TY, int TZ>
static hipError_t lists_tile(const GridArgs& a, int* counts, hipStream_t s) {
    const dim3 grid((unsigned)((long long)a.nb * a.t.ntiles));
    k_grid_lists<KW, TX, TY, TZ><<<grid, CGS_THREADS, 0, s>>>(a, counts);
    return hipGetLastError();
}

template <int KW>
static hipError_t lists_kw(int tcfg, const GridArgs& a, int* counts, hipStream_t s) {
    if (tcfg < 0 || tcfg >= grid_tile_configs()) return hipErrorInvalidValue;
    const int tx = kTileCfg[tcfg][0], ty = kTileCfg[tcfg][1], tz = kTileCfg[tcfg][2];
    if (tx == 8 && ty == 8 && tz == 8) return lists_tile<KW, 8, 8, 8>(a, counts, s);
    if (tx == 8 && ty == 8 && tz == 4) return lists_tile<KW, 8, 8, 4>(a, counts, s);
    if (tx == 4 && ty == 8 && tz == 8) return lists_tile<KW, 4, 8, 8>(a, counts, s);
    if (tx == 4 && ty == 4 && tz == 8) return lists_tile<KW, 4, 4, 8>(a, counts, s);
    return hipErrorInvalidValue;
}

hipError_t launch_grid_lists(int kw, int tcfg, const GridArgs& a, int* counts, hipStream_t s) {
    switch (kw) {
        case 3: return lists_kw<3>(tcfg, a, counts, s);
        case 5: return lists_kw<5>(tcfg, a, counts, s);
        case 7: return lists_kw<7>(tcfg, a, counts, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_grad_k(int kw, const GradKArgs& a, hipStream_t s) {
    if (a.m == 0) return hipSuccess;
    const int cw = pow2_at_least(a.g.C, 64);
    const dim3 grid((unsigned)((a.m * cw + CGS_THREADS - 1) / CGS_THREADS));
#define CGS_GK(KWv, CWv) k_grad_k<KWv, CWv><<<grid, CGS_THREADS, 0, s>>>(a)
#define CGS_GK_CW(KWv)                         \
    switch (cw) {                              \
        case 1: CGS_GK(KWv, 1); break;         \
        case 2: CGS_GK(KWv, 2); break;         \
        case 4: CGS_GK(KWv, 4); break;         \
        case 8: CGS_GK(KWv, 8); break;         \
        case 16: CGS_GK(KWv, 16); break;       \
        case 32: CGS_GK(KWv, 32); break;       \
        default: CGS_GK(KWv, 64); break;       \
    }
    switch (kw) {
        case 3: CGS_GK_CW(3); break;
        case 5: CGS_GK_CW(5); break;
        case 7: CGS_GK_CW(7); break;
        default: return hipErrorInvalidValue;
    }
#undef CGS_GK_CW
#undef CGS_GK
    return hipGetLastError();
}

}  // namespace cgs
