// cgs_internal.h -- device data structures and helpers shared by the HIP
// translation units of libcgs_hip.so (gfx950 / CDNA4).
//
// HBM layouts (all complex values are float2 = complex64, interleaved):
//   coil grids  G[bin][x][y][z][c]   coil index fastest, so one grid cell of all
//                                    coils is one contiguous C*8-byte run
//   coil maps   S[x][y][z][c]        same interleaving (shared by all bins)
//   CG vectors  x,r,p,q[bin][x][y][z]
//   samples     per bin, sorted by gridding tile:
//                 Y[m][c]            complex data (density weight applied)
//                 base[m] = int4(bx, by, bz, original sample index)
//                 vals[m][VS]        per-axis KB factors in FP64: x at [0,KW),
//                                    y at [KW,2KW), z at [2KW,3KW); VS = val_stride(KW)
//               tile_off[ntiles+1]   CSR offsets of each tile's samples
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define CGS_THREADS 256
#define CGS_MAX_KW 7          // widest per-axis kernel window supported

namespace cgs {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
    return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// doubles per sample in vals[]: 3 axes x KW, rounded up to an even count so
// every sample's row starts 16-byte aligned.
__host__ __device__ constexpr int val_stride(int kw) { return (3 * kw + 1) & ~1; }

// Geometry of one plan, passed by value to kernels.
struct Geom {
    int nx, ny, nz;      // oversampled grid
    int C;               // coils
    long long ng;        // nx*ny*nz
    long long xp;        // coil-grid x-plane pitch in elements (>= ny*nz*C)
    long long bp;        // coil-grid pitch of one bin (nx*xp)
};

// Element (x, y, z, coil 0) of one bin's coil grid [x][y][z][c]: the x planes
// are xp elements apart (padded, see coil_grid_pitch), y and z are dense.
__host__ __device__ inline long long grid_at(const Geom& g, long long x, long long y, long long z) {
    return x * g.xp + (y * g.nz + z) * g.C;
}

// One gridding-tile decomposition of the grid.
struct Tiles {
    int tx, ty, tz;      // tile extent in cells
    int ntx, nty, ntz;   // tiles per axis
    int ntiles;
};

}  // namespace cgs
