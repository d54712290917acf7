// cgs_xlists.hip -- per-cell sample lists of the fused gridding (inverse x
// pass with the scatter of E^H done as a gather, k_fft_xgrid in cgs_fft.hip).
//
// Gridding adds every acquired sample into the KW^3 grid cells of its KB
// window (Reco.cpp:492-504).  The inverse x pass transforms, per (y, z) cell
// column, the whole line along x; the fused pass forms that line itself from
// the samples whose window covers (y, z): for each x, the samples with
// bx <= x < bx + KW (bx = the window's first x).  The lists hold, for every
// bin and (y, z), those samples sorted by (bx, sample index), so that each x
// finds its samples by a binary search and sums them in a fixed order.  The
// trajectory is fixed across CG iterations, so the lists are built once when
// the bins are packed: every sample emits a key (cell, bx) per (y, z) of its
// window whose y and z factors are non-zero, a stable radix sort orders the
// keys (ties keep the sample order), each sorted entry gets a record with
// everything the fused pass needs but the sample's values (bx, its float
// weights along x for that cell, the sample index), and each cell's first
// entry is found by a binary search.
#include <hipcub/hipcub.hpp>

#include "cgs_internal.h"
#include "cgs_kernels.h"

namespace cgs {

namespace {

// One key per (sample, oy, oz): (y * nz + z) << XBITS | (bx + XOFF), or
// 0xffffffff (sorted last) where the tap is off the grid or its y / z factor
// is zero; the value is the sample's global index.
template <int KW>
__global__ void k_xlist_keys(const int4* __restrict__ base, const double* __restrict__ vals, long long first,
                             long long m, int ny, int nz, unsigned* __restrict__ keys, int* __restrict__ idx) {
    constexpr int VS = val_stride(KW);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m * KW * KW) return;
    const long long si = t / (KW * KW);
    const int k = (int)(t % (KW * KW)), oy = k / KW, oz = k % KW;
    const long long s = first + si;
    const int4 b = base[s];
    const int y = b.y + oy, z = b.z + oz;
    unsigned key = 0xffffffffu;
    if (y >= 0 && y < ny && z >= 0 && z < nz && vals[s * VS + KW + oy] != 0.0 && vals[s * VS + 2 * KW + oz] != 0.0)
        key = ((unsigned)(y * nz + z) << kXListBits) | (unsigned)(b.x + kXListOff);
    keys[t] = key;
    idx[t] = (int)s;
}

// Record of sorted entry t (a tap of a sample in cell (y, z) = key >> XBITS):
// bx, the KW weights along x times the cell's y and z factors, rounded to
// float as the gridding kernels round them, and the sample index.  Entries
// with the empty key sort after every cell and are never read.
template <int KW>
__global__ void k_xlist_records(const unsigned* __restrict__ keys, const int* __restrict__ idx, long long n,
                                const int4* __restrict__ base, const double* __restrict__ vals, int nz,
                                int* __restrict__ rec) {
    constexpr int VS = val_stride(KW);
    constexpr int RS = xrec_stride(KW);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const unsigned key = keys[t];
    if (key == 0xffffffffu) return;
    const int cell = (int)(key >> kXListBits), y = cell / nz, z = cell - (cell / nz) * nz;
    const int sidx = idx[t];
    const int4 b = base[sidx];
    const double* v = vals + (long long)sidx * VS;
    const double vy = v[KW + (y - b.y)], vz = v[2 * KW + (z - b.z)];
    int r[RS];
    r[0] = b.x;
#pragma unroll
    for (int ox = 0; ox < KW; ++ox) r[1 + ox] = __float_as_int((float)((v[ox] * vy) * vz));
    r[KW + 1] = sidx;
#pragma unroll
    for (int q = KW + 2; q < RS; ++q) r[q] = 0;
    int4* o = reinterpret_cast<int4*>(rec + t * RS);
#pragma unroll
    for (int q = 0; q < RS / 4; ++q) o[q] = make_int4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
}

// off[c] = first sorted position with a key of cell c or above (c <= ncell).
__global__ void k_xlist_offsets(const unsigned* __restrict__ keys, long long n, int ncell, long long* __restrict__ off,
                                long long off0) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncell) return;
    const unsigned target = (unsigned)c << kXListBits;
    long long lo = 0, hi = n;
    while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if (keys[mid] < target) lo = mid + 1; else hi = mid;
    }
    off[c] = off0 + lo;
}

}  // namespace

bool xlist_supported(int nx, int ny, int nz) {
    return nx + kXListOff < (1 << kXListBits) && (long long)ny * nz < (1LL << (32 - kXListBits)) - 1;
}

size_t xlist_sort_bytes(long long n) {
    size_t tb = 0;
    if (n > 0 &&
        hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (const unsigned*)nullptr, (unsigned*)nullptr,
                                           (const int*)nullptr, (int*)nullptr, (int)n, 0, 32) != hipSuccess)
        return 0;
    return tb;
}

hipError_t build_xlists_bin(int kw, const int4* base, const double* vals, long long first, long long m, int ny, int nz,
                            unsigned* keys_tmp, int* idx_tmp, void* sort_tmp, size_t sort_bytes, unsigned* keys_out,
                            int* idx_out, int* rec, long long* off, long long off0, hipStream_t s) {
    const int ncell = ny * nz;
    const long long n = m * kw * kw;
    hipError_t e = hipSuccess;
    if (n > 0) {
        const unsigned blocks = (unsigned)((n + 255) / 256);
        switch (kw) {
            case 3: k_xlist_keys<3><<<blocks, 256, 0, s>>>(base, vals, first, m, ny, nz, keys_tmp, idx_tmp); break;
            case 5: k_xlist_keys<5><<<blocks, 256, 0, s>>>(base, vals, first, m, ny, nz, keys_tmp, idx_tmp); break;
            case 7: k_xlist_keys<7><<<blocks, 256, 0, s>>>(base, vals, first, m, ny, nz, keys_tmp, idx_tmp); break;
            default: return hipErrorInvalidValue;
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
        // stable LSD radix sort (empty taps, all ones, sort last and are never
        // inside a cell's range)
        size_t tb = sort_bytes;
        if ((e = hipcub::DeviceRadixSort::SortPairs(sort_tmp, tb, keys_tmp, keys_out, idx_tmp, idx_out, (int)n, 0, 32,
                                                    s)) != hipSuccess)
            return e;
        switch (kw) {
            case 3: k_xlist_records<3><<<blocks, 256, 0, s>>>(keys_out, idx_out, n, base, vals, nz, rec); break;
            case 5: k_xlist_records<5><<<blocks, 256, 0, s>>>(keys_out, idx_out, n, base, vals, nz, rec); break;
            case 7: k_xlist_records<7><<<blocks, 256, 0, s>>>(keys_out, idx_out, n, base, vals, nz, rec); break;
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    k_xlist_offsets<<<(unsigned)((ncell + 1 + 255) / 256), 256, 0, s>>>(keys_out, n, ncell, off, off0);
    return hipGetLastError();
}

}  // namespace cgs
