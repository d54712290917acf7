// cgs_voronoi_cell.h -- one Voronoi cell built by a whole wave (64 lanes).
//
// Device code of the Voronoi density compensation (estimateDensityV3D,
// functions.h:161-258; cell volume functions.h:128-159).  A cell is built by
// half-space clipping: start from a box around the point, intersect it with
// the bisector half-spaces of the other points, and sum area * |q - p| / 6
// over the faces.  The 64 lanes of a wave share one cell:
//
//  * candidate points come 64 at a time, one per lane; every lane tests its
//    point's bisector against the cell (a bounding-sphere and bounding-box
//    bound first, then every vertex by broadcast reads); the points that cut
//    are then clipped one after another, nearest first;
//  * a clip runs over the vertex pool (one vertex per lane), over the ring
//    entries of all faces (one entry per lane) and over the faces, with
//    wave prefix sums placing the results, so the cell stays in a canonical
//    form: kept vertices in pool order, then one new vertex per distinct cut
//    edge in ascending edge-key order; faces in order with their rings back
//    to back; the new face's points sorted by angle through per-lane ranks.
//
// The same statement runs sequentially on the CPU (oracle/voronoi_clip.cpp);
// every floating-point value is formed by the same separately rounded
// operations in the same order (the file is compiled with
// -ffp-contract=off), so the volumes agree bit for bit.  The bounding-box
// test is exact to the vertex test it precedes: with lo <= v <= hi per axis,
// every rounded product and sum of the bound is >= the same rounded term of
// n.v, so a bound inside the half-space means every vertex is.
//
// Every __syncthreads() and wave vote below sits in control flow that is
// uniform across the wave: loop bounds and branch conditions are values that
// all lanes hold identically (counts from ballots and scans, the cell's
// uniform state).  Kernels using this file run one wave per workgroup.
#pragma once

#include <cstdint>
#include <type_traits>

namespace cgs {
namespace vor {

constexpr int RMAX = 10;    // hash shells visited at most (first box)
constexpr int WAVE = 64;

// Capacities of a cell: faces, points of the new face, pool vertices, ring
// entries (all faces' rings back to back) and cut-edge occurrences of one
// clip, and the pool index type.
template <int F, int V, int P, int R, int O, class I>
struct Limits {
    static constexpr int FMAX = F;
    static constexpr int VMAX = V;
    static constexpr int PMAX = P;
    static constexpr int NRING = R;
    static constexpr int OCC = O;
    using Idx = I;
};
using SmallLimits = Limits<64, 48, 128, 384, 128, uint8_t>;
using MediumLimits = Limits<128, 96, 254, 768, 256, uint8_t>;
using LargeLimits = Limits<384, 256, 2048, 4864, 4864, uint16_t>;

struct V3 {
    double x, y, z;
};

// Cell and clip scratch of one wave.  Face f's ring is
// ring[cur][fo[cur][f] .. fo[cur][f + 1]) and efc[cur][e] is the face of ring
// entry e; a clip writes the new pool, faces and rings into the other halves.
template <class LM>
struct WCell {
    using Idx = typename LM::Idx;
    using FIdx = typename std::conditional<(LM::FMAX < 256), uint8_t, uint16_t>::type;
    static constexpr int FMAX = LM::FMAX, VMAX = LM::VMAX, PMAX = LM::PMAX, NRING = LM::NRING, OCC = LM::OCC;
    V3 v[2][PMAX];
    double fh[2][FMAX];
    int fid[2][FMAX];
    uint16_t fo[2][FMAX + 1];
    Idx ring[2][NRING];
    FIdx efc[2][NRING];
    double s[PMAX];              // signed distances; face terms of the volume
    uint16_t eo[NRING + 1];      // outputs of the ring entries before entry e
    uint16_t eoc[NRING];         // occurrence number of a cut entry
    uint8_t ef[NRING];           // bit 0: entry's vertex kept, bit 1: entry's edge cut
    uint32_t key[OCC];           // cut-edge keys in ring order
    uint16_t rank[OCC];          // their new pool indices
    uint8_t first[OCC];
    int16_t nfi[FMAX];           // new index of a surviving face, -1 when dropped
    int16_t nbase[FMAX];         // its new ring offset less eo at its first entry
    Idx map[PMAX];
    Idx cand[PMAX];
    uint8_t kon[PMAX];
    double pa[VMAX], pb[VMAX];
    uint16_t rk[VMAX];
    Idx u[VMAX], sorted[VMAX];
    int hst[WAVE], hoff[WAVE];   // a group of hash cells: first point, offset in the group's list
    double hd2[WAVE];            // a batch's cutting points: distances in lane order
    uint8_t hord[WAVE];          // and their lanes nearest first
};

// The wave's copy of the cell's scalar state (identical in every lane).
// rmax2 and the bounding box lo/hi are refreshed lazily (dirty after a clip).
struct WState {
    int nf, nvert, nring, cur, nbox, overflow, dirty;
    double rmax2;
    double lo[3], hi[3];
};

__device__ __forceinline__ double dot3(const V3& a, const V3& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 sub3(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 cross3(const V3& a, const V3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ bool same3(const V3& a, const V3& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }

__device__ __forceinline__ bool ang_less(double a1, double b1, double a2, double b2) {
    const bool u1 = b1 > 0.0 || (b1 == 0.0 && a1 > 0.0);
    const bool u2 = b2 > 0.0 || (b2 == 0.0 && a2 > 0.0);
    if (u1 != u2) return u1;
    return a1 * b2 - a2 * b1 > 0.0;
}

__device__ __forceinline__ bool is_cut(double sa, double sb, double tol) {
    return (sa < -tol && sb > tol) || (sa > tol && sb < -tol);
}

__device__ __forceinline__ uint64_t lanes_below(int lane) { return (1ull << lane) - 1ull; }

// Exclusive prefix sum over the wave; *total gets the sum of all lanes.
__device__ __forceinline__ int wave_excl_scan(int v, int lane, int* total) {
    int x = v;
    for (int d = 1; d < WAVE; d <<= 1) {
        const int y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    *total = __shfl(x, WAVE - 1);
    return x - v;
}

// Butterfly sum slot[l] + slot[l ^ d], d = 32 .. 1: every lane ends with the
// same bits.
__device__ __forceinline__ double wave_sum(double v) {
    for (int d = WAVE / 2; d >= 1; d >>= 1) v = v + __shfl_xor(v, d);
    return v;
}

// Hash cells of shell r (Chebyshev distance r from the centre cell): their
// number, and the k-th in lexicographic (dx, dy, dz) order.
__device__ __forceinline__ long long shell_size(int r) {
    if (r == 0) return 1;
    const long long m = 2LL * r + 1;
    return 6 * m * m - 12 * m + 8;
}

__device__ __forceinline__ void shell_cell(int r, long long k, int* dx, int* dy, int* dz) {
    if (r == 0) { *dx = *dy = *dz = 0; return; }
    const long long m = 2LL * r + 1, face = m * m, ring = 4 * m - 4;
    if (k < face) { *dx = -r; *dy = (int)(k / m) - r; *dz = (int)(k % m) - r; return; }
    k -= face;
    if (k < (m - 2) * ring) {
        *dx = (int)(k / ring) - r + 1;
        long long t = k % ring;
        if (t < m) { *dy = -r; *dz = (int)t - r; return; }
        t -= m;
        if (t < 2 * (m - 2)) { *dy = (int)(t / 2) - r + 1; *dz = (t & 1) ? r : -r; return; }
        t -= 2 * (m - 2);
        *dy = r; *dz = (int)t - r;
        return;
    }
    k -= (m - 2) * ring;
    *dx = r; *dy = (int)(k / m) - r; *dz = (int)(k % m) - r;
}

// The box [lo, hi] relative to the cell's point; corner k has coordinates
// (bit 2, bit 1, bit 0) of k; every ring runs counter-clockwise seen from
// outside.
template <class LM>
__device__ void winit_box(WCell<LM>& C, WState& S, const double* lo, const double* hi, int lane) {
    const unsigned char q[6][4] = {{4, 6, 7, 5}, {0, 1, 3, 2}, {2, 3, 7, 6},
                                   {0, 4, 5, 1}, {1, 5, 7, 3}, {0, 2, 6, 4}};
    if (lane < 8)
        C.v[0][lane] = {(lane & 4) ? hi[0] : lo[0], (lane & 2) ? hi[1] : lo[1], (lane & 1) ? hi[2] : lo[2]};
    if (lane < 6) {
        C.fid[0][lane] = -1;
        C.fh[0][lane] = 0.0;
        C.fo[0][lane] = (uint16_t)(4 * lane);
    }
    if (lane == 6) C.fo[0][6] = 24;
    if (lane < 24) {
        C.ring[0][lane] = q[lane >> 2][lane & 3];
        C.efc[0][lane] = (typename WCell<LM>::FIdx)(lane >> 2);
    }
    double r2 = 0.0;
    for (int k = 0; k < 8; k++) {
        const V3 c = {(k & 4) ? hi[0] : lo[0], (k & 2) ? hi[1] : lo[1], (k & 1) ? hi[2] : lo[2]};
        const double d = dot3(c, c);
        if (d > r2) r2 = d;
    }
    S.nf = 6;
    S.nvert = 8;
    S.nring = 24;
    S.cur = 0;
    S.nbox = 6;
    S.overflow = 0;
    S.dirty = 0;
    S.rmax2 = r2;
    for (int a = 0; a < 3; a++) {
        S.lo[a] = lo[a];
        S.hi[a] = hi[a];
    }
    __syncthreads();
}

// rmax2 = largest |v|^2 over the pool, lo/hi = its bounding box (exact
// minima and maxima, so independent of the reduction order).
template <class LM>
__device__ void wrefresh(const WCell<LM>& C, WState& S, int lane) {
    const V3* v = C.v[S.cur];
    double r2 = 0.0, lx = v[0].x, ly = v[0].y, lz = v[0].z, hx = lx, hy = ly, hz = lz;
    for (int j = lane; j < S.nvert; j += WAVE) {
        const V3 x = v[j];
        r2 = dmax(r2, dot3(x, x));
        lx = dmin(lx, x.x); ly = dmin(ly, x.y); lz = dmin(lz, x.z);
        hx = dmax(hx, x.x); hy = dmax(hy, x.y); hz = dmax(hz, x.z);
    }
    for (int d = WAVE / 2; d >= 1; d >>= 1) {
        r2 = dmax(r2, __shfl_xor(r2, d));
        lx = dmin(lx, __shfl_xor(lx, d));
        ly = dmin(ly, __shfl_xor(ly, d));
        lz = dmin(lz, __shfl_xor(lz, d));
        hx = dmax(hx, __shfl_xor(hx, d));
        hy = dmax(hy, __shfl_xor(hy, d));
        hz = dmax(hz, __shfl_xor(hz, d));
    }
    S.rmax2 = r2;
    S.lo[0] = lx; S.lo[1] = ly; S.lo[2] = lz;
    S.hi[0] = hx; S.hi[1] = hy; S.hi[2] = hz;
    S.dirty = 0;
}

// Intersect the cell with the half-space n.x <= |n|^2/2 (nid: the point the
// new face belongs to).  Vertices with signed distance s within +-tol of the
// plane are kept and belong to the new face; an edge whose ends lie strictly
// on both sides gets one new vertex, interpolated from its inner end.  Sets
// S.overflow when the result outgrows the limits (the cell is then void).
template <class LM>
__device__ void wclip(WCell<LM>& C, WState& S, const V3 n, int nid, int lane) {
    using Idx = typename LM::Idx;
    using FIdx = typename WCell<LM>::FIdx;
    constexpr int FMAX = LM::FMAX, VMAX = LM::VMAX, PMAX = LM::PMAX, NRING = LM::NRING, OCC = LM::OCC;
    constexpr int SH = 8 * (int)sizeof(Idx);
    constexpr uint32_t LOW = (1u << SH) - 1u;
    const double cc = 0.5 * dot3(n, n);
    const double tol = 1e-10 * cc;
    const int cur = S.cur, nxt = cur ^ 1;
    const V3* ov = C.v[cur];
    V3* nvx = C.v[nxt];
    const int nvert = S.nvert, nf = S.nf, nr = S.nring;

    // signed distances (lane i % 64 owns vertex i)
    bool out = false;
    for (int i = lane; i < nvert; i += WAVE) {
        const double si = dot3(n, ov[i]) - cc;
        C.s[i] = si;
        out = out || si > tol;
    }
    if (!__any(out)) return;
    // kept vertices in pool order
    int nk = 0;
    for (int i0 = 0; i0 < nvert; i0 += WAVE) {
        const int i = i0 + lane;
        const bool keep = i < nvert && C.s[i] <= tol;
        const uint64_t m = __ballot(keep);
        if (keep) {
            const int pos = nk + __popcll(m & lanes_below(lane));
            C.map[i] = (Idx)pos;
            nvx[pos] = ov[i];
            C.kon[pos] = C.s[i] >= -tol;
        }
        nk += __popcll(m);
    }
    __syncthreads();
    // ring entries (lane e % 64 owns entry e): the entry's vertex kept, its
    // edge to the next entry of the face cut; outputs before each entry; cut
    // keys in ring order
    int nocc = 0, nout = 0;
    for (int e0 = 0; e0 < nr; e0 += WAVE) {
        const int e = e0 + lane;
        int kept = 0, cut = 0;
        uint32_t key = 0;
        if (e < nr) {
            const int f = C.efc[cur][e];
            const int nx = (e + 1 == C.fo[cur][f + 1]) ? C.fo[cur][f] : e + 1;
            const int a = C.ring[cur][e], b = C.ring[cur][nx];
            const double sa = C.s[a], sb = C.s[b];
            kept = sa <= tol;
            cut = is_cut(sa, sb, tol);
            const int iu = sa < 0.0 ? a : b;
            const int iw = sa < 0.0 ? b : a;
            key = ((uint32_t)iu << SH) | (uint32_t)iw;
        }
        const uint64_t mc = __ballot(cut);
        const int oc = nocc + __popcll(mc & lanes_below(lane));
        int tot;
        const int ex = wave_excl_scan(kept + cut, lane, &tot);
        if (e < nr) {
            C.eo[e] = (uint16_t)(nout + ex);
            C.ef[e] = (uint8_t)(kept | (cut << 1));
            if (cut) {
                C.eoc[e] = (uint16_t)oc;
                if (oc < OCC) C.key[oc] = key;
            }
        }
        nocc += __popcll(mc);
        nout += tot;
    }
    if (lane == 0) C.eo[nr] = (uint16_t)nout;
    if (nocc > OCC) { S.overflow = 1; return; }
    __syncthreads();
    // one new vertex per distinct key, numbered in ascending key order
    int nnew = 0;
    for (int t0 = 0; t0 < nocc; t0 += WAVE) {
        const int t = t0 + lane;
        bool first = false;
        if (t < nocc) {
            const uint32_t k = C.key[t];
            first = true;
            for (int u = 0; u < t; u++)
                if (C.key[u] == k) { first = false; break; }
            C.first[t] = first;
        }
        nnew += __popcll(__ballot(first));
    }
    if (nk + nnew > PMAX) { S.overflow = 1; return; }
    __syncthreads();
    for (int t = lane; t < nocc; t += WAVE) {
        const uint32_t k = C.key[t];
        int r = 0;
        for (int u = 0; u < nocc; u++) r += (C.first[u] && C.key[u] < k) ? 1 : 0;
        C.rank[t] = (uint16_t)(nk + r);
        if (C.first[t]) {
            const int iu = (int)(k >> SH), iw = (int)(k & LOW);
            const V3 u = ov[iu];
            const V3 q = ov[iw];
            const double su = C.s[iu];
            const double sw = C.s[iw];
            const double tt = su / (su - sw);
            nvx[nk + r] = {u.x + tt * (q.x - u.x), u.y + tt * (q.y - u.y), u.z + tt * (q.z - u.z)};
        }
    }
    const int nvert2 = nk + nnew;
    // faces (lane f % 64 owns face f) keeping 3 or more points, in order
    int nf2 = 0, nr2 = 0, nbox = 0;
    for (int f0 = 0; f0 < nf; f0 += WAVE) {
        const int f = f0 + lane;
        int no = 0, b0 = 0;
        if (f < nf) {
            b0 = C.fo[cur][f];
            no = C.eo[C.fo[cur][f + 1]] - C.eo[b0];
        }
        const bool keepf = no >= 3;
        const uint64_t m = __ballot(keepf);
        int tot;
        const int ex = wave_excl_scan(keepf ? no : 0, lane, &tot);
        if (f < nf) {
            if (keepf) {
                const int nfi = nf2 + __popcll(m & lanes_below(lane));
                C.nfi[f] = (int16_t)nfi;
                C.nbase[f] = (int16_t)(nr2 + ex - C.eo[b0]);
                C.fid[nxt][nfi] = C.fid[cur][f];
                C.fh[nxt][nfi] = C.fh[cur][f];
                C.fo[nxt][nfi] = (uint16_t)(nr2 + ex);
            } else {
                C.nfi[f] = -1;
            }
        }
        nbox += __popcll(__ballot(keepf && C.fid[cur][f] < 0));
        nf2 += __popcll(m);
        nr2 += tot;
    }
    if (nr2 > NRING) { S.overflow = 1; return; }
    __syncthreads();
    // every entry of a surviving face writes its kept vertex and its cut point
    for (int e = lane; e < nr; e += WAVE) {
        const int f = C.efc[cur][e];
        const int nfi = C.nfi[f];
        if (nfi >= 0) {
            const int fl = C.ef[e];
            int pos = C.nbase[f] + C.eo[e];
            if (fl & 1) {
                C.ring[nxt][pos] = C.map[C.ring[cur][e]];
                C.efc[nxt][pos] = (FIdx)nfi;
                pos++;
            }
            if (fl & 2) {
                C.ring[nxt][pos] = (Idx)C.rank[C.eoc[e]];
                C.efc[nxt][pos] = (FIdx)nfi;
            }
        }
    }
    // the new face: on-plane kept and new vertices in pool order, repeated values dropped
    int ncand = 0;
    for (int j0 = 0; j0 < nvert2; j0 += WAVE) {
        const int j = j0 + lane;
        const bool c = j < nvert2 && (j >= nk || C.kon[j]);
        const uint64_t m = __ballot(c);
        if (c) C.cand[ncand + __popcll(m & lanes_below(lane))] = (Idx)j;
        ncand += __popcll(m);
    }
    __syncthreads();
    int nu = 0;
    for (int t0 = 0; t0 < ncand; t0 += WAVE) {
        const int t = t0 + lane;
        bool keepc = false;
        if (t < ncand) {
            const V3 x = nvx[C.cand[t]];
            keepc = true;
            for (int u = 0; u < t; u++)
                if (same3(nvx[C.cand[u]], x)) { keepc = false; break; }
        }
        const uint64_t m = __ballot(keepc);
        const int pos = nu + __popcll(m & lanes_below(lane));
        if (keepc && pos < VMAX) C.u[pos] = C.cand[t];
        nu += __popcll(m);
    }
    if (nu > VMAX) { S.overflow = 1; return; }
    __syncthreads();
    if (nu >= 3) {
        double px = 0.0, py = 0.0, pz = 0.0;
        for (int t = lane; t < nu; t += WAVE) {
            const V3 x = nvx[C.u[t]];
            px = px + x.x;
            py = py + x.y;
            pz = pz + x.z;
        }
        const V3 m = {wave_sum(px) / nu, wave_sum(py) / nu, wave_sum(pz) / nu};
        const V3 e1 = sub3(nvx[C.u[0]], m);
        const V3 e2 = cross3(n, e1);
        for (int t = lane; t < nu; t += WAVE) {
            const V3 d = sub3(nvx[C.u[t]], m);
            C.pa[t] = dot3(d, e1);
            C.pb[t] = dot3(d, e2);
        }
        __syncthreads();
        // rank of each point: the points before it, then equal points of lower index
        for (int t = lane; t < nu; t += WAVE) {
            const double ta = C.pa[t], tb = C.pb[t];
            int r = 0;
            for (int j = 0; j < nu; j++) {
                if (j == t) continue;
                const bool jt = ang_less(C.pa[j], C.pb[j], ta, tb);
                const bool tj = ang_less(ta, tb, C.pa[j], C.pb[j]);
                r += (jt || (j < t && !tj)) ? 1 : 0;
            }
            C.rk[t] = (uint16_t)r;
        }
        __syncthreads();
        bool clash = false;
        for (int t = lane; t < nu; t += WAVE)
            for (int j = 0; j < nu; j++)
                if (j != t && C.rk[j] == C.rk[t]) { clash = true; break; }
        if (!__any(clash)) {
            for (int t = lane; t < nu; t += WAVE) C.sorted[C.rk[t]] = C.u[t];
        } else if (lane == 0) {
            // inconsistent angular order (rounding): insertion sort in cap order
            for (int t = 0; t < nu; t++) C.sorted[t] = C.u[t];
            for (int i = 1; i < nu; i++) {
                const Idx tv = C.sorted[i];
                const double ta = C.pa[i], tb = C.pb[i];
                int j = i - 1;
                while (j >= 0 && ang_less(ta, tb, C.pa[j], C.pb[j])) {
                    C.sorted[j + 1] = C.sorted[j];
                    C.pa[j + 1] = C.pa[j];
                    C.pb[j + 1] = C.pb[j];
                    j--;
                }
                C.sorted[j + 1] = tv;
                C.pa[j + 1] = ta;
                C.pb[j + 1] = tb;
            }
        }
        __syncthreads();
        if (nf2 == FMAX || nr2 + nu > NRING) { S.overflow = 1; return; }
        if (lane == 0) {
            C.fid[nxt][nf2] = nid;
            C.fh[nxt][nf2] = sqrt(dot3(n, n));
            C.fo[nxt][nf2] = (uint16_t)nr2;
        }
        for (int t = lane; t < nu; t += WAVE) {
            C.ring[nxt][nr2 + t] = C.sorted[t];
            C.efc[nxt][nr2 + t] = (FIdx)nf2;
        }
        nf2++;
        nr2 += nu;
    }
    if (lane == 0) C.fo[nxt][nf2] = (uint16_t)nr2;
    S.cur = nxt;
    S.nvert = nvert2;
    S.nf = nf2;
    S.nring = nr2;
    S.nbox = nbox;
    S.dirty = 1;
    __syncthreads();
}

// Volume of the finished cell: 0 while it keeps a box face, else the face
// sum area * |q - p| / 6 with fan areas (lane f % 64 forms face f's term,
// lane 0 adds them in face order).
template <class LM>
__device__ double wcell_volume(WCell<LM>& C, const WState& S, int lane) {
    if (S.nbox > 0) return 0.0;
    const int cur = S.cur;
    const V3* v = C.v[cur];
    for (int f = lane; f < S.nf; f += WAVE) {
        const int b0 = C.fo[cur][f], len = C.fo[cur][f + 1] - b0;
        const V3 v0 = v[C.ring[cur][b0]];
        double area = 0.0;
        for (int k = 1; k < len - 1; k++) {
            const V3 cr = cross3(sub3(v[C.ring[cur][b0 + k]], v0), sub3(v[C.ring[cur][b0 + k + 1]], v0));
            area += sqrt(dot3(cr, cr)) / 2.0;
        }
        C.s[f] = area * C.fh[cur][f];
    }
    __syncthreads();
    double vol = 0.0;
    if (lane == 0)
        for (int f = 0; f < S.nf; f++) vol += C.s[f];
    vol = __shfl(vol, 0);
    __syncthreads();
    return vol / 6.0;
}

// Volume of point i's cell (see the header).  Hash: the uniform grid over the
// points (lo, h, n, start, idx; boxes blo/bhi and wlo/whi), hash cell c
// holding idx[start[c] .. start[c + 1]) in ascending point index.  wide =
// false: the points' box widened by one hash cell, at most RMAX shells; wide
// = true (hull rule, second stage): the box widened by kVorWideBox times the
// points' extent, every shell.  On return S.overflow tells whether the cell
// outgrew the limits; with final_pass a cell that still reached the box when
// it did is taken as unbounded (0, overflow cleared).
template <class LM, class Hash>
__device__ double wpoint_volume(const double* __restrict__ p, long long i, const Hash& H, WCell<LM>& C,
                                WState& S, bool final_pass, bool wide, int lane) {
    const V3 pi = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    int c3[3];
    for (int a = 0; a < 3; a++) {
        int k = (int)((p[3 * i + a] - H.lo[a]) / H.h);
        c3[a] = k < 0 ? 0 : (k >= H.n[a] ? H.n[a] - 1 : k);
    }
    double blo[3], bhi[3];
    for (int a = 0; a < 3; a++) {
        blo[a] = (wide ? H.wlo[a] : H.blo[a]) - p[3 * i + a];
        bhi[a] = (wide ? H.whi[a] : H.bhi[a]) - p[3 * i + a];
    }
    winit_box(C, S, blo, bhi, lane);
    int rlim = RMAX;
    if (wide)
        for (int a = 0; a < 3; a++) rlim = H.n[a] > rlim ? H.n[a] : rlim;
    // shells beyond the farthest grid face hold no hash cells
    int rin = 0;
    for (int a = 0; a < 3; a++) {
        const int e = c3[a] > H.n[a] - 1 - c3[a] ? c3[a] : H.n[a] - 1 - c3[a];
        rin = e > rin ? e : rin;
    }
    if (rlim > rin) rlim = rin;
    for (int r = 0; r <= rlim; r++) {
        const long long ns = shell_size(r);
        // the shell's hash cells in groups of 64 (one per lane), each group's
        // points in batches of 64 (one per lane)
        for (long long h0 = 0; h0 < ns; h0 += WAVE) {
            const long long k = h0 + lane;
            int cnt = 0, st = 0;
            if (k < ns) {
                int dx, dy, dz;
                shell_cell(r, k, &dx, &dy, &dz);
                const int x = c3[0] + dx, y = c3[1] + dy, z = c3[2] + dz;
                if (x >= 0 && y >= 0 && z >= 0 && x < H.n[0] && y < H.n[1] && z < H.n[2]) {
                    const long long cell = ((long long)x * H.n[1] + y) * H.n[2] + z;
                    st = H.start[cell];
                    cnt = H.start[cell + 1] - st;
                }
            }
            int tot;
            const int ex = wave_excl_scan(cnt, lane, &tot);
            if (tot == 0) continue;
            C.hst[lane] = st;
            C.hoff[lane] = ex;
            __syncthreads();
            for (int q0 = 0; q0 < tot; q0 += WAVE) {
                if (S.dirty) wrefresh(C, S, lane);
                const int t = q0 + lane;
                int j = -1;
                V3 nrm = {0.0, 0.0, 0.0};
                double d2 = 0.0;
                bool hit = false;
                if (t < tot) {
                    // the group's last hash cell starting at or before t
                    int lo = 0, hi = WAVE - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (C.hoff[mid] <= t) lo = mid; else hi = mid - 1;
                    }
                    j = H.idx[C.hst[lo] + (t - C.hoff[lo])];
                    const V3 q = {p[3LL * j], p[3LL * j + 1], p[3LL * j + 2]};
                    nrm = sub3(q, pi);
                    d2 = dot3(nrm, nrm);
                    // duplicate points (d2 == 0) share their vertex and cut nothing
                    if (j != i && d2 != 0.0 && 0.25 * d2 < S.rmax2) {
                        const double cc = 0.5 * d2;
                        const double tol = 1e-10 * cc;
                        const double bound = (dmax(nrm.x * S.lo[0], nrm.x * S.hi[0]) +
                                              dmax(nrm.y * S.lo[1], nrm.y * S.hi[1])) +
                                             dmax(nrm.z * S.lo[2], nrm.z * S.hi[2]);
                        if (bound - cc > tol) {
                            const V3* vv = C.v[S.cur];
                            for (int v = 0; v < S.nvert; v++)
                                if (dot3(nrm, vv[v]) - cc > tol) { hit = true; break; }
                        }
                    }
                }
                const uint64_t mask = __ballot(hit);
                const int nh = __popcll(mask);
                if (nh == 0) continue;
                if (nh > 1) {
                    // nearest first, ties in batch order
                    const int hp = __popcll(mask & lanes_below(lane));
                    if (hit) C.hd2[hp] = d2;
                    __syncthreads();
                    if (hit) {
                        int r = 0;
                        for (int u = 0; u < nh; u++) {
                            const double du = C.hd2[u];
                            r += (du < d2 || (du == d2 && u < hp)) ? 1 : 0;
                        }
                        C.hord[r] = (uint8_t)lane;
                    }
                    __syncthreads();
                }
                for (int o = 0; o < nh; o++) {
                    const int b = nh == 1 ? __ffsll((long long)mask) - 1 : C.hord[o];
                    const V3 nb = {__shfl(nrm.x, b), __shfl(nrm.y, b), __shfl(nrm.z, b)};
                    const int jb = __shfl(j, b);
                    const bool box = S.nbox > 0;
                    wclip(C, S, nb, jb, lane);
                    if (S.overflow) {
                        // a cell still reaching the bounding box that outgrows the limits
                        // is taken as unbounded (0); a closed cell's overflow stays an error
                        if (box && final_pass) S.overflow = 0;
                        __syncthreads();
                        return 0.0;
                    }
                }
            }
            __syncthreads();
        }
        if (S.dirty) wrefresh(C, S, lane);
        const double reach = (double)r * H.h;
        if (0.25 * reach * reach >= S.rmax2) break;
    }
    return wcell_volume(C, S, lane);
}

}  // namespace vor
}  // namespace cgs
