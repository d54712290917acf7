// cgs_voronoi.hip -- Voronoi density compensation (estimateDensityV3D,
// functions.h:161-258) by half-space clipping, one cell per wave (first pass).
//
// The reference takes each sample's Voronoi cell volume from a CGAL Delaunay
// triangulation (sum over Delaunay edges of dual-facet area * edge length / 6,
// 0 for points on the convex hull) and divides by the maximum.  Here every
// cell is built on its own: start from the bounding box of all points widened
// by one hash cell and intersect it with the bisector half-spaces of the other
// points, visiting a uniform hash grid shell by shell in a fixed order until
// no unvisited point can cut the cell (the farthest vertex is closer than half
// the distance of the next shell).  Clipped rule (per-bin weights): cells
// that keep a face of the box get volume 0 -- the hull points, and interior
// points whose Voronoi region reaches more than one hash cell beyond the
// point cloud.  Hull rule (the reference's): those cells are built again
// from a box wider than the point cloud by kVorWideBox times its extent,
// visiting every hash cell until the radius bound closes, and only the cells
// that still keep a box face (the unbounded ones) get 0.  The
// volume is the reference's face sum area * |q - p| / 6 with areas as
// triangle fans from the first vertex.  A cell is a pool of vertices and
// faces as rings of pool indices (about 12.7 KB with the first-pass limits,
// held in LDS).
//
// Built with -ffp-contract=off; every operation is a separately rounded
// IEEE double operation in a fixed order, so the weights are reproducible
// bit for bit by a CPU evaluation of the same statements.
#include "cgs_internal.h"
#include "cgs_kernels.h"

namespace cgs {
namespace vor {

constexpr int RMAX = 10;   // hash shells visited at most

// Fixed capacities of a cell: faces, vertices of one face, distinct vertices
// (a simple polyhedron with F faces has at most 2F - 4), and the index type of
// the vertex pool.  Every cell is built with the small limits first; a cell
// that outgrows them is built again with the large ones.  The arithmetic does
// not depend on the limits, so a cell that fits both gets the same bits.
template <int F, int V, int P, class I>
struct Limits {
    static constexpr int FMAX = F;
    static constexpr int VMAX = V;
    static constexpr int PMAX = P;
    using Idx = I;
};
using SmallLimits = Limits<64, 48, 128, uint8_t>;
using MediumLimits = Limits<128, 96, 254, uint8_t>;
using LargeLimits = Limits<384, 256, 2048, uint16_t>;

struct V3 {
    double x, y, z;
};

// Faces are rings of indices into a vertex pool; clipping writes the new
// pool into the other half of v[2][] and flips `cur`.
template <class LM>
struct Cell {
    using Idx = typename LM::Idx;
    int nf;
    int nvert;
    int cur;
    int overflow;
    double rmax2;
    int fid[LM::FMAX];
    double fh[LM::FMAX];
    uint16_t nv[LM::FMAX];
    Idx ring[LM::FMAX][LM::VMAX];
    V3 v[2][LM::PMAX];
};

// Working arrays of one clip (kept apart from the cell so that large-limit
// cells can live in a heap or device workspace).
template <class LM>
struct ClipScratch {
    using Idx = typename LM::Idx;
    double s[LM::PMAX];
    Idx map[LM::PMAX];
    uint32_t edge[LM::PMAX];
    Idx cap[LM::FMAX * 2];
    Idx out[LM::VMAX];
    Idx u[LM::VMAX];
    double pa[LM::VMAX], pb[LM::VMAX];
};

__device__ __forceinline__ double dot3(const V3& a, const V3& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 sub3(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 cross3(const V3& a, const V3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

template <class LM>
__device__ void init_box(Cell<LM>& c, const double* lo, const double* hi) {
    // the bounding box [lo, hi] of all points (plus margin), relative to the
    // cell's point; corner k has coordinates (bit 2, bit 1, bit 0) of k
    const unsigned char q[6][4] = {{4, 6, 7, 5}, {0, 1, 3, 2}, {2, 3, 7, 6},
                                   {0, 4, 5, 1}, {1, 5, 7, 3}, {0, 2, 6, 4}};
    c.nf = 6;
    c.nvert = 8;
    c.cur = 0;
    c.overflow = 0;
    for (int k = 0; k < 8; k++)
        c.v[0][k] = {(k & 4) ? hi[0] : lo[0], (k & 2) ? hi[1] : lo[1], (k & 1) ? hi[2] : lo[2]};
    double r2 = 0.0;
    for (int f = 0; f < 6; f++) {
        c.fid[f] = -1;
        c.fh[f] = 0.0;
        c.nv[f] = 4;
        for (int k = 0; k < 4; k++) {
            c.ring[f][k] = q[f][k];
            const double d = dot3(c.v[0][q[f][k]], c.v[0][q[f][k]]);
            if (d > r2) r2 = d;
        }
    }
    c.rmax2 = r2;
}

// Angular order of cap points around their centroid in the plane with normal n.
__device__ __forceinline__ bool ang_less(double a1, double b1, double a2, double b2) {
    const bool u1 = b1 > 0.0 || (b1 == 0.0 && a1 > 0.0);
    const bool u2 = b2 > 0.0 || (b2 == 0.0 && a2 > 0.0);
    if (u1 != u2) return u1;
    return a1 * b2 - a2 * b1 > 0.0;
}

__device__ __forceinline__ bool same3(const V3& a, const V3& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// Intersect the cell with the half-space n.x <= |n|^2/2.  Vertices with
// signed distance within +-tol of the plane are kept and belong to the cap;
// an edge whose ends lie strictly on both sides gets one new vertex,
// interpolated from its inner end, shared by the two faces of the edge.  The
// cap (the new face) is the set of distinct cap vertices, in first-seen order
// over the faces, sorted by angle around their mean.
template <class LM>
__device__ void clip(Cell<LM>& c, ClipScratch<LM>& w, const V3& n, int nid) {
    using Idx = typename LM::Idx;
    constexpr int FMAX = LM::FMAX, VMAX = LM::VMAX, PMAX = LM::PMAX;
    constexpr Idx NONE = (Idx)~(Idx)0;
    constexpr uint32_t NOEDGE = 0xffffffffu;
    constexpr int SH = 8 * (int)sizeof(Idx);
    const double cc = 0.5 * dot3(n, n);
    const double tol = 1e-10 * cc;
    const V3* __restrict__ ov = c.v[c.cur];
    V3* __restrict__ nvx = c.v[c.cur ^ 1];
    double* s = w.s;
    for (int i = 0; i < c.nvert; i++) s[i] = dot3(n, ov[i]) - cc;
    bool any_out = false;
    for (int f = 0; f < c.nf && !any_out; f++)
        for (int k = 0; k < c.nv[f]; k++)
            if (s[c.ring[f][k]] > tol) { any_out = true; break; }
    if (!any_out) return;
    // new-pool index of every kept old vertex (NONE = not yet placed), and for
    // each new vertex the old edge it was cut from (NOEDGE for kept vertices)
    Idx* map = w.map;
    uint32_t* edge = w.edge;
    for (int i = 0; i < c.nvert; i++) map[i] = NONE;
    int nn = 0;
    Idx* cap = w.cap;
    int ncap = 0;
    int nf = 0;
    for (int f = 0; f < c.nf; f++) {
        const int k0 = c.nv[f];
        Idx* out = w.out;
        int no = 0;
        for (int k = 0; k < k0; k++) {
            const int ia = c.ring[f][k];
            const int ib = c.ring[f][(k + 1) % k0];
            const double sa = s[ia];
            const double sb = s[ib];
            if (sa <= tol) {
                if (map[ia] == NONE) {
                    if (nn >= PMAX) { c.overflow = 1; return; }
                    nvx[nn] = ov[ia];
                    edge[nn] = NOEDGE;
                    map[ia] = (Idx)nn++;
                }
                if (no < VMAX) out[no++] = map[ia]; else c.overflow = 1;
                if (sa >= -tol) {
                    if (ncap < FMAX * 2) cap[ncap++] = map[ia]; else c.overflow = 1;
                }
            }
            if ((sa < -tol && sb > tol) || (sa > tol && sb < -tol)) {
                const int iu = sa < 0.0 ? ia : ib;
                const int iw = sa < 0.0 ? ib : ia;
                const uint32_t key = ((uint32_t)iu << SH) | (uint32_t)iw;
                int ix = -1;
                for (int t = 0; t < nn; t++)
                    if (edge[t] == key) { ix = t; break; }
                if (ix < 0) {
                    if (nn >= PMAX) { c.overflow = 1; return; }
                    const V3 u = ov[iu];
                    const V3 w = ov[iw];
                    const double su = s[iu];
                    const double sw = s[iw];
                    const double t = su / (su - sw);
                    nvx[nn] = {u.x + t * (w.x - u.x), u.y + t * (w.y - u.y), u.z + t * (w.z - u.z)};
                    edge[nn] = key;
                    ix = nn++;
                }
                if (no < VMAX) out[no++] = (Idx)ix; else c.overflow = 1;
                if (ncap < FMAX * 2) cap[ncap++] = (Idx)ix; else c.overflow = 1;
            }
        }
        if (no >= 3) {
            c.fid[nf] = c.fid[f];
            c.fh[nf] = c.fh[f];
            c.nv[nf] = (uint16_t)no;
            for (int k = 0; k < no; k++) c.ring[nf][k] = out[k];
            nf++;
        }
    }
    c.cur ^= 1;
    c.nvert = nn;
    // distinct cap points (by value)
    Idx* u = w.u;
    int nu = 0;
    for (int i = 0; i < ncap; i++) {
        bool dup = false;
        for (int j = 0; j < nu; j++)
            if (same3(nvx[cap[i]], nvx[u[j]])) { dup = true; break; }
        if (!dup) {
            if (nu == VMAX) { c.overflow = 1; return; }   // cap face too large
            u[nu++] = cap[i];
        }
    }
    if (nu >= 3) {
        V3 m = {0.0, 0.0, 0.0};
        for (int i = 0; i < nu; i++) { m.x += nvx[u[i]].x; m.y += nvx[u[i]].y; m.z += nvx[u[i]].z; }
        m.x /= nu; m.y /= nu; m.z /= nu;
        const V3 e1 = sub3(nvx[u[0]], m);
        const V3 e2 = cross3(n, e1);
        double* pa = w.pa;
        double* pb = w.pb;
        for (int i = 0; i < nu; i++) {
            const V3 d = sub3(nvx[u[i]], m);
            pa[i] = dot3(d, e1);
            pb[i] = dot3(d, e2);
        }
        for (int i = 1; i < nu; i++) {   // insertion sort by angle
            const Idx tv = u[i];
            const double ta = pa[i], tb = pb[i];
            int j = i - 1;
            while (j >= 0 && ang_less(ta, tb, pa[j], pb[j])) {
                u[j + 1] = u[j]; pa[j + 1] = pa[j]; pb[j + 1] = pb[j];
                j--;
            }
            u[j + 1] = tv; pa[j + 1] = ta; pb[j + 1] = tb;
        }
        if (nf < FMAX) {
            c.fid[nf] = nid;
            c.fh[nf] = sqrt(dot3(n, n));
            c.nv[nf] = (uint16_t)nu;
            for (int k = 0; k < nu; k++) c.ring[nf][k] = u[k];
            nf++;
        } else {
            c.overflow = 1;
        }
    }
    c.nf = nf;
    double r2 = 0.0;
    for (int f = 0; f < nf; f++)
        for (int k = 0; k < c.nv[f]; k++) {
            const double d = dot3(nvx[c.ring[f][k]], nvx[c.ring[f][k]]);
            if (d > r2) r2 = d;
        }
    c.rmax2 = r2;
}

template <class LM>
__device__ bool has_box_face(const Cell<LM>& c) {
    for (int f = 0; f < c.nf; f++)
        if (c.fid[f] < 0) return true;
    return false;
}

template <class LM>
__device__ double cell_volume(const Cell<LM>& c) {
    const V3* v = c.v[c.cur];
    double vol = 0.0;
    for (int f = 0; f < c.nf; f++) {
        if (c.fid[f] < 0) return 0.0;
        const V3 v0 = v[c.ring[f][0]];
        double area = 0.0;
        for (int k = 1; k < c.nv[f] - 1; k++) {
            const V3 cr = cross3(sub3(v[c.ring[f][k]], v0), sub3(v[c.ring[f][k + 1]], v0));
            area += sqrt(dot3(cr, cr)) / 2.0;
        }
        vol += area * c.fh[f];
    }
    return vol / 6.0;
}

// Volume of point i's cell.  With the final (large) limits a cell that still
// has a box face when it outgrows them is taken as unbounded (0, overflow
// cleared) and a closed cell's overflow stays set; with the first (small)
// limits any overflow is left set so that the cell is built again.
// wide = false: the box of the points widened by one hash cell, at most RMAX
// shells; wide = true (hull rule, second stage): the wide box, every shell.
template <class LM>
__device__ double point_volume(const double* __restrict__ p, long long i, const VorHash& H, Cell<LM>& c,
                               ClipScratch<LM>& w, bool final_pass, bool wide) {
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
    init_box(c, blo, bhi);
    int rlim = RMAX;
    if (wide)
        for (int a = 0; a < 3; a++) rlim = H.n[a] > rlim ? H.n[a] : rlim;
    for (int r = 0; r <= rlim; r++) {
        for (int dx = -r; dx <= r; dx++)
            for (int dy = -r; dy <= r; dy++)
                for (int dz = -r; dz <= r; dz++) {
                    const int ax = dx < 0 ? -dx : dx, ay = dy < 0 ? -dy : dy, az = dz < 0 ? -dz : dz;
                    const int cheb = ax > ay ? (ax > az ? ax : az) : (ay > az ? ay : az);
                    if (cheb != r) continue;
                    const int x = c3[0] + dx, y = c3[1] + dy, z = c3[2] + dz;
                    if (x < 0 || y < 0 || z < 0 || x >= H.n[0] || y >= H.n[1] || z >= H.n[2]) continue;
                    const long long cell = ((long long)x * H.n[1] + y) * H.n[2] + z;
                    for (int t = H.start[cell]; t < H.start[cell + 1]; t++) {
                        const int j = H.idx[t];
                        if (j == i) continue;
                        const V3 q = {p[3LL * j], p[3LL * j + 1], p[3LL * j + 2]};
                        const V3 n = sub3(q, pi);
                        const double d2 = dot3(n, n);
                        if (d2 == 0.0) continue;
                        if (0.25 * d2 >= c.rmax2) continue;
                        const bool box = has_box_face(c);
                        clip(c, w, n, j);
                        if (c.overflow) {
                            // a cell still reaching the bounding box that outgrows the limits
                            // is taken as unbounded (0); a closed cell's overflow stays an error
                            if (box && final_pass) c.overflow = 0;
                            return 0.0;
                        }
                    }
                }
        const double reach = (double)r * H.h;
        if (0.25 * reach * reach >= c.rmax2) break;
    }
    return cell_volume(c);
}

// First pass: one cell per workgroup of one wave, built by its first lane in
// LDS (about 12.7 KB with the small limits, so about twelve cells are built
// at once on a CU); the other lanes idle.  The clipping is a long chain of
// dependent, data-dependent accesses to the cell, which LDS serves far faster
// than private memory.  Grid-stride over the points; cells that outgrow the
// limits are listed for the second pass.
__global__ __launch_bounds__(64) void k_voronoi(const double* __restrict__ p, long long n,
                                               const int* __restrict__ list, bool wide, VorHash H,
                                               double* __restrict__ vol, int* __restrict__ redo,
                                               int* __restrict__ nredo) {
    __shared__ Cell<SmallLimits> c;
    __shared__ ClipScratch<SmallLimits> w;
    if (threadIdx.x != 0) return;
    for (long long k = blockIdx.x; k < n; k += gridDim.x) {
        const long long i = list ? (long long)list[k] : k;
        vol[i] = point_volume(p, i, H, c, w, false, wide);
        if (c.overflow) redo[atomicAdd(nredo, 1)] = (int)i;
    }
}

// Middle pass over the listed points with the medium limits, again one cell
// per wave in LDS (about 32 KB); cells that outgrow these too are listed for
// the last pass.
__global__ __launch_bounds__(64) void k_voronoi_mid(const double* __restrict__ p, bool wide, VorHash H,
                                                   const int* __restrict__ redo, int nredo,
                                                   double* __restrict__ vol, int* __restrict__ redo2,
                                                   int* __restrict__ nredo2) {
    __shared__ Cell<MediumLimits> c;
    __shared__ ClipScratch<MediumLimits> w;
    if (threadIdx.x != 0) return;
    for (long long k = blockIdx.x; k < nredo; k += gridDim.x) {
        const long long i = redo[k];
        vol[i] = point_volume(p, i, H, c, w, false, wide);
        if (c.overflow) redo2[atomicAdd(nredo2, 1)] = (int)i;
    }
}

// Last pass over the listed points with the large limits; each thread's
// cell and working arrays live in a device workspace (too large for private
// memory).  Closed cells that outgrow these limits as well are counted.
__global__ __launch_bounds__(64) void k_voronoi_large(const double* __restrict__ p, bool wide, VorHash H,
                                                     const int* __restrict__ redo, int nredo,
                                                     Cell<LargeLimits>* __restrict__ cells,
                                                     ClipScratch<LargeLimits>* __restrict__ scratch,
                                                     double* __restrict__ vol, int* __restrict__ overflow) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    Cell<LargeLimits>& c = cells[t];
    ClipScratch<LargeLimits>& w = scratch[t];
    for (long long k = t; k < nredo; k += (long long)gridDim.x * blockDim.x) {
        const long long i = redo[k];
        vol[i] = point_volume(p, i, H, c, w, true, wide);
        if (c.overflow) atomicAdd(overflow, 1);
    }
}

}  // namespace vor

hipError_t launch_voronoi(const double* pts, long long n, const int* list, bool wide, const VorHash& H, double* vol,
                          int* redo, int* nredo, hipStream_t s) {
    // one wave per cell in flight; 8192 workgroups cover every CU several
    // times over and grid-stride over the rest
    const long long blocks = std::min<long long>(n, 8192);
    if (blocks == 0) return hipSuccess;
    vor::k_voronoi<<<(unsigned)blocks, 64, 0, s>>>(pts, n, list, wide, H, vol, redo, nredo);
    return hipGetLastError();
}

hipError_t launch_voronoi_mid(const double* pts, bool wide, const VorHash& H, const int* redo, int nredo, double* vol,
                              int* redo2, int* nredo2, hipStream_t s) {
    if (nredo == 0) return hipSuccess;
    vor::k_voronoi_mid<<<(unsigned)std::min(nredo, 8192), 64, 0, s>>>(pts, wide, H, redo, nredo, vol, redo2,
                                                                     nredo2);
    return hipGetLastError();
}

void voronoi_large_sizes(size_t* cell, size_t* scratch) {
    *cell = sizeof(vor::Cell<vor::LargeLimits>);
    *scratch = sizeof(vor::ClipScratch<vor::LargeLimits>);
}

hipError_t launch_voronoi_large(const double* pts, bool wide, const VorHash& H, const int* redo, int nredo,
                                int nthreads, void* cells, void* scratch, double* vol, int* overflow, hipStream_t s) {
    if (nredo == 0) return hipSuccess;
    vor::k_voronoi_large<<<(unsigned)(nthreads / 64), 64, 0, s>>>(
        pts, wide, H, redo, nredo, (vor::Cell<vor::LargeLimits>*)cells, (vor::ClipScratch<vor::LargeLimits>*)scratch, vol,
        overflow);
    return hipGetLastError();
}

}  // namespace cgs
