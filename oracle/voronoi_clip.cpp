// voronoi_clip.cpp -- Voronoi cell volumes by half-space clipping, CPU
// statement of the algorithm that the device kernel cgs_voronoi.hip runs.
//
// TEST INFRASTRUCTURE ONLY (part of the oracle library).  The reference
// computes the same volumes from a CGAL Delaunay triangulation
// (functions.h:75-258); that restatement lives in oracle/voronoi_ref.py
// (Qhull).  This file states the build's own clipping algorithm with the
// exact IEEE operation sequence of the device kernel, so the device weights
// can be checked bit for bit; voronoi_ref.py pins the geometry.
//
// Cell of point p: start from the bounding box of all points widened by one
// hash cell and intersect it with
// the half-spaces n.(x - p) <= |n|^2/2, n = q - p, of the other points q,
// visiting hash cells shell by shell around p's cell (in a fixed order) until
// the farthest cell vertex is within half the distance that any unvisited
// point can have.  A cell that still has a face of the box is taken as
// unbounded and gets volume 0, as points on the convex hull do in
// volume_of_finite_dual_cell (functions.h:148-150).  Clipped rule: this also
// zeroes the cells of interior points whose Voronoi region reaches more than
// one hash cell beyond the point cloud (near-degenerate boundary cells whose
// CGAL volume exceeds the whole k-space).  Hull rule (the reference's): such
// cells are built again from a box kWideBox times wider than the point cloud
// over every hash cell, and only the cells still keeping a box face get 0.
// Volume = sum over faces of area *
// |q - p| / 6 (functions.h:156-158), face areas as triangle fans from the
// first vertex (functions.h:117-124).  Cells are built with small fixed
// limits (faces, vertices per face, distinct vertices) and, when they outgrow
// them, again with large ones; a cell's bits do not depend on the limits it
// fits in.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

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

inline double dot3(const V3& a, const V3& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 sub3(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross3(const V3& a, const V3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

template <class LM>
void init_box(Cell<LM>& c, const double* lo, const double* hi) {
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
inline bool ang_less(double a1, double b1, double a2, double b2) {
    const bool u1 = b1 > 0.0 || (b1 == 0.0 && a1 > 0.0);
    const bool u2 = b2 > 0.0 || (b2 == 0.0 && a2 > 0.0);
    if (u1 != u2) return u1;
    return a1 * b2 - a2 * b1 > 0.0;
}

inline bool same3(const V3& a, const V3& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// Intersect the cell with the half-space n.x <= |n|^2/2.  Vertices with
// signed distance within +-tol of the plane are kept and belong to the cap;
// an edge whose ends lie strictly on both sides gets one new vertex,
// interpolated from its inner end, shared by the two faces of the edge.  The
// cap (the new face) is the set of distinct cap vertices, in first-seen order
// over the faces, sorted by angle around their mean.
template <class LM>
void clip(Cell<LM>& c, ClipScratch<LM>& w, const V3& n, int nid) {
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
            c.fh[nf] = std::sqrt(dot3(n, n));
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
bool has_box_face(const Cell<LM>& c) {
    for (int f = 0; f < c.nf; f++)
        if (c.fid[f] < 0) return true;
    return false;
}

template <class LM>
double cell_volume(const Cell<LM>& c) {
    const V3* v = c.v[c.cur];
    double vol = 0.0;
    for (int f = 0; f < c.nf; f++) {
        if (c.fid[f] < 0) return 0.0;
        const V3 v0 = v[c.ring[f][0]];
        double area = 0.0;
        for (int k = 1; k < c.nv[f] - 1; k++) {
            const V3 cr = cross3(sub3(v[c.ring[f][k]], v0), sub3(v[c.ring[f][k + 1]], v0));
            area += std::sqrt(dot3(cr, cr)) / 2.0;
        }
        vol += area * c.fh[f];
    }
    return vol / 6.0;
}

// Hull rule: a cell that still has a face of the first box is built again
// from a box wider than the point cloud by kWideBox times its largest extent
// on every side, visiting every hash cell until the radius bound closes; one
// that still keeps a box face then is unbounded (a hull point).
constexpr double kWideBox = 65536.0;

struct Hash {
    double lo[3];
    double blo[3], bhi[3];   // bounding box of the points widened by one hash cell
    double wlo[3], whi[3];   // the same box widened by kWideBox times the largest extent
    double h;
    int n[3];
    std::vector<int> start, idx;
};

Hash build_hash(const double* p, long np) {
    Hash H;
    double hi[3];
    for (int a = 0; a < 3; a++) { H.lo[a] = p[a]; hi[a] = p[a]; }
    for (long i = 0; i < np; i++)
        for (int a = 0; a < 3; a++) {
            H.lo[a] = std::fmin(H.lo[a], p[3 * i + a]);
            hi[a] = std::fmax(hi[a], p[3 * i + a]);
        }
    double vol = 1.0;
    for (int a = 0; a < 3; a++) vol *= std::fmax(hi[a] - H.lo[a], 1e-12);
    H.h = std::cbrt(vol * 6.0 / (double)np);
    for (int a = 0; a < 3; a++) H.n[a] = (int)((hi[a] - H.lo[a]) / H.h) + 1;
    for (int a = 0; a < 3; a++) { H.blo[a] = H.lo[a] - H.h; H.bhi[a] = hi[a] + H.h; }
    double ext = 0.0;
    for (int a = 0; a < 3; a++) ext = std::fmax(ext, hi[a] - H.lo[a]);
    const double wide = kWideBox * ext;
    for (int a = 0; a < 3; a++) { H.wlo[a] = H.lo[a] - wide; H.whi[a] = hi[a] + wide; }
    const long nc = (long)H.n[0] * H.n[1] * H.n[2];
    std::vector<int> cnt(nc + 1, 0), key(np);
    for (long i = 0; i < np; i++) {
        int c3[3];
        for (int a = 0; a < 3; a++) {
            int k = (int)((p[3 * i + a] - H.lo[a]) / H.h);
            c3[a] = k < 0 ? 0 : (k >= H.n[a] ? H.n[a] - 1 : k);
        }
        key[i] = (c3[0] * H.n[1] + c3[1]) * H.n[2] + c3[2];
        cnt[key[i] + 1]++;
    }
    for (long c = 0; c < nc; c++) cnt[c + 1] += cnt[c];
    H.start = cnt;
    H.idx.assign(np, 0);
    std::vector<int> cur(cnt.begin(), cnt.end() - 1);
    for (long i = 0; i < np; i++) H.idx[cur[key[i]]++] = (int)i;   // ascending index per cell
    return H;
}

// Volume of point i's cell.  With the final (large) limits a cell that still
// has a box face when it outgrows them is taken as unbounded (0, overflow
// cleared) and a closed cell's overflow stays set; with the first (small)
// limits any overflow is left set so that the cell is built again.
// wide = false: the box of the points widened by one hash cell, at most RMAX
// shells; wide = true (hull rule, second stage): the box widened by kWideBox
// times the points' extent, every shell.
template <class LM>
double point_volume(const double* p, long i, const Hash& H, Cell<LM>& c, ClipScratch<LM>& w, bool final_pass,
                    bool wide = false) {
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
                    const long cell = ((long)x * H.n[1] + y) * H.n[2] + z;
                    for (int t = H.start[cell]; t < H.start[cell + 1]; t++) {
                        const int j = H.idx[t];
                        if (j == i) continue;
                        const V3 q = {p[3 * j], p[3 * j + 1], p[3 * j + 2]};
                        const V3 n = sub3(q, pi);
                        const double d2 = dot3(n, n);
                        if (d2 == 0.0) continue;                 // duplicate point: same vertex
                        if (0.25 * d2 >= c.rmax2) continue;
                        const bool box = has_box_face(c);
                        clip(c, w, n, j);
                        if (c.overflow) {
                            // a cell still reaching the bounding cube that outgrows the limits
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

}  // namespace vor

namespace vor {

// Volumes of the listed points (all when list is null) with the small limits,
// cells that outgrow them again with the large ones; returns the closed cells
// that outgrew the large limits.
long volumes_pass(const double* pts, long n, const long* list, const Hash& H, bool wide, double* vol, int nthreads) {
    long over = 0;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : omp_get_max_threads()) reduction(+ : over)
    {
        auto* cs = new Cell<SmallLimits>;
        auto* ws = new ClipScratch<SmallLimits>;
        Cell<LargeLimits>* cl = nullptr;
        ClipScratch<LargeLimits>* wl = nullptr;
#pragma omp for schedule(dynamic, 16)
        for (long k = 0; k < n; k++) {
            const long i = list ? list[k] : k;
            vol[i] = point_volume(pts, i, H, *cs, *ws, false, wide);
            if (cs->overflow) {   // second pass with the large limits
                if (!cl) {
                    cl = new Cell<LargeLimits>;
                    wl = new ClipScratch<LargeLimits>;
                }
                vol[i] = point_volume(pts, i, H, *cl, *wl, true, wide);
                over += cl->overflow;
            }
        }
        delete cs;
        delete ws;
        delete cl;
        delete wl;
    }
    return over;
}

}  // namespace vor

extern "C" {

// Unnormalised cell volumes of np points (np x 3 doubles); returns the number
// of cells that overflowed the fixed face/vertex limits.
//   rule 0 (hull, the reference's functions.h:148-150): 0 only for cells
//     that are unbounded, i.e. points on the convex hull.  Cells that keep a
//     face of the first box (the points' bounding box widened by one hash
//     cell) are built again from the wide box; those that still keep a box
//     face are the unbounded ones.
//   rule 1 (clipped): 0 for every cell that keeps a face of the first box
//     (hull points and interior points whose cell reaches beyond it).
long orc_voronoi_volumes_rule(const double* pts, long np, int rule, double* vol, int nthreads) {
    vor::Hash H = vor::build_hash(pts, np);
    long over = vor::volumes_pass(pts, np, nullptr, H, false, vol, nthreads);
    if (rule == 0) {
        std::vector<long> open;
        for (long i = 0; i < np; i++)
            if (vol[i] == 0.0) open.push_back(i);
        over += vor::volumes_pass(pts, (long)open.size(), open.data(), H, true, vol, nthreads);
    }
    return over;
}

long orc_voronoi_volumes(const double* pts, long np, double* vol, int nthreads) {
    return orc_voronoi_volumes_rule(pts, np, 1, vol, nthreads);
}

// The tail of estimateDensityV3D (functions.h:212-246) on given volumes:
// float cell volumes; with average, per-class float sums over i % sX (the
// reference sums them in a parallel loop with a data race; here in index
// order), their double total, and every volume replaced by
// total^-1 * sum[min(i % sX, (int)((sX - 1) * cutoff))]; finally the volumes
// divided by their maximum (double reciprocal, float result).
void orc_density_from_volumes(const double* vol, long np, int average, int sx, double cutoff, float* dcf) {
    for (long i = 0; i < np; i++) dcf[i] = (float)vol[i];
    if (average && sx > 0) {
        std::vector<float> sum(sx, 0.0f);
        for (long i = 0; i < np; i++) sum[i % sx] += dcf[i];
        double tot = 0.0;
        for (int k = 0; k < sx; k++) tot += sum[k];
        tot = 1.0 / tot;
        const int cut = (int)((sx - 1) * cutoff);
        for (long i = 0; i < np; i++) {
            const int k = (int)(i % sx) < cut ? (int)(i % sx) : cut;
            dcf[i] = (float)(tot * sum[k]);
        }
    }
    double mx = 0.0;
    for (long i = 0; i < np; i++)
        if (dcf[i] > mx) mx = dcf[i];
    const double inv = 1.0 / mx;
    for (long i = 0; i < np; i++) dcf[i] = (float)((double)dcf[i] * inv);
}

// estimateDensityV3D: cell volumes by the given rule, then the tail above.
long orc_voronoi_density_rule(const double* pts, long np, int rule, int average, int sx, double cutoff, float* dcf,
                              int nthreads) {
    std::vector<double> vol(np);
    const long over = orc_voronoi_volumes_rule(pts, np, rule, vol.data(), nthreads);
    orc_density_from_volumes(vol.data(), np, average, sx, cutoff, dcf);
    return over;
}

long orc_voronoi_density(const double* pts, long np, int average, int sx, double cutoff, float* dcf, int nthreads) {
    return orc_voronoi_density_rule(pts, np, 1, average, sx, cutoff, dcf, nthreads);
}

}  // extern "C"
