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
// hash cell and intersect it with the half-spaces n.(x - p) <= |n|^2/2,
// n = q - p, of the other points q, visiting hash cells shell by shell
// around p's cell until the farthest cell vertex is within half the distance
// that any unvisited point can have.  A shell's points are taken in a fixed
// order (hash cells in lexicographic (dx, dy, dz) order over the shell's
// surface, points of a hash cell by ascending index): the shell's hash cells
// in groups of 64, each group's points in batches of 64.  Every point of a
// batch is first tested against the cell as it is at the start of the batch
// (a point whose bisector leaves every vertex inside cannot cut any later,
// smaller cell), and the points that cut it are then clipped one after
// another, nearest first (the device kernel takes a group with one hash cell
// per lane and tests a batch with one point per lane).  A cell that still
// has a face of the box is taken as unbounded and gets volume 0, as points
// on the convex hull do in volume_of_finite_dual_cell (functions.h:148-150).  Clipped rule: this also
// zeroes the cells of interior points whose Voronoi region reaches more than
// one hash cell beyond the point cloud (near-degenerate boundary cells whose
// CGAL volume exceeds the whole k-space).  Hull rule (the reference's): such
// cells are built again from a box kWideBox times wider than the point cloud
// over every hash cell, and only the cells still keeping a box face get 0.
// Volume = sum over faces of area * |q - p| / 6 (functions.h:156-158), face
// areas as triangle fans from the first vertex (functions.h:117-124).
//
// A clip keeps the cell in a canonical form that a wave of 64 lanes can
// produce in parallel: the new vertex pool lists the kept vertices in pool
// order, then one new vertex per distinct cut edge in ascending order of the
// edge key (inner vertex, outer vertex); faces keep their order, rings are
// stored back to back (CSR); the new face's points are the pool's on-plane
// and new vertices in pool order without repeated values, sorted by angle
// around their mean (a 64-slot butterfly sum) through each point's rank, or
// by insertion when rounding makes the angular order inconsistent.  Cells
// are built with small fixed limits (faces, face points, pool vertices, ring
// entries) and, when they outgrow them, again with large ones; a cell's bits
// do not depend on the limits it fits in.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace vor {

constexpr int RMAX = 10;    // hash shells visited at most (first box)
constexpr int BATCH = 64;   // points tested against one state of the cell

// Fixed capacities of a cell: faces, points of the new face, pool vertices
// and ring entries (all faces' rings back to back), and the pool index type.
template <int F, int V, int P, int R, class I>
struct Limits {
    static constexpr int FMAX = F;
    static constexpr int VMAX = V;
    static constexpr int PMAX = P;
    static constexpr int NRING = R;
    using Idx = I;
};
using SmallLimits = Limits<64, 48, 128, 384, uint8_t>;
using LargeLimits = Limits<384, 256, 2048, 4864, uint16_t>;

struct V3 {
    double x, y, z;
};

// Face f's ring is ring[cur][fo[cur][f] .. fo[cur][f + 1]); clipping writes
// the new pool, faces and rings into the other halves and flips `cur`.
template <class LM>
struct Cell {
    using Idx = typename LM::Idx;
    int nf;
    int nvert;
    int cur;
    int overflow;
    double rmax2;
    int fid[2][LM::FMAX];
    double fh[2][LM::FMAX];
    int fo[2][LM::FMAX + 1];
    Idx ring[2][LM::NRING];
    V3 v[2][LM::PMAX];
};

// Working arrays of one clip.
template <class LM>
struct ClipScratch {
    using Idx = typename LM::Idx;
    double s[LM::PMAX];
    Idx map[LM::PMAX];
    uint8_t kon[LM::PMAX];
    uint32_t key[LM::NRING];
    uint8_t first[LM::NRING];
    int rank[LM::NRING];
    Idx u[LM::VMAX];
    double pa[LM::VMAX], pb[LM::VMAX];
    int rk[LM::VMAX];
    Idx sorted[LM::VMAX];
};

inline double dot3(const V3& a, const V3& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 sub3(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross3(const V3& a, const V3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

template <class LM>
void init_box(Cell<LM>& c, const double* lo, const double* hi) {
    // the bounding box [lo, hi] of all points (plus margin), relative to the
    // cell's point; corner k has coordinates (bit 2, bit 1, bit 0) of k; every
    // ring runs counter-clockwise seen from outside
    const unsigned char q[6][4] = {{4, 6, 7, 5}, {0, 1, 3, 2}, {2, 3, 7, 6},
                                   {0, 4, 5, 1}, {1, 5, 7, 3}, {0, 2, 6, 4}};
    c.nf = 6;
    c.nvert = 8;
    c.cur = 0;
    c.overflow = 0;
    for (int k = 0; k < 8; k++)
        c.v[0][k] = {(k & 4) ? hi[0] : lo[0], (k & 2) ? hi[1] : lo[1], (k & 1) ? hi[2] : lo[2]};
    for (int f = 0; f < 6; f++) {
        c.fid[0][f] = -1;
        c.fh[0][f] = 0.0;
        c.fo[0][f] = 4 * f;
        for (int k = 0; k < 4; k++) c.ring[0][4 * f + k] = q[f][k];
    }
    c.fo[0][6] = 24;
    double r2 = 0.0;
    for (int k = 0; k < 8; k++) {
        const double d = dot3(c.v[0][k], c.v[0][k]);
        if (d > r2) r2 = d;
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

// Sum of x[0..n) as 64 slots (slot l adds x[l], x[l + 64], ... in order,
// from 0.0) joined by the butterfly slot[l] + slot[l ^ d], d = 32 .. 1.
inline double slot_sum(const double* x, int n) {
    double a[64], b[64];
    for (int l = 0; l < 64; l++) a[l] = 0.0;
    for (int c = 0; c < n; c++) a[c & 63] = a[c & 63] + x[c];
    for (int d = 32; d >= 1; d >>= 1) {
        for (int l = 0; l < 64; l++) b[l] = a[l] + a[l ^ d];
        for (int l = 0; l < 64; l++) a[l] = b[l];
    }
    return a[0];
}

// True when the plane n.x = |n|^2/2 leaves a pool vertex strictly outside.
template <class LM>
bool cuts(const Cell<LM>& c, const V3& n) {
    const double cc = 0.5 * dot3(n, n);
    const double tol = 1e-10 * cc;
    const V3* ov = c.v[c.cur];
    for (int i = 0; i < c.nvert; i++)
        if (dot3(n, ov[i]) - cc > tol) return true;
    return false;
}

// Intersect the cell with the half-space n.x <= |n|^2/2.  Vertices with
// signed distance s within +-tol of the plane are kept and belong to the
// cap; an edge whose ends lie strictly on both sides gets one new vertex,
// interpolated from its inner end.
template <class LM>
void clip(Cell<LM>& c, ClipScratch<LM>& w, const V3& n, int nid) {
    using Idx = typename LM::Idx;
    constexpr int FMAX = LM::FMAX, VMAX = LM::VMAX, PMAX = LM::PMAX, NRING = LM::NRING;
    constexpr int SH = 8 * (int)sizeof(Idx);
    constexpr uint32_t LOW = (1u << SH) - 1u;
    const double cc = 0.5 * dot3(n, n);
    const double tol = 1e-10 * cc;
    const int cur = c.cur, nxt = cur ^ 1;
    const V3* ov = c.v[cur];
    V3* nvx = c.v[nxt];
    double* s = w.s;
    bool any_out = false;
    for (int i = 0; i < c.nvert; i++) {
        s[i] = dot3(n, ov[i]) - cc;
        if (s[i] > tol) any_out = true;
    }
    if (!any_out) return;
    // kept vertices in pool order
    int nk = 0;
    for (int i = 0; i < c.nvert; i++)
        if (s[i] <= tol) {
            w.map[i] = (Idx)nk;
            nvx[nk] = ov[i];
            w.kon[nk] = s[i] >= -tol;
            nk++;
        }
    // cut edge occurrences in face / ring order
    auto is_cut = [&](int a, int b) {
        return (s[a] < -tol && s[b] > tol) || (s[a] > tol && s[b] < -tol);
    };
    int nocc = 0;
    for (int f = 0; f < c.nf; f++) {
        const int b0 = c.fo[cur][f], len = c.fo[cur][f + 1] - b0;
        for (int k = 0; k < len; k++) {
            const int a = c.ring[cur][b0 + k], b = c.ring[cur][b0 + (k + 1) % len];
            if (is_cut(a, b)) {
                if (nocc == NRING) { c.overflow = 1; return; }
                const int iu = s[a] < 0.0 ? a : b;
                const int iw = s[a] < 0.0 ? b : a;
                w.key[nocc++] = ((uint32_t)iu << SH) | (uint32_t)iw;
            }
        }
    }
    // one new vertex per distinct key, numbered in ascending key order
    int nnew = 0;
    for (int t = 0; t < nocc; t++) {
        bool first = true;
        for (int u = 0; u < t; u++)
            if (w.key[u] == w.key[t]) { first = false; break; }
        w.first[t] = first;
        nnew += first;
    }
    if (nk + nnew > PMAX) { c.overflow = 1; return; }
    for (int t = 0; t < nocc; t++) {
        int r = 0;
        for (int u = 0; u < nocc; u++)
            if (w.first[u] && w.key[u] < w.key[t]) r++;
        w.rank[t] = nk + r;
    }
    for (int t = 0; t < nocc; t++) {
        if (!w.first[t]) continue;
        const int iu = (int)(w.key[t] >> SH), iw = (int)(w.key[t] & LOW);
        const V3 u = ov[iu];
        const V3 q = ov[iw];
        const double su = s[iu];
        const double sw = s[iw];
        const double tt = su / (su - sw);
        nvx[w.rank[t]] = {u.x + tt * (q.x - u.x), u.y + tt * (q.y - u.y), u.z + tt * (q.z - u.z)};
    }
    const int nvert2 = nk + nnew;
    // faces: kept vertices and cut points in ring order; fewer than 3 drops the face
    int nf2 = 0, nr2 = 0, o = 0;
    for (int f = 0; f < c.nf; f++) {
        const int b0 = c.fo[cur][f], len = c.fo[cur][f + 1] - b0;
        int no = 0;
        for (int k = 0; k < len; k++) {
            const int a = c.ring[cur][b0 + k], b = c.ring[cur][b0 + (k + 1) % len];
            no += s[a] <= tol;
            no += is_cut(a, b);
        }
        if (no < 3) {
            for (int k = 0; k < len; k++)
                o += is_cut(c.ring[cur][b0 + k], c.ring[cur][b0 + (k + 1) % len]);
            continue;
        }
        if (nr2 + no > NRING) { c.overflow = 1; return; }
        c.fid[nxt][nf2] = c.fid[cur][f];
        c.fh[nxt][nf2] = c.fh[cur][f];
        c.fo[nxt][nf2] = nr2;
        for (int k = 0; k < len; k++) {
            const int a = c.ring[cur][b0 + k], b = c.ring[cur][b0 + (k + 1) % len];
            if (s[a] <= tol) c.ring[nxt][nr2++] = w.map[a];
            if (is_cut(a, b)) c.ring[nxt][nr2++] = (Idx)w.rank[o++];
        }
        nf2++;
    }
    // the cap: on-plane kept and new vertices in pool order, repeated values dropped
    int nu = 0;
    for (int j = 0; j < nvert2; j++) {
        if (j < nk && !w.kon[j]) continue;
        bool dup = false;
        for (int t = 0; t < nu; t++)
            if (same3(nvx[w.u[t]], nvx[j])) { dup = true; break; }
        if (dup) continue;
        if (nu == VMAX) { c.overflow = 1; return; }
        w.u[nu++] = (Idx)j;
    }
    if (nu >= 3) {
        double cx[LM::VMAX], cy[LM::VMAX], cz[LM::VMAX];
        for (int t = 0; t < nu; t++) {
            cx[t] = nvx[w.u[t]].x;
            cy[t] = nvx[w.u[t]].y;
            cz[t] = nvx[w.u[t]].z;
        }
        const V3 m = {slot_sum(cx, nu) / nu, slot_sum(cy, nu) / nu, slot_sum(cz, nu) / nu};
        const V3 e1 = sub3(nvx[w.u[0]], m);
        const V3 e2 = cross3(n, e1);
        for (int t = 0; t < nu; t++) {
            const V3 d = sub3(nvx[w.u[t]], m);
            w.pa[t] = dot3(d, e1);
            w.pb[t] = dot3(d, e2);
        }
        // rank of each point: the points before it, then equal points of lower index
        bool perm = true;
        for (int t = 0; t < nu; t++) {
            int r = 0;
            for (int j = 0; j < nu; j++) {
                if (j == t) continue;
                const bool jt = ang_less(w.pa[j], w.pb[j], w.pa[t], w.pb[t]);
                const bool tj = ang_less(w.pa[t], w.pb[t], w.pa[j], w.pb[j]);
                r += jt || (j < t && !tj);
            }
            w.rk[t] = r;
        }
        for (int t = 0; t < nu && perm; t++)
            for (int j = 0; j < t; j++)
                if (w.rk[j] == w.rk[t]) { perm = false; break; }
        if (perm) {
            for (int t = 0; t < nu; t++) w.sorted[w.rk[t]] = w.u[t];
        } else {
            // inconsistent angular order (rounding): insertion sort in cap order
            for (int t = 0; t < nu; t++) w.sorted[t] = w.u[t];
            for (int i = 1; i < nu; i++) {
                const Idx tv = w.sorted[i];
                const double ta = w.pa[i], tb = w.pb[i];
                int j = i - 1;
                while (j >= 0 && ang_less(ta, tb, w.pa[j], w.pb[j])) {
                    w.sorted[j + 1] = w.sorted[j]; w.pa[j + 1] = w.pa[j]; w.pb[j + 1] = w.pb[j];
                    j--;
                }
                w.sorted[j + 1] = tv; w.pa[j + 1] = ta; w.pb[j + 1] = tb;
            }
        }
        if (nf2 == FMAX || nr2 + nu > NRING) { c.overflow = 1; return; }
        c.fid[nxt][nf2] = nid;
        c.fh[nxt][nf2] = std::sqrt(dot3(n, n));
        c.fo[nxt][nf2] = nr2;
        for (int t = 0; t < nu; t++) c.ring[nxt][nr2++] = w.sorted[t];
        nf2++;
    }
    c.fo[nxt][nf2] = nr2;
    c.cur = nxt;
    c.nvert = nvert2;
    c.nf = nf2;
    double r2 = 0.0;
    for (int j = 0; j < nvert2; j++) {
        const double d = dot3(nvx[j], nvx[j]);
        if (d > r2) r2 = d;
    }
    c.rmax2 = r2;
}

template <class LM>
bool has_box_face(const Cell<LM>& c) {
    for (int f = 0; f < c.nf; f++)
        if (c.fid[c.cur][f] < 0) return true;
    return false;
}

template <class LM>
double cell_volume(const Cell<LM>& c) {
    const int cur = c.cur;
    const V3* v = c.v[cur];
    double vol = 0.0;
    for (int f = 0; f < c.nf; f++) {
        if (c.fid[cur][f] < 0) return 0.0;
        const int b0 = c.fo[cur][f], len = c.fo[cur][f + 1] - b0;
        const V3 v0 = v[c.ring[cur][b0]];
        double area = 0.0;
        for (int k = 1; k < len - 1; k++) {
            const V3 cr = cross3(sub3(v[c.ring[cur][b0 + k]], v0), sub3(v[c.ring[cur][b0 + k + 1]], v0));
            area += std::sqrt(dot3(cr, cr)) / 2.0;
        }
        vol += area * c.fh[cur][f];
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

// Hash cells of shell r (Chebyshev distance r from the centre cell): their
// number, and the k-th in lexicographic (dx, dy, dz) order.
inline long shell_size(int r) {
    if (r == 0) return 1;
    const long m = 2L * r + 1;
    return 6 * m * m - 12 * m + 8;
}

inline void shell_cell(int r, long k, int* dx, int* dy, int* dz) {
    if (r == 0) { *dx = *dy = *dz = 0; return; }
    const long m = 2L * r + 1, face = m * m, ring = 4 * m - 4;
    if (k < face) { *dx = -r; *dy = (int)(k / m) - r; *dz = (int)(k % m) - r; return; }
    k -= face;
    if (k < (m - 2) * ring) {
        *dx = (int)(k / ring) - r + 1;
        long t = k % ring;
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

// One batch of up to 64 points j[0..nb): every point is tested against the
// cell as it is at the start of the batch, then the points that cut it are
// clipped nearest first (ascending |q - p|^2, ties in batch order).  Returns
// false when the cell outgrew the limits (c.overflow stays set unless a
// final pass takes a cell that still reaches the box as unbounded, which
// clears it).
template <class LM>
bool clip_batch(const double* p, long i, const V3& pi, const int* j, int nb, Cell<LM>& c, ClipScratch<LM>& w,
                bool final_pass) {
    V3 nn[BATCH];
    double d2[BATCH];
    int hit[BATCH];
    int nh = 0;
    const double rm = c.rmax2;
    for (int t = 0; t < nb; t++) {
        const V3 q = {p[3L * j[t]], p[3L * j[t] + 1], p[3L * j[t] + 2]};
        nn[t] = sub3(q, pi);
        d2[t] = dot3(nn[t], nn[t]);
        // duplicate points (d2 == 0) share their vertex and cut nothing
        if (j[t] != i && d2[t] != 0.0 && 0.25 * d2[t] < rm && cuts(c, nn[t])) hit[nh++] = t;
    }
    for (int a = 1; a < nh; a++) {   // stable insertion sort by distance
        const int x = hit[a];
        int b = a - 1;
        while (b >= 0 && d2[hit[b]] > d2[x]) { hit[b + 1] = hit[b]; b--; }
        hit[b + 1] = x;
    }
    for (int h = 0; h < nh; h++) {
        const int t = hit[h];
        const bool box = has_box_face(c);
        clip(c, w, nn[t], j[t]);
        if (c.overflow) {
            // a cell still reaching the bounding box that outgrows the limits
            // is taken as unbounded (0); a closed cell's overflow stays an error
            if (box && final_pass) c.overflow = 0;
            return false;
        }
    }
    return true;
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
    // shells beyond the farthest grid face hold no hash cells
    int rin = 0;
    for (int a = 0; a < 3; a++) {
        const int e = c3[a] > H.n[a] - 1 - c3[a] ? c3[a] : H.n[a] - 1 - c3[a];
        rin = e > rin ? e : rin;
    }
    if (rlim > rin) rlim = rin;
    std::vector<int> cand;
    for (int r = 0; r <= rlim; r++) {
        const long ns = shell_size(r);
        // the shell's hash cells in groups of 64, each group's points in batches of 64
        for (long h0 = 0; h0 < ns; h0 += BATCH) {
            cand.clear();
            for (long k = h0; k < ns && k < h0 + BATCH; k++) {
                int dx, dy, dz;
                shell_cell(r, k, &dx, &dy, &dz);
                const int x = c3[0] + dx, y = c3[1] + dy, z = c3[2] + dz;
                if (x < 0 || y < 0 || z < 0 || x >= H.n[0] || y >= H.n[1] || z >= H.n[2]) continue;
                const long cell = ((long)x * H.n[1] + y) * H.n[2] + z;
                for (int t = H.start[cell]; t < H.start[cell + 1]; t++) cand.push_back(H.idx[t]);
            }
            for (size_t b0 = 0; b0 < cand.size(); b0 += BATCH) {
                const size_t b1 = b0 + BATCH < cand.size() ? b0 + BATCH : cand.size();
                if (!clip_batch(p, i, pi, cand.data() + b0, (int)(b1 - b0), c, w, final_pass)) return 0.0;
            }
        }
        const double reach = (double)r * H.h;
        if (0.25 * reach * reach >= c.rmax2) break;
    }
    return cell_volume(c);
}

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
    if (np < 2) {
        for (long i = 0; i < np; i++) vol[i] = 0.0;
        return 0;
    }
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
