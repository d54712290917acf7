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
// Cell of point p: start from the cube |x - p|_inf <= R0 and intersect it with
// the half-spaces n.(x - p) <= |n|^2/2, n = q - p, of the other points q,
// visiting hash cells shell by shell around p's cell (in a fixed order) until
// the farthest cell vertex is within half the distance that any unvisited
// point can have.  A cell that still has a cube face is unbounded (p on the
// convex hull) and gets volume 0, as in volume_of_finite_dual_cell
// (functions.h:148-150).  Volume = sum over faces of area * |q - p| / 6
// (functions.h:156-158), face areas as triangle fans from the first vertex
// (functions.h:117-124).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace vor {

constexpr int FMAX = 48;
constexpr int VMAX = 32;
constexpr int RMAX = 10;   // hash shells visited at most

struct V3 {
    double x, y, z;
};

struct Cell {
    int nf;
    int fid[FMAX];
    double fh[FMAX];
    int nv[FMAX];
    V3 v[FMAX][VMAX];
    double rmax2;
    int overflow;
};

inline double dot3(const V3& a, const V3& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 sub3(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross3(const V3& a, const V3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

void init_box(Cell& c, double r0) {
    static const int q[6][4][3] = {
        {{1, -1, -1}, {1, 1, -1}, {1, 1, 1}, {1, -1, 1}},     {{-1, -1, -1}, {-1, -1, 1}, {-1, 1, 1}, {-1, 1, -1}},
        {{-1, 1, -1}, {-1, 1, 1}, {1, 1, 1}, {1, 1, -1}},     {{-1, -1, -1}, {1, -1, -1}, {1, -1, 1}, {-1, -1, 1}},
        {{-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1}},     {{-1, -1, -1}, {-1, 1, -1}, {1, 1, -1}, {1, -1, -1}}};
    c.nf = 6;
    c.overflow = 0;
    for (int f = 0; f < 6; f++) {
        c.fid[f] = -1;
        c.fh[f] = 0.0;
        c.nv[f] = 4;
        for (int k = 0; k < 4; k++) c.v[f][k] = {q[f][k][0] * r0, q[f][k][1] * r0, q[f][k][2] * r0};
    }
    c.rmax2 = 3.0 * r0 * r0;
}

// Angular order of cap points around their centroid in the plane with normal n.
inline bool ang_less(double a1, double b1, double a2, double b2) {
    const bool u1 = b1 > 0.0 || (b1 == 0.0 && a1 > 0.0);
    const bool u2 = b2 > 0.0 || (b2 == 0.0 && a2 > 0.0);
    if (u1 != u2) return u1;
    return a1 * b2 - a2 * b1 > 0.0;
}

void clip(Cell& c, const V3& n, int nid) {
    const double cc = 0.5 * dot3(n, n);
    const double tol = 1e-10 * cc;
    bool any_out = false;
    for (int f = 0; f < c.nf && !any_out; f++)
        for (int k = 0; k < c.nv[f]; k++)
            if (dot3(n, c.v[f][k]) - cc > tol) { any_out = true; break; }
    if (!any_out) return;
    V3 cap[FMAX * 2];
    int ncap = 0;
    int nf = 0;
    for (int f = 0; f < c.nf; f++) {
        const int k0 = c.nv[f];
        V3 out[VMAX];
        int no = 0;
        for (int k = 0; k < k0; k++) {
            const V3 a = c.v[f][k];
            const V3 b = c.v[f][(k + 1) % k0];
            const double sa = dot3(n, a) - cc;
            const double sb = dot3(n, b) - cc;
            if (sa <= tol) {
                if (no < VMAX) out[no++] = a; else c.overflow = 1;
                if (sa >= -tol) {
                    if (ncap < FMAX * 2) cap[ncap++] = a; else c.overflow = 1;
                }
            }
            if ((sa < -tol && sb > tol) || (sa > tol && sb < -tol)) {
                const V3 u = sa < 0.0 ? a : b;
                const V3 w = sa < 0.0 ? b : a;
                const double su = sa < 0.0 ? sa : sb;
                const double sw = sa < 0.0 ? sb : sa;
                const double t = su / (su - sw);
                const V3 x = {u.x + t * (w.x - u.x), u.y + t * (w.y - u.y), u.z + t * (w.z - u.z)};
                if (no < VMAX) out[no++] = x; else c.overflow = 1;
                if (ncap < FMAX * 2) cap[ncap++] = x; else c.overflow = 1;
            }
        }
        if (no >= 3) {
            c.fid[nf] = c.fid[f];
            c.fh[nf] = c.fh[f];
            c.nv[nf] = no;
            for (int k = 0; k < no; k++) c.v[nf][k] = out[k];
            nf++;
        }
    }
    // distinct cap points
    V3 u[FMAX * 2];
    int nu = 0;
    for (int i = 0; i < ncap; i++) {
        bool dup = false;
        for (int j = 0; j < nu; j++)
            if (cap[i].x == u[j].x && cap[i].y == u[j].y && cap[i].z == u[j].z) { dup = true; break; }
        if (!dup) u[nu++] = cap[i];
    }
    if (nu >= 3) {
        V3 m = {0.0, 0.0, 0.0};
        for (int i = 0; i < nu; i++) { m.x += u[i].x; m.y += u[i].y; m.z += u[i].z; }
        m.x /= nu; m.y /= nu; m.z /= nu;
        const V3 e1 = sub3(u[0], m);
        const V3 e2 = cross3(n, e1);
        double pa[FMAX * 2], pb[FMAX * 2];
        for (int i = 0; i < nu; i++) {
            const V3 d = sub3(u[i], m);
            pa[i] = dot3(d, e1);
            pb[i] = dot3(d, e2);
        }
        for (int i = 1; i < nu; i++) {   // insertion sort by angle
            const V3 tv = u[i];
            const double ta = pa[i], tb = pb[i];
            int j = i - 1;
            while (j >= 0 && ang_less(ta, tb, pa[j], pb[j])) {
                u[j + 1] = u[j]; pa[j + 1] = pa[j]; pb[j + 1] = pb[j];
                j--;
            }
            u[j + 1] = tv; pa[j + 1] = ta; pb[j + 1] = tb;
        }
        if (nf < FMAX && nu <= VMAX) {
            c.fid[nf] = nid;
            c.fh[nf] = std::sqrt(dot3(n, n));
            c.nv[nf] = nu;
            for (int k = 0; k < nu; k++) c.v[nf][k] = u[k];
            nf++;
        } else {
            c.overflow = 1;
        }
    }
    c.nf = nf;
    double r2 = 0.0;
    for (int f = 0; f < nf; f++)
        for (int k = 0; k < c.nv[f]; k++) {
            const double d = dot3(c.v[f][k], c.v[f][k]);
            if (d > r2) r2 = d;
        }
    c.rmax2 = r2;
}

bool has_box_face(const Cell& c) {
    for (int f = 0; f < c.nf; f++)
        if (c.fid[f] < 0) return true;
    return false;
}

double cell_volume(const Cell& c) {
    double vol = 0.0;
    for (int f = 0; f < c.nf; f++) {
        if (c.fid[f] < 0) return 0.0;
        const V3 v0 = c.v[f][0];
        double area = 0.0;
        for (int k = 1; k < c.nv[f] - 1; k++) {
            const V3 cr = cross3(sub3(c.v[f][k], v0), sub3(c.v[f][k + 1], v0));
            area += std::sqrt(dot3(cr, cr)) / 2.0;
        }
        vol += area * c.fh[f];
    }
    return vol / 6.0;
}

struct Hash {
    double lo[3];
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

double point_volume(const double* p, long i, const Hash& H, double r0, Cell& c) {
    const V3 pi = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    int c3[3];
    for (int a = 0; a < 3; a++) {
        int k = (int)((p[3 * i + a] - H.lo[a]) / H.h);
        c3[a] = k < 0 ? 0 : (k >= H.n[a] ? H.n[a] - 1 : k);
    }
    init_box(c, r0);
    for (int r = 0; r <= RMAX; r++) {
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
                        clip(c, n, j);
                        if (c.overflow) {
                            // a cell still reaching the bounding cube that outgrows the limits
                            // is taken as unbounded (0); a closed cell's overflow stays an error
                            if (box) c.overflow = 0;
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

extern "C" {

// Unnormalised cell volumes (0 on the hull) of np points (np x 3 doubles);
// returns the number of cells that overflowed the fixed face/vertex limits.
long orc_voronoi_volumes(const double* pts, long np, double* vol, int nthreads) {
    vor::Hash H = vor::build_hash(pts, np);
    double ext = 0.0;
    for (int a = 0; a < 3; a++) ext = std::fmax(ext, H.h * H.n[a]);
    const double r0 = 4.0 * ext;
    long over = 0;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : omp_get_max_threads()) reduction(+ : over)
    {
        vor::Cell* c = new vor::Cell;
#pragma omp for schedule(dynamic, 256)
        for (long i = 0; i < np; i++) {
            vol[i] = vor::point_volume(pts, i, H, r0, *c);
            over += c->overflow;
        }
        delete c;
    }
    return over;
}

// estimateDensityV3D without averaging (functions.h:234-245): float volumes
// divided by their maximum.
long orc_voronoi_density(const double* pts, long np, float* dcf, int nthreads) {
    std::vector<double> vol(np);
    const long over = orc_voronoi_volumes(pts, np, vol.data(), nthreads);
    double mx = 0.0;
    for (long i = 0; i < np; i++) {
        dcf[i] = (float)vol[i];
        if (dcf[i] > mx) mx = dcf[i];
    }
    const double inv = 1.0 / mx;
    for (long i = 0; i < np; i++) dcf[i] = (float)((double)dcf[i] * inv);
    return over;
}

}  // extern "C"
