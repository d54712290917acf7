// cgs_voronoi.hip -- Voronoi density compensation (estimateDensityV3D,
// functions.h:161-258) by half-space clipping, one cell per thread.
//
// The reference takes each sample's Voronoi cell volume from a CGAL Delaunay
// triangulation (sum over Delaunay edges of dual-facet area * edge length / 6,
// 0 for points on the convex hull) and divides by the maximum.  Here every
// cell is built on its own: start from a cube around the point and intersect
// it with the bisector half-spaces of the other points, visiting a uniform
// hash grid shell by shell in a fixed order until no unvisited point can cut
// the cell (the farthest vertex is closer than half the distance of the next
// shell).  Cells that keep a face of the cube after the last shell are taken
// as unbounded (volume 0).  The volume is the reference's face sum
// area * |q - p| / 6 with areas as triangle fans from the first vertex.
//
// Built with -ffp-contract=off; every operation is a separately rounded
// IEEE double operation in a fixed order, so the weights are reproducible
// bit for bit by a CPU evaluation of the same statements.
#include "cgs_internal.h"
#include "cgs_kernels.h"

namespace cgs {
namespace vor {

constexpr int FMAX = 48;
constexpr int VMAX = 32;
constexpr int RMAX = 10;

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

__device__ __forceinline__ double dot3(const V3& a, const V3& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 sub3(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 cross3(const V3& a, const V3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

__device__ void init_box(Cell& c, double r0) {
    const signed char q[6][4][3] = {
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

__device__ __forceinline__ bool ang_less(double a1, double b1, double a2, double b2) {
    const bool u1 = b1 > 0.0 || (b1 == 0.0 && a1 > 0.0);
    const bool u2 = b2 > 0.0 || (b2 == 0.0 && a2 > 0.0);
    if (u1 != u2) return u1;
    return a1 * b2 - a2 * b1 > 0.0;
}

__device__ void clip(Cell& c, const V3& n, int nid) {
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
        for (int i = 1; i < nu; i++) {
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
            c.fh[nf] = sqrt(dot3(n, n));
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

__device__ bool has_box_face(const Cell& c) {
    for (int f = 0; f < c.nf; f++)
        if (c.fid[f] < 0) return true;
    return false;
}

__device__ double cell_volume(const Cell& c) {
    double vol = 0.0;
    for (int f = 0; f < c.nf; f++) {
        if (c.fid[f] < 0) return 0.0;
        const V3 v0 = c.v[f][0];
        double area = 0.0;
        for (int k = 1; k < c.nv[f] - 1; k++) {
            const V3 cr = cross3(sub3(c.v[f][k], v0), sub3(c.v[f][k + 1], v0));
            area += sqrt(dot3(cr, cr)) / 2.0;
        }
        vol += area * c.fh[f];
    }
    return vol / 6.0;
}

__global__ __launch_bounds__(64) void k_voronoi(const double* __restrict__ p, long long np, VorHash H, double r0,
                                               double* __restrict__ vol, int* __restrict__ overflow) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    Cell c;
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
                        clip(c, n, j);
                        if (c.overflow) {
                            // a cell still reaching the bounding cube that outgrows the limits
                            // is taken as unbounded (0); a closed cell's overflow stays an error
                            if (box) c.overflow = 0;
                            vol[i] = 0.0;
                            if (c.overflow) atomicAdd(overflow, 1);
                            return;
                        }
                    }
                }
        const double reach = (double)r * H.h;
        if (0.25 * reach * reach >= c.rmax2) break;
    }
    vol[i] = cell_volume(c);
    if (c.overflow) atomicAdd(overflow, 1);
}

}  // namespace vor

hipError_t launch_voronoi(const double* pts, long long np, const VorHash& H, double r0, double* vol, int* overflow,
                          hipStream_t s) {
    // 64 KB of dynamic LDS per block keeps at most two blocks per CU, which
    // bounds the private (scratch) memory the runtime reserves for the cells.
    vor::k_voronoi<<<(unsigned)((np + 63) / 64), 64, 65536, s>>>(pts, np, H, r0, vol, overflow);
    return hipGetLastError();
}

}  // namespace cgs
