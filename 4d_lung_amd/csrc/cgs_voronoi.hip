// cgs_voronoi.hip -- Voronoi density compensation (estimateDensityV3D,
// functions.h:161-258): the kernels that build the cells, one wave per cell.
//
// The reference takes each sample's Voronoi cell volume from a CGAL Delaunay
// triangulation (sum over Delaunay edges of dual-facet area * edge length / 6,
// 0 for points on the convex hull) and divides by the maximum.  Here every
// cell is built on its own by half-space clipping (cgs_voronoi_cell.h): start
// from the bounding box of all points widened by one hash cell and intersect
// it with the bisector half-spaces of the other points, visiting a uniform
// hash grid shell by shell until no unvisited point can cut the cell (the
// farthest vertex is closer than half the distance of the next shell).
// Clipped rule (per-bin weights): cells that keep a face of the box get
// volume 0 -- the hull points, and interior points whose Voronoi region
// reaches more than one hash cell beyond the point cloud.  Hull rule (the
// reference's): those cells are built again from a box wider than the point
// cloud by kVorWideBox times its extent, visiting every hash cell until the
// radius bound closes, and only the cells that still keep a box face (the
// unbounded ones) get 0.
//
// The 64 lanes of a wave build one cell together: candidate points are
// tested 64 at a time (one per lane) and each clip runs lane-parallel over
// the vertex pool, the ring entries and the faces.  Cells live in LDS with
// small limits (16 KB, ten cells per CU); cells that outgrow them are listed
// and built again with medium limits (31 KB of LDS), and what outgrows those
// with large limits in a device workspace, one cell per wave.  Built with -ffp-contract=off: every operation is a
// separately rounded IEEE double operation in a fixed order, so the CPU
// statement of the same algorithm reproduces the volumes bit for bit, and a
// cell's bits do not depend on the limits it was built with.
#include "cgs_internal.h"
#include "cgs_kernels.h"
#include "cgs_voronoi_cell.h"

namespace cgs {
namespace vor {

// Cells in LDS, grid-stride over the listed points (every point when list
// is null); points whose cell outgrew the limits are appended to redo.
template <class LM>
__global__ __launch_bounds__(64) void k_voronoi_lds(const double* __restrict__ p, long long n,
                                                   const int* __restrict__ list, bool wide, VorHash H,
                                                   double* __restrict__ vol, int* __restrict__ redo,
                                                   int* __restrict__ nredo) {
    __shared__ WCell<LM> C;
    const int lane = threadIdx.x;
    WState S;
    for (long long k = blockIdx.x; k < n; k += gridDim.x) {
        const long long i = list ? (long long)list[k] : k;
        const double v = wpoint_volume(p, i, H, C, S, false, wide, lane);
        if (lane == 0) {
            vol[i] = v;
            if (S.overflow) redo[atomicAdd(nredo, 1)] = (int)i;
        }
    }
}

// Last tier: large limits, each wave's cell in its own slot of a device
// workspace, over the listed points (every point when redo is null).  Closed
// cells that outgrow these limits too are counted.
__global__ __launch_bounds__(64) void k_voronoi_global(const double* __restrict__ p, bool wide, VorHash H,
                                                      const int* __restrict__ redo, long long nredo,
                                                      WCell<LargeLimits>* __restrict__ cells,
                                                      double* __restrict__ vol, int* __restrict__ overflow) {
    WCell<LargeLimits>& C = cells[blockIdx.x];
    const int lane = threadIdx.x;
    WState S;
    for (long long k = blockIdx.x; k < nredo; k += gridDim.x) {
        const long long i = redo ? (long long)redo[k] : k;
        const double v = wpoint_volume(p, i, H, C, S, true, wide, lane);
        if (lane == 0) {
            vol[i] = v;
            if (S.overflow) atomicAdd(overflow, 1);
        }
    }
}

}  // namespace vor

// Workgroups of the LDS tiers: one cell per wave; more workgroups than fit
// at once, so that the hardware keeps every CU busy until the list drains.
static unsigned vor_blocks(long long n) { return (unsigned)std::min<long long>(n, 1LL << 16); }

hipError_t launch_voronoi(const double* pts, long long n, const int* list, bool wide, const VorHash& H, double* vol,
                          int* redo, int* nredo, hipStream_t s) {
    if (n == 0) return hipSuccess;
    vor::k_voronoi_lds<vor::SmallLimits><<<vor_blocks(n), 64, 0, s>>>(pts, n, list, wide, H, vol, redo, nredo);
    return hipGetLastError();
}

hipError_t launch_voronoi_mid(const double* pts, long long n, const int* list, bool wide, const VorHash& H,
                              double* vol, int* redo, int* nredo, hipStream_t s) {
    if (n == 0) return hipSuccess;
    vor::k_voronoi_lds<vor::MediumLimits><<<vor_blocks(n), 64, 0, s>>>(pts, n, list, wide, H, vol, redo, nredo);
    return hipGetLastError();
}

size_t voronoi_large_cell_bytes() { return sizeof(vor::WCell<vor::LargeLimits>); }

hipError_t launch_voronoi_large(const double* pts, bool wide, const VorHash& H, const int* redo, long long nredo,
                                int nwaves, void* cells, double* vol, int* overflow, hipStream_t s) {
    if (nredo == 0 || nwaves <= 0) return hipSuccess;
    vor::k_voronoi_global<<<(unsigned)nwaves, 64, 0, s>>>(pts, wide, H, redo, nredo,
                                                          (vor::WCell<vor::LargeLimits>*)cells, vol, overflow);
    return hipGetLastError();
}

}  // namespace cgs
