"""Voronoi density compensation restated from the reference (test
infrastructure only).

estimateDensityV3D (functions.h:161-258) builds a 3D Delaunay triangulation
of all trajectory points (CGAL), and for every point takes the volume of its
Voronoi cell as  sum over Delaunay edges (p, q) of  area(dual facet) * |p - q| / 6
(volume_of_finite_dual_cell, functions.h:128-159), where the dual facet is the
polygon of the circumcentres of the tetrahedra around the edge, its area a fan
of triangles from the first circumcentre (get_area_of_dual_delaunay_edge,
functions.h:75-125).  A point on the convex hull (an edge to the infinite
vertex) gets 0.  The volumes are then divided by their maximum
(functions.h:234-245).

CGAL is not available; the same quantities are taken here from Qhull through
scipy.spatial.Voronoi: every ridge is the dual facet of one Delaunay edge,
``ridge_vertices`` lists its circumcentres in cyclic order and ``-1`` marks an
unbounded ridge.
"""
import numpy as np
from scipy.spatial import Voronoi


def voronoi_volumes(points):
    """Unnormalised Voronoi cell volumes (0 for hull points), float64."""
    pts = np.asarray(points, np.float64)
    uniq, inv = np.unique(pts, axis=0, return_inverse=True)
    inv = np.asarray(inv).reshape(-1)
    vor = Voronoi(uniq)
    vol = np.zeros(len(uniq))
    unbounded = np.zeros(len(uniq), bool)
    for (a, b), rv in zip(vor.ridge_points, vor.ridge_vertices):
        if -1 in rv:
            unbounded[a] = unbounded[b] = True
            continue
        poly = vor.vertices[rv]
        v0 = poly[0]
        area = 0.0
        for i in range(1, len(poly) - 1):
            area += np.linalg.norm(np.cross(poly[i] - v0, poly[i + 1] - v0)) / 2.0
        h = np.linalg.norm(uniq[a] - uniq[b])
        vol[a] += area * h
        vol[b] += area * h
    vol /= 6.0
    vol[unbounded] = 0.0
    return vol[inv]


def normalise(vol):
    """The tail of estimateDensityV3D without averaging (functions.h:234-245):
    float32 cell volumes times the double reciprocal of their maximum."""
    cell = np.asarray(vol).astype(np.float32)
    mx = float(cell.max())
    return (cell.astype(np.float64) * (1.0 / mx)).astype(np.float32)


def estimate_density(points):
    """estimateDensityV3D without averaging: float32 weights = volume / max."""
    return normalise(voronoi_volumes(points))
