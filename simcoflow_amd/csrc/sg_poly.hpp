// Cut-cell polygons of CellGeoCal (PredictorUV_mod.f90:2258-2352,
// Clsvof_Mod.f90:2159-2253) held in registers.
//
// The box of a cell is cut by the solid line; the fluid side is a polygon
// of at most five vertices, counter-clockwise from the SE corner, and the
// liquid part of it is that polygon clipped by the liquid line.  Both areas
// are shoelace sums.  The reference keeps the vertices in lists and walks
// them afterwards; a list indexed at run time would live in scratch memory
// here, so the polygons are walked vertex by vertex instead: each new vertex
// closes an edge, which adds its shoelace term and is clipped against the
// liquid line at once, and only the first and the previous vertex of each
// polygon are held.  Every term is computed by the same expression and
// summed in the same order as over the lists, so the results are the
// lists' bit for bit (the files including this are compiled with FP
// contraction off).
#pragma once

namespace sg {

// the liquid polygon's shoelace sum, fed one clipped vertex at a time
struct ShoelaceClip {
    int m = 0;                  // vertices so far
    double gx = 0.0, gy = 0.0;  // first vertex
    double hx = 0.0, hy = 0.0;  // previous vertex
    double area = 0.0;
    __device__ __forceinline__ void vertex(double x, double y) {
        if (m == 0) {
            gx = x;
            gy = y;
        } else {
            area = area + 0.5 * (hx * y - x * hy);
        }
        hx = x;
        hy = y;
        ++m;
    }
    // the part of edge (ax, ay) -> (bx, by) with liquid distance <= 0; ad and
    // bd are the ends' distances
    __device__ __forceinline__ void edge(double ax, double ay, double ad, double bx, double by, double bd) {
        if (ad <= 0.0) vertex(ax, ay);
        if (ad * bd < 0.0) {
            const double a = fabs(ad), b = fabs(bd);
            vertex(ax + (bx - ax) * a / (a + b), ay + (by - ay) * a / (a + b));
        }
    }
    __device__ __forceinline__ double close() {
        if (m > 0) area = area + 0.5 * (hx * gy - gx * hy);
        return area;
    }
};

// the fluid polygon walked vertex by vertex, clipped by the liquid line
// nxl x + nyl y + phil <= 0 on the way
struct PolyWalk {
    double nxl, nyl, phil;
    int n = 0;                            // fluid polygon vertices so far
    double fx = 0.0, fy = 0.0, fd = 0.0;  // first vertex and its liquid distance
    double qx = 0.0, qy = 0.0, qd = 0.0;  // previous vertex
    double vol = 0.0;
    ShoelaceClip liq;
    __device__ PolyWalk(double a, double b, double c) : nxl(a), nyl(b), phil(c) {}
    __device__ __forceinline__ void put(double x, double y) {
        if (n >= 5) return;
        const double d = x * nxl + y * nyl + phil;
        if (n == 0) {
            fx = x;
            fy = y;
            fd = d;
        } else {
            vol = vol + 0.5 * (qx * y - x * qy);
            liq.edge(qx, qy, qd, x, y, d);
        }
        qx = x;
        qy = y;
        qd = d;
        ++n;
    }
    __device__ __forceinline__ void close() {
        if (n > 0) {
            vol = vol + 0.5 * (qx * fy - fx * qy);
            liq.edge(qx, qy, qd, fx, fy, fd);
        }
        liq.close();
    }
};

// the fluid polygon's vertices in the reference's order: calls put(x, y)
// for each (the solid distances of the corners decide which exist)
template <class Put>
__device__ __forceinline__ void fluid_vertices(double nxs, double nys, double phis, double dx, double dy, Put&& put) {
    const double dx2 = dx / 2.0, dy2 = dy / 2.0;
    const double d0 = dx2 * nxs - dy2 * nys + phis;
    const double d1 = -dx2 * nxs - dy2 * nys + phis;
    const double d2 = -dx2 * nxs + dy2 * nys + phis;
    const double d3 = dx2 * nxs + dy2 * nys + phis;
    if (d0 >= 0.0) put(dx2, -dy2);
    if (d0 * d1 < 0.0) put((dy2 * nys - phis) / nxs, -dy2);
    if (d1 >= 0.0) put(-dx2, -dy2);
    if (d1 * d2 < 0.0) put(-dx2, (dx2 * nxs - phis) / nys);
    if (d2 >= 0.0) put(-dx2, dy2);
    if (d2 * d3 < 0.0) put((-dy2 * nys - phis) / nxs, dy2);
    if (d3 >= 0.0) put(dx2, dy2);
    if (d3 * d0 < 0.0) put(dx2, (-dx2 * nxs - phis) / nys);
}

// CellGeoCal: the fluid (vols = 1 - |fluid area| / box) and liquid
// (volf = |liquid area| / box) fractions of a dx x dy box whose solid line
// is (nxs, nys, phis) and liquid line (nxl, nyl, phil), both relative to
// the box centre
__device__ __forceinline__ void cut_cell_fractions(double nxs, double nys, double phis, double nxl, double nyl,
                                                   double phil, double dx, double dy, double& vols, double& volf) {
    PolyWalk w(nxl, nyl, phil);
    fluid_vertices(nxs, nys, phis, dx, dy, [&](double x, double y) { w.put(x, y); });
    w.close();
    vols = 1.0 - fabs(w.vol / (dx * dy));
    volf = fabs(w.liq.area / (dx * dy));
}

}  // namespace sg
