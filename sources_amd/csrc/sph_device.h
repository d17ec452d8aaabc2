// sph_device.h -- device building blocks of the per-step particle loop:
// the smoothing kernel, the pair predicate and the cell-list neighbour walk.
// All arithmetic is FP64 (the reference's real(dp)).  Constants that the
// reference writes as default-real (FP32) literals are the FP32 values widened
// to double, written here as exact hex-float literals.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SPH_PI_F   0x1.921fb6p+1    // pi = 3.14159265358979323846 (FP32 literal), m_particles.f90:1483
#define SPH_C23_F  0x1.555556p-1    // 2./3. (FP32), m_particles.f90:1508, :2902
#define SPH_YIELD_TOL_F 0x1.4f8b58p-17  // 1.0e-5 (FP32), m_particles.f90:3492
#define SPH_EPSP_F 0x1.4f8b58p-18   // 0.000005 (FP32), m_particles.f90:3756

// ---------------------------------------------------------------------------
// Smoothing kernel W and grad W for dim = 2, kernel (m_particles.f90:1469-1602).
// Evaluation order follows the Fortran expressions left to right; integer
// powers are repeated squaring.  `contract(off)` keeps every product and sum
// separately rounded so W and grad W agree with the CPU restatement.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void sph_kernel(int skf, double r, double dx, double dy, double hsml,
                                           double& w, double& dwx, double& dwy)
{
#pragma clang fp contract(off)
    const double q = r / hsml;
    w = 0.0; dwx = 0.0; dwy = 0.0;
    if (skf == 1) {
        const double factor = 15.0 / (7.0 * SPH_PI_F * hsml * hsml);
        if (q >= 0.0 && q <= 1.0) {
            w = factor * (SPH_C23_F - q * q + ((q * q) * q) / 2.0);
            const double g = factor * (-2.0 + 1.5 * q) / (hsml * hsml);
            dwx = g * dx;
            dwy = g * dy;
        } else if (q > 1.0 && q <= 2.0) {
            const double t = 2.0 - q;
            w = factor * 1.0 / 6.0 * ((t * t) * t);
            const double g = factor * 1.0 / 6.0 * 3.0 * (t * t) / hsml;
            dwx = -(g * (dx / r));
            dwy = -(g * (dy / r));
        }
    } else if (skf == 2) {
        const double factor = 1.0 / ((hsml * hsml) * SPH_PI_F);
        if (q >= 0.0 && q <= 3.0) {
            w = factor * exp(-q * q);
            dwx = w * (-2.0 * dx / hsml / hsml);
            dwy = w * (-2.0 * dy / hsml / hsml);
        }
    } else if (skf == 3) {
        const double factor = 7.0 / (478.0 * SPH_PI_F * hsml * hsml);
        const double a = 3.0 - q, b = 2.0 - q, c = 1.0 - q;
        const double a2 = a * a, b2 = b * b, c2 = c * c;
        const double a4 = a2 * a2, b4 = b2 * b2, c4 = c2 * c2;
        if (q >= 0.0 && q <= 1.0) {
            w = factor * (a4 * a - 6.0 * (b4 * b) + 15.0 * (c4 * c));
            const double g = (-120.0 + 120.0 * q - 50.0 * (q * q)) / (hsml * hsml);
            dwx = factor * (g * dx);
            dwy = factor * (g * dy);
        } else if (q > 1.0 && q <= 2.0) {
            w = factor * (a4 * a - 6.0 * (b4 * b));
            const double g = factor * (-(5.0 * a4) + 30.0 * b4) / hsml;
            dwx = g * (dx / r);
            dwy = g * (dy / r);
        } else if (q > 2.0 && q <= 3.0) {
            w = factor * (a4 * a);
            const double g = factor * (-(5.0 * a4)) / hsml;
            dwx = g * (dx / r);
            dwy = g * (dy / r);
        }
    } else if (skf == 4) {
        const double factor = 7.0 / (4.0 * SPH_PI_F * hsml * hsml);
        if (q >= 0.0 && q <= 2.0) {
            const double t = 1.0 - q / 2.0;
            const double t2 = t * t;
            w = factor * ((t2 * t2) * (1.0 + 2.0 * q));
            const double g = factor * (-5.0 + 15.0 * q / 2.0 - 15.0 * (q * q) / 4.0 + 5.0 * ((q * q) * q) / 8.0)
                             / (hsml * hsml);
            dwx = g * dx;
            dwy = g * dy;
        }
    }
}

// Pair predicate of direct_find (m_particles.f90:965-973): dx = x_i - x_j,
// driac = dx*dx + dy*dy (unfused), sqrt(driac) < scale_k * mhsml.  The
// correctly rounded device sqrt makes the pair set bit-identical to the CPU.
__device__ __forceinline__ bool sph_pair_test(double xa, double ya, double xb, double yb, double reach,
                                              double& dx, double& dy, double& r)
{
#pragma clang fp contract(off)
    dx = xa - xb;
    dy = ya - yb;
    const double dr = dx * dx + dy * dy;
    r = __builtin_sqrt(dr);
    return r < reach;
}

__device__ __forceinline__ double sph_mhsml(double ha, double hb)
{
#pragma clang fp contract(off)
    return (ha + hb) / 2.0;
}

// ---------------------------------------------------------------------------
// Uniform cell grid shared by both phases.  Cell size >= the largest support
// radius, so every neighbour lies in the 3 x 3 block of cells around a
// particle.  Keys are row-major (key = cy * ncx + cx): the three cells of a
// block row are consecutive keys and therefore one contiguous index range.
// ---------------------------------------------------------------------------
struct Grid {
    double x0, y0, x1, y1, inv_cs;
    int ncx, ncy;
    // reference link_list sorting grid (init_grid), used when nnps == 2
    double lmin_x, lmin_y, lscale_x, lscale_y;  // ngridx/dgeomx per axis
    int lreach_x, lreach_y;                      // ghsmlx
    int link_list;                               // 1: restrict pairs to the reference grid reach
};

__device__ __forceinline__ void grid_cell(const Grid& g, double x, double y, int& cx, int& cy)
{
    // Clamp in floating point first: the double -> int conversion of a huge or
    // NaN coordinate is then well defined.
    const double fx = fmin(fmax(floor((x - g.x0) * g.inv_cs), -1.0), (double)g.ncx);
    const double fy = fmin(fmax(floor((y - g.y0) * g.inv_cs), -1.0), (double)g.ncy);
    cx = (x == x) ? (int)fx : -1;
    cy = (y == y) ? (int)fy : -1;
}

// Reference grid_geom cell (m_particles.f90:1137-1138), 1-based.
__device__ __forceinline__ int ref_gcell(double scale, double mn, double x)
{
#pragma clang fp contract(off)
    return (int)(scale * (x - mn) + 1.0);
}

struct CellList {
    const int* start;   // ncell + 1 entries; cell c holds order[start[c] .. start[c+1])
    const int* order;   // particle indices sorted by cell, ascending index inside a cell
};

// Visit every particle of `cl` in the 3 x 3 cell block around (x, y).
template <class F>
__device__ __forceinline__ void for_each_candidate(const Grid& g, const CellList& cl, double x, double y, F&& f)
{
    int cx, cy;
    grid_cell(g, x, y, cx, cy);
    // A particle outside the grid has already raised the out-of-range flag;
    // clamping keeps its walk inside the cell arrays.
    cx = min(max(cx, 0), g.ncx - 1);
    cy = min(max(cy, 0), g.ncy - 1);
    const int c0 = max(cx - 1, 0), c1 = min(cx + 1, g.ncx - 1);
#pragma unroll 1
    for (int row = max(cy - 1, 0); row <= min(cy + 1, g.ncy - 1); ++row) {
        const int k0 = cl.start[row * g.ncx + c0];
        const int k1 = cl.start[row * g.ncx + c1 + 1];
#pragma unroll 1
        for (int k = k0; k < k1; ++k) f(cl.order[k]);
    }
}
