// sph_fused.h -- the fused two-phase step (FP64), tuned for CDNA4.
//
// Per phase and step:
//   pack      thread t of the cell order: half-step update of particle order[t]
//             and its neighbour record t, cell-sorted and stored as chunk
//             planes (chunk c of record t is element t of plane c).  Record
//             chunks (16 bytes): 0 x,y | 1 vx,vy | 2 rho,m | 3 vof,p | 4 s3,c |
//             5 s1,s2 | 6 h,itype; m carries the sign of itype (negative for
//             virtual particles).  The lanes of a wave hold nearby particles
//             whose k-th partners are nearby in the cell order too, so one
//             chunk load of all lanes touches a few cache lines of a plane (a
//             128-byte record per particle would touch one line per lane: the
//             address unit serialises on those).  The loops load only the
//             chunks they use: 1/rho and m/rho of a partner are recomputed;
//   stage A   candidate scan of the 3 cell rows with a sqrt-free exact
//             predicate (dr < T, T precomputed so that dr < T <=> sqrt(dr) < R),
//             accepted partners kept as one concatenated 128-bit mask; then a
//             dense loop over the mask bits evaluates the pair terms.  The masks
//             are stored for stage B, so each pair is tested once per step;
//   stage B   dense loop over the stored masks (water).  Soil stage B needs
//             nothing from the stage A of other soil particles, so soil runs
//             both stages in one kernel over one partner loop.
// Lanes of a wavefront hold consecutive cell-sorted particles; the dense loops
// diverge only by the lanes' neighbour-count spread.
#pragma once
#include "sph_kernels.h"

// record slots (doubles)
enum : int {
    R_X = 0, R_Y, R_VX, R_VY, R_RHO, R_M, R_VOF, R_P, R_S3, R_C, R_S1, R_S2, R_H, R_IT, R_PAD0, R_PAD1,
    R_STRIDE = 16
};

#ifndef SPHG_WPB
#define SPHG_WPB 4           // waves per workgroup of the stage kernels (consecutive particles share a CU)
#endif
constexpr int STAGE_BLOCK = 64 * SPHG_WPB;
#ifndef SPHG_STAGE_WAVES
#define SPHG_STAGE_WAVES 0   // 0: compiler's choice; n: ask for at least n waves per SIMD
#endif
#if SPHG_STAGE_WAVES > 0
#define STAGE_ATTR __launch_bounds__(STAGE_BLOCK) __attribute__((amdgpu_waves_per_eu(SPHG_STAGE_WAVES, 8)))
#else
#define STAGE_ATTR __launch_bounds__(STAGE_BLOCK)
#endif
#ifndef SPHG_AB_WAVES
#define SPHG_AB_WAVES 0      // the same for the one-pass soil kernel
#endif
#if SPHG_AB_WAVES > 0
#define STAGE_ATTR_AB __launch_bounds__(STAGE_BLOCK) __attribute__((amdgpu_waves_per_eu(SPHG_AB_WAVES, 8)))
#else
#define STAGE_ATTR_AB __launch_bounds__(STAGE_BLOCK)
#endif

struct FusedK {
    // dr thresholds and uniform-h flags for same-phase (water, soil) and cross pairs
    double thr_w, thr_s, thr_x;
    int uni_w, uni_s, uni_x;
};

// q^n for an integer 0 <= n <= 64 by binary powering (within a few ulp of pow).
__device__ __forceinline__ double ipow(double q, int n)
{
    double r = 1.0, b = q;
#pragma unroll 1
    for (; n > 0; n >>= 1) {
        if (n & 1) r *= b;
        b *= b;
    }
    return r;
}

// 1/d and 1/sqrt(d) from the hardware estimates plus two Newton steps
// (each step doubles the correct bits), i.e. to about one ulp; far shorter
// dependency chains than IEEE-exact FP64 division and square root.
__device__ __forceinline__ double fast_rcp(double d)
{
    double x = __builtin_amdgcn_rcp(d);
    double e = fma(-d, x, 1.0);
    x = fma(x, e, x);
    e = fma(-d, x, 1.0);
    return fma(x, e, x);
}

__device__ __forceinline__ double fast_rsq(double d)
{
    double y = __builtin_amdgcn_rsq(d);
    const double h = 0.5 * d;
    y = y * fma(-h, y * y, 1.5);
    y = y * fma(-h, y * y, 1.5);
    return y;
}

// r, 1/r and 1/r^2 of a squared distance dr > 0.
struct Dist {
    double r, ir, idr;
};

__device__ __forceinline__ Dist dist_of(double dr)
{
    const double ir = fast_rsq(dr);
    return Dist{dr * ir, ir, ir * ir};
}

// Kernel constants for one smoothing length (uniform-h fast path).
struct KC {
    double inv_h, f, fh1, fh2;
};

__device__ __forceinline__ KC make_kc(int skf, double h)
{
    KC k;
    k.inv_h = 1.0 / h;
    double f;
    if (skf == 1) f = 15.0 / (7.0 * SPH_PI_F * h * h);
    else if (skf == 2) f = 1.0 / ((h * h) * SPH_PI_F);
    else if (skf == 3) f = 7.0 / (478.0 * SPH_PI_F * h * h);
    else f = 7.0 / (4.0 * SPH_PI_F * h * h);
    k.f = f;
    k.fh1 = f * k.inv_h;
    k.fh2 = f * k.inv_h * k.inv_h;
    return k;
}

// W and grad W (m_particles.f90:1469-1602) with precomputed constants;
// inv_r = 1/r is only used where the reference divides by r.
template <int SKF, bool NEED_W>
__device__ __forceinline__ void kern_fast(const KC& k, double r, double inv_r, double dx, double dy,
                                          double& w, double& gx, double& gy)
{
    const double q = r * k.inv_h;
    w = 0.0; gx = 0.0; gy = 0.0;
    if constexpr (SKF == 1) {
        // both pieces, then a select: no divergent branch in the pair loop
        const bool in1 = q <= 1.0, in2 = q <= 2.0;
        const double t = 2.0 - q;
        const double s1 = k.fh2 * (-2.0 + 1.5 * q);
        const double s2 = -(0.5 * k.fh1) * (t * t) * inv_r;
        const double s = in1 ? s1 : (in2 ? s2 : 0.0);
        if (NEED_W) {
            const double w1 = k.f * (SPH_C23_F - q * q + ((q * q) * q) * 0.5);
            const double w2 = k.f * (1.0 / 6.0) * ((t * t) * t);
            w = in1 ? w1 : (in2 ? w2 : 0.0);
        }
        gx = s * dx; gy = s * dy;
    } else if constexpr (SKF == 4) {
        const bool in2 = q <= 2.0;
        const double t = 1.0 - 0.5 * q;
        const double t2 = t * t;
        if (NEED_W) w = in2 ? k.f * ((t2 * t2) * (1.0 + 2.0 * q)) : 0.0;
        double s = k.fh2 * (-5.0 + 7.5 * q - 3.75 * (q * q) + 0.625 * ((q * q) * q));
        s = in2 ? s : 0.0;
        gx = s * dx; gy = s * dy;
    } else if constexpr (SKF == 2) {
        if (q <= 3.0) {
            const double e = k.f * exp(-q * q);
            if (NEED_W) w = e;
            const double s = -2.0 * e * (k.inv_h * k.inv_h);
            gx = s * dx; gy = s * dy;
        }
    } else {
        const double a = 3.0 - q, b = 2.0 - q, c = 1.0 - q;
        const double a2 = a * a, b2 = b * b, c2 = c * c;
        const double a4 = a2 * a2, b4 = b2 * b2, c4 = c2 * c2;
        if (q <= 1.0) {
            if (NEED_W) w = k.f * (a4 * a - 6.0 * (b4 * b) + 15.0 * (c4 * c));
            const double s = k.fh2 * (-120.0 + 120.0 * q - 50.0 * (q * q));  // reference polynomial
            gx = s * dx; gy = s * dy;
        } else if (q <= 2.0) {
            if (NEED_W) w = k.f * (a4 * a - 6.0 * (b4 * b));
            const double s = k.fh1 * (-(5.0 * a4) + 30.0 * b4) * inv_r;
            gx = s * dx; gy = s * dy;
        } else if (q <= 3.0) {
            if (NEED_W) w = k.f * (a4 * a);
            const double s = k.fh1 * (-(5.0 * a4)) * inv_r;
            gx = s * dx; gy = s * dy;
        }
    }
}

// The records are stored as chunk planes: chunk c of record k is element k
// of plane c (n 16-byte elements per plane).  The lanes of a wave read one
// chunk of nearby partners at a time, so a plane layout touches a few cache
// lines per load where a 128-byte-per-record layout touches one per lane.
struct RecV {
    const double2* p;
    int n;
    __device__ __forceinline__ double2 operator()(int k, int c) const { return p[(size_t)c * n + k]; }
};

__device__ __forceinline__ RecV recv(const double* rec, int n)
{
    return RecV{reinterpret_cast<const double2*>(rec), n};
}

// Exact pair predicate: sqrt(dx^2+dy^2) < scale_k * mh  (direct_find) or
// < scale_k * h(1) within the reference grid reach (link_list).
__device__ __forceinline__ bool pred(bool uni, double thr, double dr, double ha, double hb, const KParams& K)
{
    if (uni) return dr < thr;
    return __builtin_sqrt(dr) < K.scale_k * sph_mhsml(ha, hb);
}

// Cell-row ranges of the 3 x 3 block around (x, y) in the cell order.
struct Rows {
    int k0[3], k1[3];
};

// Row ranges of the 3 x 3 block around cell (cx, cy).
__device__ __forceinline__ Rows block_rows_at(const Grid& g, const int* cell_start, int cx, int cy)
{
    Rows R;
    const int c0 = max(cx - 1, 0), c1 = min(cx + 1, g.ncx - 1);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int row = cy - 1 + r;
        if (row >= 0 && row < g.ncy) {
            R.k0[r] = cell_start[row * g.ncx + c0];
            R.k1[r] = cell_start[row * g.ncx + c1 + 1];
        } else {
            R.k0[r] = 0;
            R.k1[r] = 0;
        }
    }
    return R;
}

// Cell of a position, clamped into the grid (a particle outside it has
// raised the out-of-range flag at the cell-list build).
__device__ __forceinline__ void clamped_cell(const Grid& g, double x, double y, int& cx, int& cy)
{
    grid_cell(g, x, y, cx, cy);
    cx = min(max(cx, 0), g.ncx - 1);
    cy = min(max(cy, 0), g.ncy - 1);
}

__device__ __forceinline__ Rows block_rows(const Grid& g, const int* cell_start, double x, double y)
{
    int cx, cy;
    clamped_cell(g, x, y, cx, cy);
    return block_rows_at(g, cell_start, cx, cy);
}

constexpr int Q_STRIDE = 16;   // floats per record: four float4 chunk planes

// FP32 records of the mixed-precision form, as chunk planes like RecV.
struct RecQ {
    const float4* p;
    int n;
    __device__ __forceinline__ float4 operator()(int k, int c) const { return p[(size_t)c * n + k]; }
};

__device__ __forceinline__ RecQ recqv(const float* q, int n)
{
    return RecQ{reinterpret_cast<const float4*>(q), n};
}

__device__ __forceinline__ int cell_tag(int cx, int cy, bool is_virtual)
{
    return (is_virtual ? (int)0x80000000 : 0) | (cx << 15) | cy;
}
__device__ __forceinline__ int tag_cx(int tag) { return (tag >> 15) & 0x7fff; }
__device__ __forceinline__ int tag_cy(int tag) { return tag & 0x7fff; }

// FP32 neighbour record of the mixed-precision form (the layout is
// described in sph_mixed.h): chunks q0 and q1 of the record at R (plane 0),
// planes pn elements apart.
// key >= 0: the record's cell is the row-major cell `key` (the cell of the
// particle at the neighbour-list rebuild, which it may have left since);
// otherwise the cell of its position.
__device__ __forceinline__ void write_q01(float4* R, size_t pn, const Grid& g, double cs, double x, double y, double h,
                                          bool virt, double vx, double vy, double drho, double m, int key)
{
    int cx, cy;
    if (key >= 0) {
        cy = key / g.ncx;
        cx = key - cy * g.ncx;
    } else {
        grid_cell(g, x, y, cx, cy);
        cx = min(max(cx, 0), g.ncx - 1);
        cy = min(max(cy, 0), g.ncy - 1);
    }
    const double xc = g.x0 + ((double)cx + 0.5) * cs, yc = g.y0 + ((double)cy + 0.5) * cs;
    R[0] = make_float4((float)(x - xc), (float)(y - yc), __int_as_float(cell_tag(cx, cy, virt)), (float)h);
    R[pn] = make_float4((float)vx, (float)vy, (float)drho, (float)m);
}

// ------------------------------------------------------------------------
// pack: half-step update (first_half, declarations_sph.f90:643-684, and the
// soil Drucker-Prager projections) fused with writing the neighbour record.
// ------------------------------------------------------------------------
// Displacement since the neighbour-list rebuild, per record slot t (the
// record order is fixed between rebuilds).  A rebuild (rebuild = 1) records
// the position.  Later steps raise bit 0 of flag (rebuild) when the
// displacement d passes trig, or when d plus `look` times the speed (look =
// a few steps' dt) passes near, so that a rebuild that takes effect two
// steps after the flag is raised still comes before d reaches half the
// skin; and bit 1 when d is past half the skin (d^2 > hard2): the lists may
// then have missed a pair.  One atomic per wave.  x == nullptr: no lists.
struct NRef {
    double *x, *y;
    int* flag;
    double trig, near, hard2, look;
    int rebuild;
    const int* cell;   // cell key per record slot since the rebuild (FP32 records)
};

__device__ __forceinline__ int record_cell(const NRef& R, int t) { return R.cell && !R.rebuild ? R.cell[t] : -1; }

__device__ __forceinline__ void track_displacement(const NRef& R, int t, double x, double y, double vx, double vy)
{
    if (!R.x) return;
    if (R.rebuild) {
        R.x[t] = x;
        R.y[t] = y;
        return;
    }
    const double dx = x - R.x[t], dy = y - R.y[t];
    const double d2 = dx * dx + dy * dy;
    const double d = sqrt(d2);
    const bool hard = !(d2 <= R.hard2);
    const bool due = hard || !(d <= R.trig) || !(d + R.look * sqrt(vx * vx + vy * vy) <= R.near);
    const uint64_t any_due = __ballot(due), any_hard = __ballot(hard);
    const uint64_t active = __ballot(1);
    if (any_due && (int)(threadIdx.x & 63) == __ffsll((unsigned long long)active) - 1)
        atomicOr(R.flag, any_hard ? 3 : 1);
}

// q != nullptr selects the FP32 record of the mixed-precision form.
// Pack of water particle i into record slot t from its current values x,
// v, rho, vof (the state's, or those the final update of the step before has
// just computed): first_half (do_half), the equation of state, the record.
// lean: the half-step levels that nothing reads before the next final update
// (which starts from the saved full-step levels) stay out of the state; only
// for a pack that the next step's stages read from FP64 records alone.
// lean_v: the same for the velocity alone (no stage reads the state's).
__device__ __forceinline__ void pack_water_at(const DPhase& P, const KParams& K, double* rec, int t, int i,
                                              int do_half, bool lean, bool lean_v, float* q, const Grid& g,
                                              double cs, const NRef& NR, double x, double y, double rho, double vof,
                                              double vx, double vy)
{
#pragma clang fp contract(off)
    track_displacement(NR, t, x, y, vx, vy);
    const int it = P.itype[i];
    if (do_half) {
        const double h = K.dt / 2.0;
        if (!K.summation_density) {   // first_half's rho update (declarations_sph.f90:647-650)
            P.rho_min[i] = rho;
            rho = rho + h * P.drho[i];
            if (!lean) P.rho[i] = rho;
        }
        if (K.volume_fraction && K.two_phase) {
            P.vof_min[i] = vof;
            vof = vof + h * P.dvof[i];
            if (!lean || K.summation_density) P.vof[i] = vof;   // the renormalisation reads it then
        }
        if (it > 0) {
            P.vx_min[i] = vx;
            P.vy_min[i] = vy;
            vx = vx + h * P.ax[i];
            vy = vy + h * P.ay[i];
            if (!lean_v) {
                P.vx[i] = vx;
                P.vy[i] = vy;
            }
        }
    }
    const double m = P.mass[i];
    const double irho = 1.0 / rho;
    // equation of state (pressure(), new sound speed) at the half-step
    // density: pointwise, so the soil stage can read the new water pressure
    // without waiting for the water stage
    const double rr = rho / (K.w_rho0 * vof);
    const double p = K.w_b * (pow(rr, K.w_gamma) - 1.0);
    const double cn = K.w_c * pow(rr, 3.0);
    P.p[i] = p;
    P.c_next[i] = cn;
    if (q) {
        float4* Q = reinterpret_cast<float4*>(q) + t;
        const size_t pn = (size_t)P.n;
        write_q01(Q, pn, g, cs, x, y, P.hsml[i], it < 0, vx, vy, rho - K.w_rho0, m, record_cell(NR, t));
        Q[2 * pn] = make_float4((float)vof, (float)irho, (float)(m * irho), (float)p);
        Q[3 * pn] = make_float4((float)cn, 0.0f, 0.0f, 0.0f);
        return;
    }
    double2* R = reinterpret_cast<double2*>(rec) + t;
    const size_t pn = (size_t)P.n;
    R[0 * pn] = make_double2(x, y);
    R[1 * pn] = make_double2(vx, vy);
    R[2 * pn] = make_double2(rho, it < 0 ? -m : m);   // the sign of m marks a virtual particle
    R[3 * pn] = make_double2(vof, p);
    R[4 * pn] = make_double2(0.0, cn);
    // (h and itype stay between list rebuilds, in the same slot: a lean pack
    // finds chunk 6 as the call's first pack or the rebuild's wrote it)
    if (!lean) R[6 * pn] = make_double2(P.hsml[i], (double)it);
}

__device__ __forceinline__ void pack_water_body(int blk, DPhase P, KParams K, double* rec, int do_half, float* q,
                                                Grid g, double cs, NRef NR)
{
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= P.n) return;
    const int i = P.order[t];
    pack_water_at(P, K, rec, t, i, do_half, false, false, q, g, cs, NR, P.x[i], P.y[i], P.rho[i], P.vof[i], P.vx[i], P.vy[i]);
}

__global__ void k_pack_water(DPhase P, KParams K, double* rec, int do_half, float* q, Grid g, double cs, NRef NR)
{
    pack_water_body(blockIdx.x, P, K, rec, do_half, q, g, cs, NR);
}

// The final update of step s fused with the pack of step s + 1 (first_half,
// full step): one pass over the state instead of two, with the same values
// in memory afterwards as k_final_water followed by k_pack_water(do_half =
// 1), apart from stores the pack would overwrite and the half-step levels
// the pack leaves out (lean with FP64 records, lean_v always).  The record
// order of step s + 1 must be the current one (neighbour lists kept, no
// cell-list rebuild), and step s + 1 must follow at once: its final update
// is the next reader of the state.
__device__ __forceinline__ void step_water(int blk, DPhase P, KParams K, double* rec, int first, int with_vof,
                                           int renorm, float* q, Grid g, double cs, NRef NR)
{
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= P.n) return;
    const int i = P.order[t];
    const WaterEnd e = water_end(P, K, i, first, with_vof, renorm);
    const int it = P.itype[i];
    // the final update's stores that the pack does not overwrite
    if (K.summation_density && renorm) P.rho[i] = e.rho;
    if (with_vof && !(K.volume_fraction && K.two_phase)) P.vof[i] = e.vof;
    if (it >= 0) {
        P.x[i] = e.x;
        P.y[i] = e.y;
        if (it == 0) {
            P.vx[i] = e.vx;
            P.vy[i] = e.vy;
        }
    }
    pack_water_at(P, K, rec, t, i, 1, q == nullptr, true, q, g, cs, NR, e.x, e.y, e.rho, e.vof, e.vx, e.vy);
}

__global__ void k_step_water(DPhase P, KParams K, double* rec, int first, int with_vof, int renorm, float* q, Grid g,
                             double cs, NRef NR)
{
    step_water(blockIdx.x, P, K, rec, first, with_vof, renorm, q, g, cs, NR);
}

// Pack of soil particle i (values as for water, plus stress and pressure):
// first_half (do_half), ndp Drucker-Prager projections (the last one
// single_step's failure criterion), the record.  lean as for water: the
// soil density stays (the final update's vof reads it), and the stress
// levels stay unless the final update starts from the saved ones
// (stress_integration = 2).
__device__ __forceinline__ void pack_soil_at(const DPhase& P, const KParams& K, double* rec, int t, int i,
                                             int do_half, bool lean, bool lean_v, int ndp, float* q, const Grid& g,
                                             double cs, const NRef& NR, double x, double y, double rho, double vx,
                                             double vy, double p, double sxx, double sxy, double syy)
{
#pragma clang fp contract(off)
    track_displacement(NR, t, x, y, vx, vy);
    const int it = P.itype[i];
    if (do_half) {
        const double h = K.dt / 2.0;
        if (!K.summation_density) {   // first_half's rho update (declarations_sph.f90:647-650)
            P.rho_min[i] = rho;
            rho = rho + h * P.drho[i];
            P.rho[i] = rho;
        }
        if (it > 0) {
            P.vx_min[i] = vx;
            P.vy_min[i] = vy;
            vx = vx + h * P.ax[i];
            vy = vy + h * P.ay[i];
            if (!lean_v) {
                P.vx[i] = vx;
                P.vy[i] = vy;
            }
        }
    }
    const bool keep_stress = !lean || K.stress_integration != 2;
    if (do_half && K.stress_integration == 2) {   // first_half, declarations_sph.f90:664-672
        const double h = K.dt / 2.0;
        P.sxx_min[i] = sxx; P.sxy_min[i] = sxy; P.syy_min[i] = syy; P.p_min[i] = p;
        sxx = sxx + h * P.dsxx[i];
        sxy = sxy + h * P.dsxy[i];
        syy = syy + h * P.dsyy[i];
        p = p + h * P.dp[i];
        if (keep_stress) {
            P.p[i] = p; P.sxx[i] = sxx; P.sxy[i] = sxy; P.syy[i] = syy;
        }
    }
    // the last projection is single_step's failure criterion: Mohr-Coulomb
    // when yield_criterion = 1, which may also reset rho (the record then
    // keeps the density before it in chunk 7 for the terms evaluated before
    // the criterion: con_density's delta_sph, strain rate, divergence)
    const double rho_pre = rho;
    if (ndp > 0) {
        int f = 0;
        for (int k = 0; k < ndp; ++k) {
            if (k == ndp - 1 && K.yield_criterion == 1) f = mc_project(K, p, sxx, sxy, syy, rho);
            else f = dp_project(K.s_alpha1, p, sxx, sxy, syy);
        }
        if (keep_stress) {
            P.p[i] = p; P.sxx[i] = sxx; P.sxy[i] = sxy; P.syy[i] = syy;
        }
        P.fail[i] = f;
        if (rho != rho_pre) P.rho[i] = rho;
    }
    const double m = P.mass[i];
    const double irho = 1.0 / rho;
    if (q) {
        float4* Q = reinterpret_cast<float4*>(q) + t;
        const size_t pn = (size_t)P.n;
        write_q01(Q, pn, g, cs, x, y, P.hsml[i], it < 0, vx, vy, rho - K.s_rho0, m, record_cell(NR, t));
        Q[2 * pn] = make_float4((float)P.vof[i], (float)irho, (float)(m * irho), (float)p);
        Q[3 * pn] = make_float4((float)P.c[i], (float)sxx, (float)sxy, (float)syy);
        return;
    }
    double2* R = reinterpret_cast<double2*>(rec) + t;
    const size_t pn = (size_t)P.n;
    R[0 * pn] = make_double2(x, y);
    R[1 * pn] = make_double2(vx, vy);
    R[2 * pn] = make_double2(rho, it < 0 ? -m : m);   // the sign of m marks a virtual particle
    R[3 * pn] = make_double2(P.vof[i], p);
    R[4 * pn] = make_double2(syy, P.c[i]);
    R[5 * pn] = make_double2(sxx, sxy);
    if (!lean) R[6 * pn] = make_double2(P.hsml[i], (double)it);   // (as for water)
    if (K.yield_criterion == 1) R[7 * pn] = make_double2(rho_pre, 0.0);
}

__device__ __forceinline__ void pack_soil_body(int blk, DPhase P, KParams K, double* rec, int do_half, int ndp,
                                               float* q, Grid g, double cs, NRef NR)
{
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= P.n) return;
    const int i = P.order[t];
    pack_soil_at(P, K, rec, t, i, do_half, false, false, ndp, q, g, cs, NR, P.x[i], P.y[i], P.rho[i], P.vx[i], P.vy[i], P.p[i],
                 P.sxx[i], P.sxy[i], P.syy[i]);
}

__global__ void k_pack_soil(DPhase P, KParams K, double* rec, int do_half, int ndp, float* q, Grid g, double cs,
                            NRef NR)
{
    pack_soil_body(blockIdx.x, P, K, rec, do_half, ndp, q, g, cs, NR);
}

// The soil final update of step s fused with its pack for step s + 1
// (first_half, two projections), as for water.  The pack reads vof from
// the state: the final update's vof is stored first.
__device__ __forceinline__ void step_soil(int blk, DPhase P, KParams K, double* rec, int first, float* q, Grid g,
                                          double cs, NRef NR)
{
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= P.n) return;
    const int i = P.order[t];
    const SoilEnd e = soil_end(P, K, i, first);
    if (K.volume_fraction && K.two_phase) P.vof[i] = e.vof;
    if (P.itype[i] > 0) {
        P.x[i] = e.x;
        P.y[i] = e.y;
    }
    pack_soil_at(P, K, rec, t, i, 1, q == nullptr, true, 2, q, g, cs, NR, e.x, e.y, e.rho, e.vx, e.vy, e.p, e.sxx, e.sxy, e.syy);
}

__global__ void k_step_soil(DPhase P, KParams K, double* rec, int first, float* q, Grid g, double cs, NRef NR)
{
    step_soil(blockIdx.x, P, K, rec, first, q, g, cs, NR);
}

// ------------------------------------------------------------------------
// Partner masks of stage A, reused by stage B: per particle t a 128-bit mask
// over its candidate list (the three cell rows of its 3 x 3 block, in cell
// order, concatenated): m[t] low and m[n + t] high word, and an overflow
// flag (rows walked directly).
// ------------------------------------------------------------------------
struct Masks {
    uint64_t* m;   // [3][n] (the first two words are used)
    int* ovf;      // [n]
};

// ------------------------------------------------------------------------
// Neighbour lists with a skin (FP64 fused step).  At a rebuild the cell
// list is built on a grid of cell size >= R + skin (R the largest support
// radius), the records are written in its order, and every particle keeps
// (a) the key of its cell, (b) a 128-bit mask over its candidate rows of the
// candidates within R + skin (same phase and cross), and (c) its position.
// The record order and the cell list then stay fixed until the next rebuild.
// While no particle has moved more than skin / 2 from its rebuild position,
// every pair within R lies in these masks, so each step tests only the
// listed candidates with the exact predicate instead of scanning the rows.
// ------------------------------------------------------------------------
struct NList {
    Masks same, cross;   // candidates within R + skin: same phase, other phase
    const int* cell;     // [n] key of the particle's cell at the rebuild (record order)
};

// ------------------------------------------------------------------------
// Partner iteration.  A particle's candidates are the three cell rows of
// its 3 x 3 block, each a contiguous range [k0, k1) of the cell-sorted
// records.  A scan tests every candidate with the exact predicate and keeps
// one 64-bit mask per row; the three masks are concatenated into one
// 128-bit mask over row0 ++ row1 ++ row2, and the dense loops visit its set
// bits.  A row over 64 candidates, or more than 128 in all, sets the
// overflow flag and that particle walks its rows directly.
// ------------------------------------------------------------------------
struct LaneRows {
    int k0[3], k1[3];
};

__device__ __forceinline__ LaneRows lane_rows(const Rows& R)
{
    LaneRows L;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        L.k0[r] = R.k0[r];
        L.k1[r] = R.k1[r];
    }
    return L;
}

// Scan of one candidate row: accept(k, xy) on the candidates' positions.
template <class P>
__device__ __forceinline__ uint64_t scan_row(int k0, int k1, int self, const RecV& rec, P&& accept)
{
    uint64_t m = 0;
    if (k1 - k0 > 64) return 0;
#pragma unroll 4
    for (int k = k0; k < k1; ++k) {
        const bool ok = k != self && accept(k, rec(k, 0));
        m |= (uint64_t)ok << (k - k0);
    }
    return m;
}

__device__ __forceinline__ int ovf_of(const LaneRows& L)
{
    int o = 0;
#pragma unroll
    for (int r = 0; r < 3; ++r)
        if (L.k1[r] - L.k0[r] > 64) o |= 1 << r;
    return o;
}

// Dense visit of the masked partners, then the direct walk of overflowed rows.
template <class IT, class LD, class B, class P>
__device__ __forceinline__ void visit(IT it, const LaneRows& L, int ovf, int self, const LD& load, P&& accept, B&& body)
{
#pragma unroll 1
    for (int k = it.next(); k >= 0; k = it.next()) body(load(k));
    if (ovf) {
        // one pass over the concatenated rows (row bounds as scalars: an
        // indexed LaneRows would be kept in scratch memory)
        const int s1 = L.k1[0] - L.k0[0], s2 = s1 + L.k1[1] - L.k0[1], tot = s2 + L.k1[2] - L.k0[2];
        const int e0 = L.k0[0], e1 = L.k0[1] - s1, e2 = L.k0[2] - s2;
#pragma unroll 1
        for (int c = 0; c < tot; ++c) {
            const int r = c < s1 ? 0 : (c < s2 ? 1 : 2);
            if (!(ovf & (1 << r))) continue;
            const int kk = c + (r == 0 ? e0 : (r == 1 ? e1 : e2));
            if (kk == self || !accept(kk)) continue;
            body(load(kk));
        }
    }
}

// The three row masks of a lane concatenated into one 128-bit mask over
// the candidate list row0 ++ row1 ++ row2 (at most 128 candidates); a lane
// with more, or with a row over 64, walks its rows directly (ovf = 7).
struct CatMask {
    uint64_t lo, hi;
    int ovf;
};

__device__ __forceinline__ CatMask cat_masks(const LaneRows& L, uint64_t m0, uint64_t m1, uint64_t m2)
{
    const int l0 = L.k1[0] - L.k0[0], l1 = L.k1[1] - L.k0[1], l2 = L.k1[2] - L.k0[2];
    CatMask c;
    c.ovf = (ovf_of(L) || l0 + l1 + l2 > 128) ? 7 : 0;
    if (c.ovf) {
        c.lo = c.hi = 0;
        return c;
    }
    unsigned __int128 v = (unsigned __int128)m2;
    v = (v << l1) | m1;
    v = (v << l0) | m0;
    c.lo = (uint64_t)v;
    c.hi = (uint64_t)(v >> 64);
    return c;
}

// Value select.  A conditional expression over two lvalues (struct fields)
// is itself an lvalue: clang then selects between their addresses and loads
// through the result, which keeps the struct on the stack and puts a
// scratch load on every partner iteration.  Arguments taken by value make
// it a register select.
template <class T>
__device__ __forceinline__ T pick(bool c, T x, T y) { return c ? x : y; }

// Iterator over the set bits of a CatMask; branch-free apart from the loop.
// next() returns the record index of the next partner (-1 at the end) and
// i = the index in the staged copy (record index + base of its row).
struct CatIter {
    uint64_t lo, hi;
    int s1, s2;        // start of row 1 and row 2 in the concatenated list
    int e0, e1, e2;    // record index - list index, per row
    __device__ __forceinline__ int next()
    {
        const bool in_lo = lo != 0;
        const uint64_t m = pick(in_lo, lo, hi);
        if (m == 0) return -1;
        const int c = __builtin_ctzll(m) + (in_lo ? 0 : 64);
        const uint64_t rest = m & (m - 1);
        lo = pick(in_lo, rest, lo);
        hi = pick(in_lo, hi, rest);
        const bool a = c < s1, b = c < s2;
        return c + pick(a, e0, pick(b, e1, e2));
    }
};

__device__ __forceinline__ CatIter cat_iter(const CatMask& c, const LaneRows& L)
{
    CatIter it;
    it.lo = c.lo;
    it.hi = c.hi;
    const int l0 = L.k1[0] - L.k0[0], l1 = L.k1[1] - L.k0[1];
    it.s1 = l0;
    it.s2 = l0 + l1;
    it.e0 = L.k0[0];
    it.e1 = L.k0[1] - l0;
    it.e2 = L.k0[2] - l0 - l1;
    return it;
}

// Partner data of the stage kernels: xy = (x, y), v = (vx, vy),
// rm = (rho, m), hv = (h, vof), pc = (p, c), s12 = (s1, s2), s3 = (s3, itype),
// iv = (1/rho, m/rho), filled from the record chunks.
struct Nb {
    double2 xy, v, rm, hv, pc, s12, s3, iv;
    double rpre;   // density before the Mohr-Coulomb criterion (record chunk 7)
    __device__ __forceinline__ void set(int c, double2 q)
    {
        switch (c) {
        case 0: xy = q; break;
        case 1: v = q; break;
        case 2: rm = q; break;
        case 3: hv.y = q.x; pc.x = q.y; break;   // vof, p
        case 4: s3.x = q.x; pc.y = q.y; break;   // s3, c
        case 5: s12 = q; break;
        case 6: hv.x = q.x; s3.y = q.y; break;   // h, itype
        default: break;
        }
    }
};

constexpr int NEED_IV = 1 << 8;   // recompute 1/rho and m/rho of the partner
constexpr int CH_H = 1 << 6;      // chunk holding h

// Loads the chunks NEED of partner record k -- plus the h chunk when with_h
// (non-uniform smoothing length); FROM = 1: all but the positions (chunk 0).
template <int NEED, int FROM = 0>
__device__ __forceinline__ Nb load_nb(const RecV& rec, int k, bool with_h)
{
    Nb b{};
#pragma unroll
    for (int c = FROM; c < 7; ++c) {
        if (!(NEED & (1 << c)) && !(c == 6 && with_h)) continue;
        b.set(c, rec(k, c));
    }
    if (NEED & 4) {
        // chunk 2 carries m with the sign of itype (virtual < 0)
        if (!((NEED & CH_H) || with_h)) b.s3.y = b.rm.y < 0.0 ? -1.0 : 1.0;
        b.rm.y = fabs(b.rm.y);
    }
    if (NEED & NEED_IV) {
        b.iv.x = fast_rcp(b.rm.x);
        b.iv.y = b.rm.y * b.iv.x;
    }
    return b;
}

// Partner loader of a dense loop: the chunks NEED of record k (chunk 7, the
// density before the Mohr-Coulomb criterion, when mc), whole or as positions
// plus the rest.
template <int NEED>
struct Loader {
    RecV rec;
    bool with_h, mc;
    __device__ __forceinline__ double2 xy(int k) const { return rec(k, 0); }
    __device__ __forceinline__ Nb rest(int k, double2 p) const
    {
        Nb b = load_nb<NEED, 1>(rec, k, with_h);
        b.xy = p;
        if (mc) b.rpre = rec(k, 7).x;
        return b;
    }
    __device__ __forceinline__ Nb operator()(int k) const
    {
        Nb b = load_nb<NEED>(rec, k, with_h);
        if (mc) b.rpre = rec(k, 7).x;
        return b;
    }
};

// Exact pair predicate on a candidate.
struct Pred {
    double xa, ya, ha, thr;
    int uni, link;
    __device__ __forceinline__ bool operator()(double2 b, double hb, const Grid& g, const KParams& K) const
    {
        const double dx = xa - b.x, dy = ya - b.y;
        double dr;
        {
#pragma clang fp contract(off)
            dr = dx * dx + dy * dy;
        }
        if (!pred(uni != 0, thr, dr, ha, hb, K)) return false;
        return !link || ll_reach(g, xa, ya, b.x, b.y);
    }
};

// Wave-level set-up shared by the stage kernels.
struct WaveCtx {
    int lane, t, tt, last;
    bool valid;
};

// blk: the workgroup's block of the phase (the XCD-aware block of a solo
// launch, the tile's block of a batched one).
__device__ __forceinline__ WaveCtx wave_ctx(int n, int blk)
{
    WaveCtx c;
    c.lane = threadIdx.x & 63;
    const int t0 = (blk * SPHG_WPB + (int)(threadIdx.x >> 6)) * 64;
    c.t = t0 + c.lane;
    c.valid = c.t < n;
    c.last = min(63, n - 1 - t0);
    c.tt = c.valid ? c.t : n - 1;
    return c;
}

__device__ __forceinline__ WaveCtx wave_ctx(int n) { return wave_ctx(n, xcd_block()); }

// Mask of a particle over its candidate rows in the records rec: scan of the
// three rows with the predicate P (self: own record index, or -1).
template <class A>
__device__ __forceinline__ CatMask scan_mask(const LaneRows& L, int self, const RecV& rec, A&& acc)
{
    const uint64_t m0 = scan_row(L.k0[0], L.k1[0], self, rec, acc);
    const uint64_t m1 = scan_row(L.k0[1], L.k1[1], self, rec, acc);
    const uint64_t m2 = scan_row(L.k0[2], L.k1[2], self, rec, acc);
    return cat_masks(L, m0, m1, m2);
}

// The exact partner mask from a neighbour-list mask: the listed candidates
// (bits of vm over the rows L) whose current positions pass accept(k, xy).
// Four candidates' positions are loaded back to back.  A list that
// overflowed (ovf) stays overflowed: the visit then walks the rows with the
// exact predicate.
template <class A>
__device__ __forceinline__ CatMask filter_mask(const LaneRows& L, const CatMask& vm, const RecV& rec, A&& acc)
{
    if (vm.ovf) return vm;
    const int s1 = L.k1[0] - L.k0[0], s2 = s1 + L.k1[1] - L.k0[1];
    const int e0 = L.k0[0], e1 = L.k0[1] - s1, e2 = L.k0[2] - s2;
    uint64_t lo = vm.lo, hi = vm.hi, olo = 0, ohi = 0;
#pragma unroll 1
    for (;;) {
        int c[4], k[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {   // next listed bit: position c, record index k (-1: none)
            const bool in_lo = lo != 0;
            const uint64_t m = in_lo ? lo : hi;
            c[j] = m ? __builtin_ctzll(m) + (in_lo ? 0 : 64) : -1;
            const uint64_t rest = m & (m - 1);
            lo = in_lo ? rest : lo;
            hi = in_lo ? hi : rest;
            k[j] = c[j] < 0 ? -1 : c[j] + (c[j] < s1 ? e0 : (c[j] < s2 ? e1 : e2));
        }
        if (k[0] < 0) break;
        double2 p[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) p[j] = rec(k[j] >= 0 ? k[j] : k[0], 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (k[j] >= 0 && acc(k[j], p[j])) {
                if (c[j] < 64) olo |= 1ull << c[j];
                else ohi |= 1ull << (c[j] - 64);
            }
        }
        if (k[3] < 0) break;
    }
    CatMask out;
    out.lo = olo;
    out.hi = ohi;
    out.ovf = 0;
    return out;
}

__device__ __forceinline__ CatMask mask_of(const Masks& M, int n, int t)
{
    CatMask c;
    c.lo = M.m[t];
    c.hi = M.m[n + t];
    c.ovf = M.ovf[t];
    return c;
}

// Candidate rows of particle tt: the 3 x 3 block of its cell at the list
// rebuild (row-major key NL.cell[tt]), or (no lists: NL.cell == nullptr) of
// its current cell.
__device__ __forceinline__ LaneRows cand_rows(const Grid& g, const int* cell_start, const NList& NL, int tt,
                                              double xa, double ya)
{
    int cx, cy;
    if (NL.cell) {
        const int key = NL.cell[tt];
        cy = key / g.ncx;
        cx = key - cy * g.ncx;
    } else {
        clamped_cell(g, xa, ya, cx, cy);
    }
    return lane_rows(block_rows_at(g, cell_start, cx, cy));
}

// Partner mask of particle tt over the rows L: the list M filtered by the
// exact predicate, or (no lists: M.m == nullptr) the scan of the rows.
template <class A>
__device__ __forceinline__ CatMask cand_mask(const LaneRows& L, const Masks& M, int n, int tt, int self,
                                             const RecV& rec, A&& acc)
{
    if (M.m) return filter_mask(L, mask_of(M, n, tt), rec, acc);
    return scan_mask(L, self, rec, acc);
}

// Neighbour lists of phase A against phase B (cell list startB on grid g,
// candidate k at state slot ordB[k], or k when ordB is null: the state in
// cell order, as a rebuild leaves it): candidates with dr < thr (thr =
// (R + skin)^2 of the pair type), the particle itself excluded when same.
// Also records each particle's cell key.  Positions come from the state, so
// the FP64 and the FP32 records share the lists.
__device__ __forceinline__ void nlist_body(int blk, const double* xA, const double* yA, const int* ordA, int nA,
                                           const double* xB, const double* yB, const int* ordB, const int* startB,
                                           Grid g, double thr, int same, Masks VM, int* cell)
{
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= nA) return;
    const int ia = ordA ? ordA[t] : t;
    const double px = xA[ia], py = yA[ia];
    int cx, cy;
    clamped_cell(g, px, py, cx, cy);
    if (cell) cell[t] = cy * g.ncx + cx;
    const LaneRows L = lane_rows(block_rows_at(g, startB, cx, cy));
    auto acc = [&](int k) {
        const int ib = ordB ? ordB[k] : k;
        const double dx = px - xB[ib], dy = py - yB[ib];
        return dx * dx + dy * dy < thr;
    };
    const int self = same ? t : -1;
    uint64_t m[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int k0 = r == 0 ? L.k0[0] : (r == 1 ? L.k0[1] : L.k0[2]);
        const int k1 = r == 0 ? L.k1[0] : (r == 1 ? L.k1[1] : L.k1[2]);
        uint64_t mm = 0;
        if (k1 - k0 <= 64) {
#pragma unroll 4
            for (int k = k0; k < k1; ++k) mm |= (uint64_t)(k != self && acc(k)) << (k - k0);
        }
        m[r] = mm;
    }
    const CatMask cm = cat_masks(L, m[0], m[1], m[2]);
    VM.m[t] = cm.lo;
    VM.m[nA + t] = cm.hi;
    VM.ovf[t] = cm.ovf;
}

__global__ __launch_bounds__(256) void k_nlist(const double* xA, const double* yA, const int* ordA, int nA,
                                               const double* xB, const double* yB, const int* ordB, const int* startB,
                                               Grid g, double thr, int same, Masks VM, int* cell)
{
    nlist_body(xcd_block(), xA, yA, ordA, nA, xB, yB, ordB, startB, g, thr, same, VM, cell);
}

// Chunks read per partner by each loop (h only when it is not uniform).
constexpr int NEED_SAME_A = 0x07 | NEED_IV;          // x,y v rho,m(itype in its sign); 1/rho, m/rho
constexpr int NEED_WX_A = 0x0F | NEED_IV;            // soil partners of water: x,y v rho,m vof
constexpr int NEED_SX_A = 0x0F | NEED_IV;            // water partners of soil: x,y v rho,m vof,p
constexpr int NEED_B = 0x3F | NEED_IV;               // x,y v rho,m vof,p s3,c s1,s2

// ------------------------------------------------------------------------
// Cross terms, water side (stage A of single_step's interaction block,
// declarations_sph.f90:1380-1410): darcy_law's drag on the water,
// volume_fraction_water2's dvof and volume_fraction_water's vof2 from the
// soil partners (direct_find_2's pairs, virtual x virtual and r = 0
// included).  Its own kernel: the loop needs few registers, so it runs at
// twice the occupancy of the same-phase stage.
// ------------------------------------------------------------------------
template <int SKF, bool U>
__device__ __forceinline__ void water_x4(int blk, DPhase W, DPhase S, const double* recW, const double* recS,
                                         NList NL, Grid g, KParams K, FusedK F)
{
    const RecV RW = recv(recW, W.n), RS = recv(recS, S.n);
    const WaveCtx wc = wave_ctx(W.n, blk);
    const double2 xy = RW(wc.tt, 0), va = RW(wc.tt, 1), rm = RW(wc.tt, 2);
    const double xa = xy.x, ya = xy.y, vxa = va.x, vya = va.y, rhoa = rm.x, vofa = RW(wc.tt, 3).x;
    const double ha = U ? K.h1_water : RW(wc.tt, 6).x;
    const double irhoa = 1.0 / rhoa;
    const int a = W.order[wc.tt];
    const bool ux = U || F.uni_x;
    const LaneRows L = cand_rows(g, S.cell_start, NL, wc.tt, xa, ya);
    const Pred PX{xa, ya, ha, F.thr_x, ux ? 1 : 0, 0};
    auto acc = [&](int k, double2 b) { return PX(b, ux ? ha : RS(k, 6).x, g, K); };
    auto acc_k = [&](int k) { return acc(k, RS(k, 0)); };
    const CatMask cm = cand_mask(L, NL.cross, W.n, wc.tt, -1, RS, acc);
    const Loader<NEED_WX_A> load{RS, !ux, false};
    const double cfa = vofa * rhoa * (-K.numeric_gravity) / K.s_perm;
    const KC kx = make_kc(SKF, sph_mhsml(ha, ux ? RS(0, 6).x : ha));
    double dg_x = 0.0, dg_y = 0.0, dvof_x = 0.0, vof2 = 0.0;
    int cntx = 0;
    visit(cat_iter(cm, L), L, cm.ovf, -1, load, acc_k, [&](const Nb& b) {
        const double dx = xa - b.xy.x, dy = ya - b.xy.y;
        const double dr = dx * dx + dy * dy;
        const double ir = (dr > 0.0) ? fast_rsq(dr) : 0.0;
        const double r = dr * ir;
        double ww, gx, gy;
        if (ux) kern_fast<SKF, true>(kx, r, ir, dx, dy, ww, gx, gy);
        else kern_fast<SKF, true>(make_kc(SKF, sph_mhsml(ha, b.hv.x)), r, ir, dx, dy, ww, gx, gy);
        ++cntx;
        const double mb = b.rm.y;
        const double dvx = vxa - b.v.x, dvy = vya - b.v.y;
        const double sc = mb * (cfa * b.hv.y) * (ww * irhoa * b.iv.x);
        dg_x += sc * dvx;
        dg_y += sc * dvy;
        dvof_x += mb * (dvx * gx + dvy * gy);
        vof2 += mb * ww;
    });
    if (!wc.valid) return;
    W.drgx[a] = dg_x;
    W.drgy[a] = dg_y;
    W.dvof_x[a] = dvof_x * (1.0 / K.s_rho0);
    W.vof2[a] = 1.0 - vof2 / K.s_rho0;
    W.countiac[a] = cntx;
}

template <int SKF, bool U>
__global__ STAGE_ATTR void k_water_x4(DPhase W, DPhase S, const double* recW, const double* recS, NList NL, Grid g,
                                      KParams K, FusedK F)
{
    water_x4<SKF, U>(xcd_block(), W, S, recW, recS, NL, g, K, F);
}

// ------------------------------------------------------------------------
// Cross terms, soil side: darcy_law's drag on the soil and
// pore_water_pressure (m_particles.f90:3832-3909) from the water partners
// (whose pressure is the new one the water pack evaluated).
// ------------------------------------------------------------------------
template <int SKF, bool U>
__device__ __forceinline__ void soil_x4(int blk, DPhase S, DPhase W, const double* recS, const double* recW, NList NL,
                                        Grid g, KParams K, FusedK F)
{
    const RecV RS = recv(recS, S.n), RW = recv(recW, W.n);
    const WaveCtx wc = wave_ctx(S.n, blk);
    const double2 xy = RS(wc.tt, 0), va = RS(wc.tt, 1), rm = RS(wc.tt, 2);
    const double xa = xy.x, ya = xy.y, vxa = va.x, vya = va.y, rhoa = rm.x, vofa = RS(wc.tt, 3).x;
    const double ha = U ? K.h1_soil : RS(wc.tt, 6).x;
    const double irhoa = 1.0 / rhoa;
    const int a = S.order[wc.tt];
    const bool ux = U || F.uni_x;
    const LaneRows L = cand_rows(g, W.cell_start, NL, wc.tt, xa, ya);
    const Pred PX{xa, ya, ha, F.thr_x, ux ? 1 : 0, 0};
    auto acc = [&](int k, double2 b) { return PX(b, ux ? ha : RW(k, 6).x, g, K); };
    auto acc_k = [&](int k) { return acc(k, RW(k, 0)); };
    const CatMask cm = cand_mask(L, NL.cross, S.n, wc.tt, -1, RW, acc);
    const Loader<NEED_SX_A> load{RW, !ux, false};
    const double cfs = vofa * (-K.numeric_gravity) / K.s_perm;
    const KC kx = make_kc(SKF, sph_mhsml(ux ? RW(0, 6).x : ha, ha));
    double dg_x = 0.0, dg_y = 0.0, po_x = 0.0, po_y = 0.0;
    int cntx = 0;
    visit(cat_iter(cm, L), L, cm.ovf, -1, load, acc_k, [&](const Nb& b) {
        const double dx = b.xy.x - xa, dy = b.xy.y - ya;   // water - soil
        const double dr = dx * dx + dy * dy;
        const double ir = (dr > 0.0) ? fast_rsq(dr) : 0.0;
        const double r = dr * ir;
        double ww, gx, gy;
        if (ux) kern_fast<SKF, true>(kx, r, ir, dx, dy, ww, gx, gy);
        else kern_fast<SKF, true>(make_kc(SKF, sph_mhsml(b.hv.x, ha)), r, ir, dx, dy, ww, gx, gy);
        ++cntx;
        const double mw = b.rm.y, rhow = b.rm.x;
        const double sc = mw * (b.hv.y * rhow * cfs) * (ww * b.iv.x * irhoa);
        dg_x += sc * (b.v.x - vxa);
        dg_y += sc * (b.v.y - vya);
        const double mprr = mw * b.pc.x * (b.iv.x * irhoa);
        po_x += mprr * gx;
        po_y += mprr * gy;
    });
    if (!wc.valid) return;
    S.drgx[a] = dg_x;
    S.drgy[a] = dg_y;
    S.porx[a] = vofa * po_x;
    S.pory[a] = vofa * po_y;
    S.countiac[a] = cntx;
}

template <int SKF, bool U>
__global__ STAGE_ATTR void k_soil_x4(DPhase S, DPhase W, const double* recS, const double* recW, NList NL, Grid g,
                                     KParams K, FusedK F)
{
    soil_x4<SKF, U>(xcd_block(), S, W, recS, recW, NL, g, K, F);
}

// ------------------------------------------------------------------------
// Stage A, water: same-phase terms (con_density, delta_sph(rho),
// shear_strain_rate, av_vel, repulsive_force), then pressure() per particle.
// Writes a-space outputs, the same-phase masks and the record slots
// vof*sigma for stage B.  cross: two-phase run (countiac then counts the
// soil partners, which k_water_x4 writes).
// ------------------------------------------------------------------------
template <int SKF, bool U>
__device__ __forceinline__ void water_a4(int blk, DPhase W, const double* recW, double* recWw, NList NL, Masks MW,
                                         Grid g, KParams K, FusedK F, int cross)
{
    const RecV RW = recv(recW, W.n);
    const WaveCtx wc = wave_ctx(W.n, blk);
    const int t = wc.t;
    auto RA = [&](int c) { return RW(wc.tt, c); };
    const double2 xy = RA(0), va = RA(1), rm = RA(2), hi = RA(6);
    const bool link = K.nnps == 2;
    // U (checked by the host): h uniform in the phase (hsml(1) everywhere, a
    // wave-uniform value), integer repulsion exponents
    const double xa = xy.x, ya = xy.y, vxa = va.x, vya = va.y, rhoa = rm.x, vofa = RA(3).x;
    const double ha = U ? K.h1_water : hi.x;
    const int ita = (int)hi.y;
    const int a = W.order[wc.tt];
    const double ca = W.c[a];
    const bool uni = U || link || F.uni_w;
    const KC kc = make_kc(SKF, link ? K.h1_water : ha);
    const LaneRows L = cand_rows(g, W.cell_start, NL, wc.tt, xa, ya);
    const Pred P{xa, ya, ha, F.thr_w, uni ? 1 : 0, link ? 1 : 0};
    auto acc = [&](int k, double2 b) { return P(b, uni ? ha : RW(k, 6).x, g, K); };
    auto acc_k = [&](int k) { return acc(k, RW(k, 0)); };
    const CatMask cm = cand_mask(L, NL.same, W.n, wc.tt, t, RW, acc);
    if (wc.valid) {
        MW.m[t] = cm.lo;
        MW.m[W.n + t] = cm.hi;
        MW.ovf[t] = cm.ovf;
    }
    const Loader<NEED_SAME_A> load{RW, !uni, false};
    double con = 0.0, drd = 0.0, txx = 0.0, txy = 0.0, tyy = 0.0, avx = 0.0, avy = 0.0, repx = 0.0, repy = 0.0;
    int cnt = 0;
    visit(cat_iter(cm, L), L, cm.ovf, t, load, acc_k, [&](const Nb& b) {
        const double dx = xa - b.xy.x, dy = ya - b.xy.y;
        const double dr = dx * dx + dy * dy;
        const Dist D = dist_of(dr);
        const double r = D.r, idr = D.idr;
        double ww, gx, gy;
        if (uni) kern_fast<SKF, true>(kc, r, D.ir, dx, dy, ww, gx, gy);
        else kern_fast<SKF, true>(make_kc(SKF, sph_mhsml(ha, b.hv.x)), r, D.ir, dx, dy, ww, gx, gy);
        ++cnt;
        const double mb = b.rm.y, Vb = b.iv.y;
        const double dvx = vxa - b.v.x, dvy = vya - b.v.y;
        con += mb * (dvx * gx + dvy * gy);
        drd += mb * (2.0 * (rhoa - b.rm.x) * idr * (dx * gx + dy * gy)) * b.iv.x;
        txx += Vb * (-2.0 * dvx * gx + dvy * gy);
        txy += Vb * (-(dvx * gy) - dvy * gx);
        tyy += Vb * (-2.0 * dvy * gy + dvx * gx);
        const double f = mb * ww * (2.0 * fast_rcp(rhoa + b.rm.x));
        avx -= dvx * f;
        avy -= dvy * f;
        if (ita > 0 && !(r > W.dspp) && (int)b.s3.y < 0) {
            const double q = W.dspp / r;
            const double fr = (U ? ipow(q, K.ip1) - ipow(q, K.ip2) : pow(q, K.p1) - pow(q, K.p2)) * idr;
            repx += K.dd * dx * fr;
            repy += K.dd * dy * fr;
        }
    });
    if (!wc.valid) return;
    const double mu = K.w_mu;
    txx *= SPH_C23_F;
    tyy *= SPH_C23_F;
    // delta_sph(rho) and av_vel only under their switches (artificial_density,
    // average_velocity); without av_vel the reference leaves av as it was
    W.drho[a] = K.artificial_density ? con + (K.delta * ha * ca) * drd : con;
    W.txx[a] = txx; W.txy[a] = txy; W.tyy[a] = tyy;
    if (ita > 0) { avx *= K.epsilon; avy *= K.epsilon; }
    if (K.average_velocity) { W.avx[a] = avx; W.avy[a] = avy; }
    W.repx[a] = repx; W.repy[a] = repy;
    if (!cross) W.countiac[a] = cnt;
    W.nsame[a] = cnt;
    // p and the new c (chunks 3 and 4) come from the pack
    double2* WP = reinterpret_cast<double2*>(recWw) + t;
    const size_t nw = (size_t)W.n;
    reinterpret_cast<double*>(WP + 4 * nw)[0] = vofa * (mu * tyy);
    WP[5 * nw] = make_double2(vofa * (mu * txx), vofa * (mu * txy));
}

template <int SKF, bool U>
__global__ STAGE_ATTR void k_water_a4(DPhase W, const double* recW, double* recWw, NList NL, Masks MW, Grid g,
                                      KParams K, FusedK F, int cross)
{
    water_a4<SKF, U>(xcd_block(), W, recW, recWw, NL, MW, g, K, F, cross);
}

// ------------------------------------------------------------------------
// Stage B, water: -vof*df(p) + df(vof*sigma) (declarations_sph.f90:1257-1265),
// art_visc, delta_sph(vof, dvof); /rho, + repulsion, + gravity, - Darcy drag.
// The partner masks come from stage A.
// ------------------------------------------------------------------------
template <int SKF, bool U>
__device__ __forceinline__ void water_b4(int blk, DPhase W, const double* recW, NList NL, Masks MW, Grid g, KParams K,
                                         FusedK F, int mask)
{
    const RecV RW = recv(recW, W.n);
    const WaveCtx wc = wave_ctx(W.n, blk);
    const int t = wc.t;
    auto RA = [&](int c) { return RW(wc.tt, c); };
    const double2 axy = RA(0), av = RA(1), arm = RA(2), a3 = RA(3), apc = RA(4), a12 = RA(5);
    // U: h uniform in the phase (checked by the host)
    const double xa = axy.x, ya = axy.y, vxa = av.x, vya = av.y, rhoa = arm.x, vofa = a3.x;
    const double ha = U ? K.h1_water : RA(6).x;
    const double pa = a3.y, ca = apc.y, fxx_a = a12.x, fxy_a = a12.y, fyy_a = apc.x;
    const double irhoa = 1.0 / rhoa;
    const int a = W.order[wc.tt];
    const bool link = K.nnps == 2;
    const bool uni = U || link || F.uni_w;
    const KC kc = make_kc(SKF, link ? K.h1_water : ha);
    const LaneRows L = cand_rows(g, W.cell_start, NL, wc.tt, xa, ya);
    const Pred P{xa, ya, ha, F.thr_w, uni ? 1 : 0, link ? 1 : 0};
    auto acc_k = [&](int k) { return P(RW(k, 0), uni ? ha : RW(k, 6).x, g, K); };
    const Loader<NEED_B> load{RW, !uni, false};
    CatMask cm{0, 0, 0};
    if (wc.valid) {
        cm.lo = MW.m[t];
        cm.hi = MW.m[W.n + t];
        cm.ovf = MW.ovf[t];
    }
    const double et2 = K.etq * K.etq;
    double Px = 0.0, Py = 0.0, Sxx = 0.0, Sxyy = 0.0, Sxyx = 0.0, Syy = 0.0;
    double artx = 0.0, arty = 0.0, dvd = 0.0;
    visit(cat_iter(cm, L), L, cm.ovf, t, load, acc_k, [&](const Nb& b) {
        const double dx = xa - b.xy.x, dy = ya - b.xy.y;
        const double dr = dx * dx + dy * dy;
        const Dist D = dist_of(dr);
        const double r = D.r, idr = D.idr;
        double ww, gx, gy;
        const double amh = uni ? ha : sph_mhsml(ha, b.hv.x);
        if (uni) kern_fast<SKF, false>(kc, r, D.ir, dx, dy, ww, gx, gy);
        else kern_fast<SKF, false>(make_kc(SKF, amh), r, D.ir, dx, dy, ww, gx, gy);
        const double Vb = b.iv.y, mb = b.rm.y;
        const double ps = pa + b.pc.x;
        Px += Vb * (ps * gx);
        Py += Vb * (ps * gy);
        const double sxy = fxy_a + b.s12.y;
        Sxx += Vb * ((fxx_a + b.s12.x) * gx);
        Sxyy += Vb * (sxy * gy);
        Sxyx += Vb * (sxy * gx);
        Syy += Vb * ((fyy_a + b.s3.x) * gy);
        const double vr = (vxa - b.v.x) * dx + (vya - b.v.y) * dy;
        {   // art_visc, branch-free: approaching pairs only (vr < 0)
            const double muv = amh * vr * fast_rcp(dr + amh * amh * et2);
            const double mc = 0.5 * (ca + b.pc.y);
            double piv = (K.beta * muv - K.alpha * mc) * muv * (2.0 * fast_rcp(rhoa + b.rm.x));
            piv = vr < 0.0 ? piv : 0.0;
            artx -= mb * piv * gx;
            arty -= mb * piv * gy;
        }
        dvd += mb * (2.0 * (vofa - b.hv.y) * idr * (dx * gx + dy * gy)) * b.iv.x;
    });
    if (!wc.valid) return;
    double ax = -(vofa * Px) + Sxx + Sxyy;
    double ay = -(vofa * Py) + Sxyx + Syy;
    if (rhoa > 0.0) {
        ax = ax * irhoa;
        ay = ay * irhoa;
    }
    if (mask & M_ART) { ax = ax + artx; ay = ay + arty; }
    ax = ax + W.repx[a];
    ay = ay + W.repy[a];
    ay = ay + K.gravity;
    double dvof = (mask & M_DELTA_FIELD) ? (K.delta * ha * ca) * dvd : 0.0;
    if (mask & M_CROSS) {
        ax = ax - W.drgx[a];
        ay = ay - W.drgy[a];
        dvof = dvof - W.dvof_x[a];
    }
    W.ax[a] = ax;
    W.ay[a] = ay;
    W.dvof[a] = dvof;
}

template <int SKF, bool U>
__global__ STAGE_ATTR void k_water_b4(DPhase W, const double* recW, NList NL, Masks MW, Grid g, KParams K, FusedK F,
                                      int mask)
{
    water_b4<SKF, U>(xcd_block(), W, recW, NL, MW, g, K, F, mask);
}

// ------------------------------------------------------------------------
// Soil, stages A and B in one kernel (SOLO: single-phase soil,
// single_step_for_soil, declarations_sph.f90:1530-1655: the df form of the
// momentum without vof factors, no repulsion, no water partners).  Soil stage B reads only what the pack
// wrote (p, sigma, vof of the partners) and the new sound speed, which is the
// constant s_c for every soil particle, so both stages share one scan and one
// dense loop over the same partner set (no stored masks, one partner load):
//   same phase  con_density, delta_sph(rho), shear_strain_rate,
//               velocity_divergence, Jaumann spin, av_vel, repulsive_force
//               (stage A); -df3(vof*p) + df3(vof*sigma) (:1270-1277),
//               art_visc, delta_sph on p, sxx, sxy, syy (stage B);
//   tail        plastic_or_not + plastic_flow_rule3 + Jaumann_rate, then
//               /rho, + repulsion, + gravity, + Darcy drag, + pore pressure
//               (the cross terms k_soil_x4 wrote; cross: two-phase run).
// ------------------------------------------------------------------------
template <int SKF, bool U, bool SOLO>
__device__ __forceinline__ void soil_ab4(int blk, DPhase S, const double* recS, NList NL, Grid g, KParams K, FusedK F,
                                         int cross, int mask)
{
    const RecV RS = recv(recS, S.n);
    const WaveCtx wc = wave_ctx(S.n, blk);
    const int t = wc.t;
    auto RA = [&](int c) { return RS(wc.tt, c); };
    const double2 xy = RA(0), va = RA(1), rm = RA(2), a3 = RA(3), a4 = RA(4), a12 = RA(5), hi = RA(6);
    // U (checked by the host): h uniform in the phase (hsml(1) everywhere, a
    // wave-uniform value), integer repulsion exponents, Drucker-Prager
    const double xa = xy.x, ya = xy.y, vxa = va.x, vya = va.y, rhoa = rm.x, vofa = a3.x;
    const double ha = U ? K.h1_soil : hi.x;
    const double pa = a3.y, sxxa = a12.x, sxya = a12.y, syya = a4.x;
    const double ca = a4.y;          // sound speed before this step (delta_sph of rho)
    const double cn = K.s_c;         // sound speed after it (stage B)
    const double irhoa = 1.0 / rhoa, ir2a = irhoa * irhoa;
    // Mohr-Coulomb (general instantiation only): the criterion may have reset
    // rho; con_density's delta_sph, the strain rate and the divergence come
    // before it in single_step and use the density before it (chunk 7)
    const bool mc = !U && K.yield_criterion == 1;
    const double rhoa_pre = mc ? RA(7).x : rhoa;
    const double fpa = vofa * pa * ir2a, fxxa = vofa * sxxa * ir2a, fxya = vofa * sxya * ir2a,
                 fyya = vofa * syya * ir2a;
    const int ita = (int)hi.y;
    const int a = S.order[wc.tt];
    const bool link = K.nnps == 2;
    const bool uni = U || link || F.uni_s;
    const KC kc = make_kc(SKF, link ? K.h1_soil : ha);
    const LaneRows L = cand_rows(g, S.cell_start, NL, wc.tt, xa, ya);
    const Pred P{xa, ya, ha, F.thr_s, uni ? 1 : 0, link ? 1 : 0};
    auto acc = [&](int k, double2 b) { return P(b, uni ? ha : RS(k, 6).x, g, K); };
    auto acc_k = [&](int k) { return acc(k, RS(k, 0)); };
    const CatMask cm = cand_mask(L, NL.same, S.n, wc.tt, t, RS, acc);
    const Loader<NEED_B> load{RS, !uni, mc};
    const double et2 = K.etq * K.etq;
    double con = 0.0, drd = 0.0, txx = 0.0, txy = 0.0, tyy = 0.0, vcc = 0.0, wxy = 0.0;
    double avx = 0.0, avy = 0.0, repx = 0.0, repy = 0.0;
    double Fx = 0.0, Fy = 0.0;
    double artx = 0.0, arty = 0.0, dP = 0.0, dXX = 0.0, dXY = 0.0, dYY = 0.0;
    int cnt = 0;
    visit(cat_iter(cm, L), L, cm.ovf, t, load, acc_k, [&](const Nb& b) {
        const double dx = xa - b.xy.x, dy = ya - b.xy.y;
        const double dr = dx * dx + dy * dy;
        const Dist D = dist_of(dr);
        const double r = D.r, idr = D.idr;
        double ww, gx, gy;
        const double amh = uni ? ha : sph_mhsml(ha, b.hv.x);
        if (uni) kern_fast<SKF, true>(kc, r, D.ir, dx, dy, ww, gx, gy);
        else kern_fast<SKF, true>(make_kc(SKF, amh), r, D.ir, dx, dy, ww, gx, gy);
        ++cnt;
        const double mb = b.rm.y, Vb = b.iv.y, irb = b.iv.x;
        // stage A
        const double dvx = vxa - b.v.x, dvy = vya - b.v.y;
        const double dvg = dvx * gx + dvy * gy;
        const double rb_pre = mc ? b.rpre : b.rm.x;
        const double irb_pre = mc ? fast_rcp(rb_pre) : irb;
        const double Vb_pre = mc ? mb * irb_pre : Vb;
        con += mb * dvg;
        drd += mb * (2.0 * (rhoa_pre - rb_pre) * idr * (dx * gx + dy * gy)) * irb_pre;
        txx += Vb_pre * (-2.0 * dvx * gx + dvy * gy);
        txy += Vb_pre * (-(dvx * gy) - dvy * gx);
        tyy += Vb_pre * (-2.0 * dvy * gy + dvx * gx);
        vcc -= Vb_pre * dvg;
        wxy += Vb * (-(dvx * gy) + dvy * gx);
        const double f = mb * ww * (2.0 * fast_rcp(rhoa + b.rm.x));
        avx -= dvx * f;
        avy -= dvy * f;
        if (!SOLO && ita > 0 && !(r > S.dspp) && (int)b.s3.y < 0) {
            const double q = S.dspp / r;
            const double fr = (U ? ipow(q, K.ip1) - ipow(q, K.ip2) : pow(q, K.p1) - pow(q, K.p2)) * idr;
            repx += K.dd * dx * fr;
            repy += K.dd * dy * fr;
        }
        // stage B
        const double vofb = b.hv.y, ir2b = irb * irb;
        const double fp = fpa + vofb * b.pc.x * ir2b;
        const double fxx = fxxa + vofb * b.s12.x * ir2b;
        const double fxy = fxya + vofb * b.s12.y * ir2b;
        const double fyy = fyya + vofb * b.s3.x * ir2b;
        if constexpr (SOLO) {
            // -df(p) + df(sigma) (single_step_for_soil :1596-1597):
            // m_b/rho_b (f_a + f_b) grad W, no vof factors
            const double ps = pa + b.pc.x, sxy = sxya + b.s12.y;
            Fx += Vb * ((sxxa + b.s12.x - ps) * gx + sxy * gy);
            Fy += Vb * ((syya + b.s3.x - ps) * gy + sxy * gx);
        } else {
            Fx += mb * ((fxx - fp) * gx + fxy * gy);
            Fy += mb * ((fyy - fp) * gy + fxy * gx);
        }
        const double vr = dvx * dx + dvy * dy;
        {   // art_visc, branch-free: approaching pairs only (vr < 0)
            const double muv = amh * vr * fast_rcp(dr + amh * amh * et2);
            double piv = (K.beta * muv - K.alpha * cn) * muv * (2.0 * fast_rcp(rhoa + b.rm.x));
            piv = vr < 0.0 ? piv : 0.0;
            artx -= mb * piv * gx;
            arty -= mb * piv * gy;
        }
        const double hh = 2.0 * idr * (dx * gx + dy * gy) * mb * irb;
        dP += hh * (pa - b.pc.x);
        dXX += hh * (sxxa - b.s12.x);
        dXY += hh * (sxya - b.s12.y);
        dYY += hh * (syya - b.s3.x);
    });

    if (!wc.valid) return;
    txx *= SPH_C23_F;
    tyy *= SPH_C23_F;
    wxy *= 0.5;
    // delta_sph(rho) and av_vel only under their switches (artificial_density,
    // average_velocity); without av_vel the reference leaves av as it was
    // single-phase soil under summation_density: sum_density replaced rho and
    // drho keeps only the delta-SPH term (declarations_sph.f90:1558-1570)
    if (SOLO && K.summation_density) con = 0.0;
    S.drho[a] = K.artificial_density ? con + (K.delta * ha * ca) * drd : con;
    S.txx[a] = txx; S.txy[a] = txy; S.tyy[a] = tyy; S.vcc[a] = vcc; S.wxy[a] = wxy;
    if (ita > 0) { avx *= K.epsilon; avy *= K.epsilon; }
    if (K.average_velocity) { S.avx[a] = avx; S.avy[a] = avy; }
    S.repx[a] = repx; S.repy[a] = repy;
    if (!cross) S.countiac[a] = cnt;
    S.nsame[a] = cnt;
    S.c_next[a] = cn;

    const double alpha = K.s_alpha1, G = K.s_G, Kb = K.s_k;
    double dsxx = 0.0, dsxy = 0.0, dsyy = 0.0, dpv = 0.0;
    // plasticity 3: plastic_or_not + plastic_flow_rule3 (failed points);
    // plasticity 1: plastic_flow_rule (m_particles.f90:3571-3633), the same
    // correction at every point, with the criterion's fail flags left as they are
    if (K.plasticity == 3 || K.plasticity == 1) {
        const double I1 = 3.0 * pa;
        const double j2r = sxxa * sxxa + 2.0 * (sxya * sxya) + syya * syya + (sxxa + syya) * (sxxa + syya);
        const double J2 = __builtin_sqrt(j2r / 2.0) + 1.0e-6;
        const int fail = fabs(J2 - alpha * I1) < SPH_YIELD_TOL_F ? 1 : 0;
        if (fail || K.plasticity == 1) {
            const double exx = txx * 0.5 + vcc * (1.0 / 3.0);
            const double exy = txy * 0.5;
            const double eyy = tyy * 0.5 + vcc * (1.0 / 3.0);
            const double sde = sxxa * exx + 2.0 * sxya * exy + syya * eyy;
            const double sj = __builtin_sqrt(j2r / 2.0 + 1.0e-10);
            const double gs = G / sj;
            const double dl = (3.0 * alpha * Kb * vcc + gs * sde) / (9.0 * (alpha * alpha) * Kb + G);
            if (!(dl < 0.0)) {
                dsxx = -dl * (gs * sxxa);
                dsxy = -dl * (gs * sxya);
                dsyy = -dl * (gs * syya);
                dpv = 3.0 * Kb * alpha * dl;
                S.eps_p[a] = S.eps_p[a] + dl * sxya / (2.0 * sj) * SPH_EPSP_F;
            }
        }
        if (K.plasticity == 3) S.fail[a] = fail;
    }
    if (K.plasticity == 2) {
        // plastic_flow_rule2 (m_particles.f90:3636-3697): at every point and
        // without the sign test; its rates act in the return mapping
        const double j2r = sxxa * sxxa + 2.0 * (sxya * sxya) + syya * syya + (sxxa + syya) * (sxxa + syya);
        const double exx = txx * 0.5 + vcc * (1.0 / 3.0);
        const double exy = txy * 0.5;
        const double eyy = tyy * 0.5 + vcc * (1.0 / 3.0);
        const double sde = sxxa * exx + 2.0 * sxya * exy + syya * eyy;
        const double sj = __builtin_sqrt(j2r / 2.0 + 1.0e-10);
        const double gs = G / sj;
        const double dl = (3.0 * alpha * Kb * vcc + gs * sde) / (9.0 * (alpha * alpha) * Kb + G);
        S.dsxx2[a] = -dl * (gs * sxxa);
        S.dsxy2[a] = -dl * (gs * sxya);
        S.dsyy2[a] = -dl * (gs * syya);
        S.dp2[a] = 3.0 * Kb * alpha * dl;
    }
    dsxx = dsxx + G * txx + 2.0 * sxya * wxy;
    dsxy = dsxy + G * txy - (sxxa - syya) * wxy;
    dsyy = dsyy + G * tyy - 2.0 * sxya * wxy;
    dpv = dpv - Kb * vcc;

    double ax = SOLO ? Fx : Fx * rhoa;
    double ay = SOLO ? Fy : Fy * rhoa;
    if (rhoa > 0.0) { ax = ax * irhoa; ay = ay * irhoa; }
    if (mask & M_ART) { ax = ax + artx; ay = ay + arty; }
    ax = ax + repx;
    ay = ay + repy;
    ay = ay + K.gravity;
    if (mask & M_CROSS) {
        ax = ax + S.drgx[a]; ay = ay + S.drgy[a];
        ax = ax + S.porx[a]; ay = ay + S.pory[a];
    }
    if (mask & M_DELTA_FIELD) {
        const double cf = K.delta * ha * cn;
        dpv = dpv + cf * dP;
        dsxx = dsxx + cf * dXX;
        dsxy = dsxy + cf * dXY;
        dsyy = dsyy + cf * dYY;
    }
    S.dsxx[a] = dsxx; S.dsxy[a] = dsxy; S.dsyy[a] = dsyy; S.dp[a] = dpv;
    S.ax[a] = ax;
    S.ay[a] = ay;
}

template <int SKF, bool U, bool SOLO>
__global__ STAGE_ATTR_AB void k_soil_ab4(DPhase S, const double* recS, NList NL, Grid g, KParams K, FusedK F,
                                         int cross, int mask)
{
    soil_ab4<SKF, U, SOLO>(xcd_block(), S, recS, NL, g, K, F, cross, mask);
}
