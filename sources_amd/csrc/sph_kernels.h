// sph_kernels.h -- HIP kernels of the per-step particle loop (FP64 path).
//
// Data layout: one struct-of-arrays per phase in HBM, indexed by particle
// number (0-based).  Each step a counting sort builds, per phase, a list of
// particle numbers ordered by grid cell (row-major cells, ascending particle
// number inside a cell).  Gather kernels walk that order, so the 64 lanes of a
// wavefront hold neighbouring particles and their neighbour loads hit the same
// cache lines.  Every pair interaction of the reference (a symmetric
// scatter over the i<j pair list) is evaluated as a gather: particle a sums
// the term the reference adds to it over all partners b.  Because
// x_a - x_b = -(x_b - x_a) exactly and grad W is odd in dx, each per-pair term
// equals the reference's bit for bit; only the summation order differs.
// No atomics are used in the physics passes.
#pragma once
#include "sph_device.h"

struct DPhase {
    int n, ntotal, material;
    double dspp;
    int *id;   // reference particle number (0-based) of each device slot
    int *itype, *fail, *countiac, *key, *slot;
    int *nsame;   // same-phase partners of the last fused step (countiac holds the cross count in two-phase runs)
    double *x, *y, *vx, *vy, *mass, *hsml;
    double *rho, *rho_min, *rho_next;
    double *vof, *vof_min, *vof2, *dvof, *dvof_x;
    double *p, *c, *c_next;
    double *drho, *ax, *ay, *avx, *avy, *vx_min, *vy_min;
    double *txx, *txy, *tyy, *vcc, *wxy;
    double *sxx, *sxy, *syy, *eps_p;
    double *dsxx, *dsxy, *dsyy, *dp;
    double *repx, *repy, *drgx, *drgy, *porx, *pory;
    double *sxx_min, *sxy_min, *syy_min, *p_min;   // stress_integration = 2
    double *dsxx2, *dsxy2, *dsyy2, *dp2;           // plastic_flow_rule2
    int *cell_count, *cell_start, *order;
};

struct KParams {
    int skf, nnps;
    double scale_k;
    double alpha, beta, etq, dd, p1, p2, delta, epsilon, gravity, numeric_gravity;
    double w_rho0, w_b, w_gamma, w_c, w_mu;
    double s_rho0, s_k, s_c, s_G, s_alpha1, s_perm;
    double dt;
    double h1_water, h1_soil;   // hsml(1) of each phase (link_list predicate)
    int artificial_density, visc_artificial, average_velocity;
    int soil_artificial_stress, water_artificial_volume, volume_fraction;
    int plasticity, yield_criterion, two_phase;
    int ip1, ip2;               // p1, p2 when they are integers in [0, 64], else -1
    int stress_integration;     // 1: sigma += dt dsigma; 2: leapfrog levels sigma_min
    int summation_density;      // the leapfrog leaves rho alone (declarations_sph.f90:647-650)
    // Mohr-Coulomb criterion (yield_criterion = 1): cohesion cos(phi), sin(phi)
    // and the tension cut-off pressure -cohesion tan(phi)**(-1.0), host-evaluated
    double mc_ccos, mc_sin, mc_pcut;
};

// Stage masks (granular entry points run a subset of a fused pass).
enum : int {
    M_CON = 1 << 0,         // con_density
    M_DRHO_DELTA = 1 << 1,  // delta_sph(rho, drho)
    M_STRAIN = 1 << 2,      // shear_strain_rate (+ velocity_divergence for soil)
    M_EOS = 1 << 3,         // pressure()
    M_XSPH = 1 << 4,        // av_vel
    M_REP = 1 << 5,         // repulsive_force
    M_CROSS = 1 << 6,       // water/soil interaction sums
    M_PLASTIC = 1 << 7,     // plastic_or_not + plastic_flow_rule3 + Jaumann_rate (soil)
    M_DS_ACCUM = 1 << 8,    // add to the existing dsxx.. / dp instead of starting at zero
    M_DRHO_ACCUM = 1 << 9,  // add to the existing drho instead of starting at zero
    M_MOMENTUM = 1 << 10,   // df / df3 internal force
    M_ART = 1 << 11,        // art_visc
    M_DELTA_FIELD = 1 << 12,// delta_sph on vof (water) / p, s (soil)
    M_GRAV = 1 << 13,       // dvx(2,:) += gravity
    M_DVX_ACCUM = 1 << 14,  // add to existing dvx instead of the momentum term
};

// XCD-aware block index: consecutive logical blocks (neighbouring cells)
// land on the same XCD, whose L2 then holds their shared neighbour rows.
__device__ __forceinline__ int xcd_block()
{
    const int nb = gridDim.x, b = blockIdx.x;
    const int xcd = b & 7, per = nb >> 3, rem = nb & 7;
    return xcd * per + min(xcd, rem) + (b >> 3);
}

// ------------------------------------------------------------------------
// utility kernels
// ------------------------------------------------------------------------
__device__ __forceinline__ void fill_int_body(int blk, int* a, int n, int v)
{
    int i = blk * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_fill_int(int* a, int n, int v)
{
    fill_int_body(blockIdx.x, a, n, v);
}

// a[0] &= keep (one thread).
__device__ __forceinline__ void keep_bits_body(int blk, int* a, int keep)
{
    if (blk == 0 && threadIdx.x == 0) a[0] &= keep;
}

__global__ void k_keep_bits(int* a, int keep)
{
    keep_bits_body(blockIdx.x, a, keep);
}

__global__ void k_fill_dbl(double* a, int n, double v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__device__ __forceinline__ void iota_body(int blk, int* a, int n)
{
    int i = blk * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

__global__ void k_iota(int* a, int n)
{
    iota_body(blockIdx.x, a, n);
}

// dst = src (element-wise copy).
template <class T>
__global__ void k_copy(const T* src, T* dst, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// dst[t] = src[idx[t]]: permutation into cell order, and host -> device
// transfer by reference particle number.
template <class T>
__global__ void k_gather(const T* src, T* dst, const int* idx, int n)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[idx[t]];
}

// Several arrays through one permutation in one pass: dst_a[t] = src_a[idx[t]]
// for the nd double and ni int arrays (idx == nullptr: a plain copy).
constexpr int PERM_MAXD = 28, PERM_MAXI = 4;
struct PermArgs {
    const double* sd[PERM_MAXD];
    double* dd[PERM_MAXD];
    const int* si[PERM_MAXI];
    int* di[PERM_MAXI];
    int nd, ni;
};

__device__ __forceinline__ void permute_multi_body(int blk, PermArgs A, const int* idx, int n)
{
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int s = idx ? idx[t] : t;
#pragma unroll
    for (int a = 0; a < PERM_MAXD; ++a)
        if (a < A.nd) A.dd[a][t] = A.sd[a][s];
#pragma unroll
    for (int a = 0; a < PERM_MAXI; ++a)
        if (a < A.ni) A.di[a][t] = A.si[a][s];
}

__global__ void k_permute_multi(PermArgs A, const int* idx, int n)
{
    permute_multi_body(blockIdx.x, A, idx, n);
}

// dst[idx[t]] = src[t]: device -> host order.
template <class T>
__global__ void k_scatter(const T* src, T* dst, const int* idx, int n)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[idx[t]] = src[t];
}

// Exclusive scan of n ints in tiles of SCAN_TILE; out has n+1 entries.
constexpr int SCAN_BLOCK = 256, SCAN_ITEMS = 8, SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

__device__ __forceinline__ int block_exclusive_scan(int v, int* lds, int& total)
{
    const int tid = threadIdx.x;
    lds[tid] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
        int t = (tid >= off) ? lds[tid - off] : 0;
        __syncthreads();
        lds[tid] += t;
        __syncthreads();
    }
    const int incl = lds[tid];
    total = lds[SCAN_BLOCK - 1];
    __syncthreads();
    return incl - v;
}

// clear (optional, the same array as in): zeroed after it is read, so that a
// counting pass can start from zeros without a fill.
__device__ __forceinline__ void scan_tile_body(int blk, const int* in, int* out, int n, int* tile_sums, int* clear)
{
    __shared__ int lds[SCAN_BLOCK];
    const int base = blk * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS], s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        v[i] = (base + i < n) ? in[base + i] : 0;
        const int t = v[i];
        v[i] = s;
        s += t;
    }
    if (clear) {
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; ++i)
            if (base + i < n) clear[base + i] = 0;
    }
    int total;
    const int pre = block_exclusive_scan(s, lds, total);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
        if (base + i < n) out[base + i] = pre + v[i];
    if (threadIdx.x == 0) tile_sums[blk] = total;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_tile(const int* in, int* out, int n, int* tile_sums, int* clear)
{
    scan_tile_body(blockIdx.x, in, out, n, tile_sums, clear);
}

// Single block: exclusive scan of the tile sums (ntiles <= SCAN_TILE), total at [ntiles].
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_sums(int* tile_sums, int ntiles)
{
    __shared__ int lds[SCAN_BLOCK];
    const int base = threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS], s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        v[i] = (base + i < ntiles) ? tile_sums[base + i] : 0;
        const int t = v[i];
        v[i] = s;
        s += t;
    }
    int total;
    const int pre = block_exclusive_scan(s, lds, total);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
        if (base + i < ntiles) tile_sums[base + i] = pre + v[i];
    if (threadIdx.x == 0) tile_sums[ntiles] = total;
}

__global__ void k_scan_add(int* out, int n, const int* tile_sums, int ntiles)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += tile_sums[i / SCAN_TILE];
    if (i == 0) out[n] = tile_sums[ntiles];
}

// ------------------------------------------------------------------------
// Cell list: key + rank (counting sort), scatter, per-cell ordering.
// ------------------------------------------------------------------------
__device__ __forceinline__ void cell_key_body(int blk, DPhase P, Grid g, int* err)
{
    const int i = blk * blockDim.x + threadIdx.x;
    const bool valid = i < P.n;
    int key = -1;
    if (valid) {
        const double x = P.x[i], y = P.y[i];
        int cx, cy;
        grid_cell(g, x, y, cx, cy);
        // grid_geom's range check (m_particles.f90:1131-1136): the grid spans the
        // geometry box [x_mingeom, x_maxgeom] x [y_mingeom, y_maxgeom].
        const bool bad = !(x >= g.x0 && x <= g.x1 && y >= g.y0 && y <= g.y1) || cx < 0 || cy < 0 ||
                         cx >= g.ncx || cy >= g.ncy;
        if (bad) atomicOr(err, 1);
        key = bad ? 0 : cy * g.ncx + cx;
        P.key[i] = key;
    }
    // Rank within the cell.  The state is kept in cell order, so lanes of a
    // wave mostly hold runs of equal keys: the head lane of each run takes
    // the run's slots with one atomic and the run's lanes share them out.  The
    // rank order inside a cell is arbitrary; k_cell_sort fixes it.
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(key, 1, 64);
    const bool head = lane == 0 || prev != key;
    const uint64_t heads = __ballot(head);
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);   // lanes 0..lane
    const int start = 63 - __clzll((long long)(heads & upto));
    const uint64_t after = heads & ~upto;
    const int end = after ? __ffsll((long long)after) - 1 : 64;
    int base = 0;
    if (head && key >= 0) base = atomicAdd(&P.cell_count[key], end - start);
    base = __shfl(base, start, 64);
    if (valid) P.slot[i] = base + (lane - start);
}

__global__ __launch_bounds__(256) void k_cell_key(DPhase P, Grid g, int* err)
{
    cell_key_body(blockIdx.x, P, g, err);
}

// Exclusive prefix of the tile sums of k_scan_tile (ntiles <= SCAN_TILE)
// into LDS tp[0..ntiles], by the whole workgroup (SCAN_BLOCK threads).
__device__ __forceinline__ void tile_prefix(const int* tile_sums, int ntiles, int* tp, int* lds)
{
    const int base = threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS], s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        v[i] = (base + i < ntiles) ? tile_sums[base + i] : 0;
        const int t = v[i];
        v[i] = s;
        s += t;
    }
    int total;
    const int pre = block_exclusive_scan(s, lds, total);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
        if (base + i <= ntiles) tp[base + i] = pre + v[i];
    if (threadIdx.x == 0 && ntiles == SCAN_TILE) tp[SCAN_TILE] = total;
    __syncthreads();
}

// Cell-list scatter on the tile-local prefix of the counts (part, from
// k_scan_tile) plus the prefix of the tile sums, computed per workgroup:
// the three-pass scan becomes one pass.
__device__ __forceinline__ void cell_scatter_t_body(int blk, DPhase P, const int* part, const int* tile_sums,
                                                    int ntiles)
{
    __shared__ int tp[SCAN_TILE + 1];
    __shared__ int lds[SCAN_BLOCK];
    tile_prefix(tile_sums, ntiles, tp, lds);
    const int i = blk * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const int key = P.key[i];
    P.order[part[key] + tp[key / SCAN_TILE] + P.slot[i]] = i;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_cell_scatter_t(DPhase P, const int* part, const int* tile_sums,
                                                               int ntiles)
{
    cell_scatter_t_body(blockIdx.x, P, part, tile_sums, ntiles);
}

// Ascending particle number inside each cell (insertion sort; the ranks of
// k_cell_key depend on atomic arrival order), which makes the neighbour
// visiting order and so every floating-point sum deterministic; and the
// final cell_start from the tile-local prefix and the tile sums.
__device__ __forceinline__ void cell_sort_t_body(int blk, DPhase P, const int* part, const int* tile_sums, int ntiles,
                                                 int ncell)
{
    __shared__ int tp[SCAN_TILE + 1];
    __shared__ int lds[SCAN_BLOCK];
    tile_prefix(tile_sums, ntiles, tp, lds);
    const int c = blk * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int s = part[c] + tp[c / SCAN_TILE];
    const int e = c + 1 < ncell ? part[c + 1] + tp[(c + 1) / SCAN_TILE] : tp[ntiles];
    P.cell_start[c] = s;
    if (c + 1 == ncell) P.cell_start[ncell] = e;
    constexpr int CS = 16;   // cells up to this size sort in registers
    if (e - s <= CS) {
        // the cell's entries in registers (independent loads), an unrolled
        // odd-even transposition network, one store each
        int v[CS];
#pragma unroll
        for (int q = 0; q < CS; ++q) v[q] = s + q < e ? P.order[s + q] : INT_MAX;
#pragma unroll
        for (int r = 0; r < CS; ++r) {
#pragma unroll
            for (int q = r & 1; q + 1 < CS; q += 2) {
                const int a = v[q], b = v[q + 1];
                v[q] = min(a, b);
                v[q + 1] = max(a, b);
            }
        }
#pragma unroll
        for (int q = 0; q < CS; ++q)
            if (s + q < e) P.order[s + q] = v[q];
        return;
    }
    for (int k = s + 1; k < e; ++k) {
        const int v = P.order[k];
        int m = k - 1;
        while (m >= s && P.order[m] > v) {
            P.order[m + 1] = P.order[m];
            --m;
        }
        P.order[m + 1] = v;
    }
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_cell_sort_t(DPhase P, const int* part, const int* tile_sums,
                                                            int ntiles, int ncell)
{
    cell_sort_t_body(blockIdx.x, P, part, tile_sums, ntiles, ncell);
}


// ------------------------------------------------------------------------
// Per-particle helpers
// ------------------------------------------------------------------------

// drucker_prager_failure_criterion for one particle (m_particles.f90:3349-3375).
__device__ __forceinline__ int dp_project(double alpha1, double& p, double& sxx, double& sxy, double& syy)
{
#pragma clang fp contract(off)
    if (p < 0.0) p = 0.0;
    const double I1 = 3.0 * p;
    double J2 = sxx * sxx + 2.0 * (sxy * sxy) + syy * syy + (sxx + syy) * (sxx + syy);
    J2 = __builtin_sqrt(J2 / 2.0) + 1.0e-6;
    if (J2 > alpha1 * I1) {
        const double skale = alpha1 * I1 / J2;
        sxx = skale * sxx;
        sxy = skale * sxy;
        syy = skale * syy;
        return 1;
    }
    return 0;
}

// mohr_coulomb_failure_criterion for one particle (m_particles.f90:3220-3266):
// where the yield stress c cos(phi) + p sin(phi) is not positive, p drops to
// the tension cut-off and rho to rho0; a maximum shear stress above the yield
// stress scales the deviator down to it.
__device__ __forceinline__ int mc_project(const KParams& K, double& p, double& sxx, double& sxy, double& syy,
                                          double& rho)
{
#pragma clang fp contract(off)
    const double h = (sxx - syy) / 2.0;
    const double tmax = __builtin_sqrt(h * h + sxy * sxy);
    if (tmax < (double)1.e-6f) return 0;
    double yield = K.mc_ccos + p * K.mc_sin;
    if (yield <= 0.0) {
        yield = 0.0;
        p = K.mc_pcut;
        rho = K.s_rho0;
    }
    if (tmax > yield) {
        const double skale = yield / tmax;
        sxx = skale * sxx;
        sxy = skale * sxy;
        syy = skale * syy;
        return 1;
    }
    return 0;
}

// Link-list reach restriction of the reference (m_particles.f90:1196-1216):
// the partner's sorting-grid cell must be within ghsmlx cells.
__device__ __forceinline__ bool ll_reach(const Grid& g, double xa, double ya, double xb, double yb)
{
    const int ax = ref_gcell(g.lscale_x, g.lmin_x, xa), bx = ref_gcell(g.lscale_x, g.lmin_x, xb);
    const int ay = ref_gcell(g.lscale_y, g.lmin_y, ya), by = ref_gcell(g.lscale_y, g.lmin_y, yb);
    return abs(ax - bx) <= g.lreach_x && abs(ay - by) <= g.lreach_y;
}

// ------------------------------------------------------------------------
// Leapfrog (time_integration, declarations_sph.f90:643-762)
// ------------------------------------------------------------------------

// first_half (:643-684) for water.
__global__ void k_first_half_water(DPhase P, KParams K)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const double h = K.dt / 2.0;
    if (!K.summation_density) {
        const double r = P.rho[i];
        P.rho_min[i] = r;
        P.rho[i] = r + h * P.drho[i];
    }
    if (K.volume_fraction && K.two_phase) {
        const double v = P.vof[i];
        P.vof_min[i] = v;
        P.vof[i] = v + h * P.dvof[i];
    }
    if (P.itype[i] < 0) return;
    const double vx = P.vx[i], vy = P.vy[i];
    P.vx_min[i] = vx;
    P.vy_min[i] = vy;
    P.vx[i] = vx + h * P.ax[i];
    P.vy[i] = vy + h * P.ay[i];
}

// first_half (:643-684) for soil followed by `ndp` Drucker-Prager
// projections (time_integration :559 and single_step :1244).
__global__ void k_first_half_soil(DPhase P, KParams K, int do_half, int ndp)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    if (do_half) {
        const double h = K.dt / 2.0;
        if (!K.summation_density) {
            const double r = P.rho[i];
            P.rho_min[i] = r;
            P.rho[i] = r + h * P.drho[i];
        }
        if (P.itype[i] > 0) {
            const double vx = P.vx[i], vy = P.vy[i];
            P.vx_min[i] = vx;
            P.vy_min[i] = vy;
            P.vx[i] = vx + h * P.ax[i];
            P.vy[i] = vy + h * P.ay[i];
        }
    }
    if (ndp > 0) {
        double p = P.p[i], sxx = P.sxx[i], sxy = P.sxy[i], syy = P.syy[i];
        int f = 0;
        for (int k = 0; k < ndp; ++k) f = dp_project(K.s_alpha1, p, sxx, sxy, syy);
        P.p[i] = p; P.sxx[i] = sxx; P.sxy[i] = sxy; P.syy[i] = syy;
        P.fail[i] = f;
    }
}

// first_step (:686-723) / second_half (:725-762) for water.  renorm: the
// volume-fraction renormalisation that single_step applies every 40th step
// (declarations_sph.f90:1414-1420) before the update, rho <- rho / vof * vof2
// with the half-step vof; second_half then recomputes vof (and, unless
// summation_density is set, rho), so only a summation-density run keeps it.
// The final update of a water particle (second_half of the leapfrog) as
// values, without writing: x, v, rho, vof after the step.
struct WaterEnd {
    double x, y, vx, vy, rho, vof;
};

__device__ __forceinline__ WaterEnd water_end(const DPhase& P, const KParams& K, int i, int first, int with_vof,
                                              int renorm)
{
#pragma clang fp contract(off)
    const double dt = K.dt, h = dt / 2.0;
    WaterEnd e;
    // (a level that the update replaces by the saved one is not read: after
    // a fused final/pack the state does not hold it)
    if (!K.summation_density && !first) {
        e.rho = P.rho_min[i] + dt * P.drho[i];
    } else {
        e.rho = P.rho[i];
        if (renorm) e.rho = e.rho / P.vof[i] * P.vof2[i];
        if (!K.summation_density) e.rho = e.rho + h * P.drho[i];
    }
    if (with_vof && !first) {
        e.vof = P.vof_min[i] + dt * P.dvof[i];
    } else {
        e.vof = P.vof[i];
        if (with_vof) e.vof = e.vof + h * P.dvof[i];
    }
    e.x = P.x[i];
    e.y = P.y[i];
    const int it = P.itype[i];
    if (it >= 0 && !first) {
        e.vx = P.vx_min[i] + dt * P.ax[i] + P.avx[i];
        e.vy = P.vy_min[i] + dt * P.ay[i] + P.avy[i];
    } else {
        e.vx = P.vx[i];
        e.vy = P.vy[i];
        if (it >= 0) {
            e.vx = e.vx + h * P.ax[i] + P.avx[i];
            e.vy = e.vy + h * P.ay[i] + P.avy[i];
        }
    }
    if (it >= 0) {
        e.x = e.x + dt * e.vx;
        e.y = e.y + dt * e.vy;
    }
    return e;
}

__device__ __forceinline__ void final_water(int blk, DPhase P, KParams K, int first, int with_vof, int renorm)
{
    const int i = blk * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const WaterEnd e = water_end(P, K, i, first, with_vof, renorm);
    if (renorm || !K.summation_density) P.rho[i] = e.rho;
    if (with_vof) P.vof[i] = e.vof;
    if (P.itype[i] < 0) return;
    P.vx[i] = e.vx;
    P.vy[i] = e.vy;
    P.x[i] = e.x;
    P.y[i] = e.y;
}

__global__ void k_final_water(DPhase P, KParams K, int first, int with_vof, int renorm)
{
    final_water(blockIdx.x, P, K, first, with_vof, renorm);
}

// The renormalisation as single_step itself leaves it (declarations_sph.f90:
// 1414-1420), for the bare right-hand side: rho <- rho / vof * vof2, vof <- vof2.
__global__ void k_renorm_water(DPhase P)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const double v2 = P.vof2[i];
    P.rho[i] = P.rho[i] / P.vof[i] * v2;
    P.vof[i] = v2;
}

// time_integration_for_water_by_verlet (declarations_sph.f90:889-974), the
// update after single_step_for_water of step `step`: real particles (the
// first ntotal of the reference numbering) move by dt v + h2 a with
// h2 = dt**2./2. and advance their other velocity level by 2 dt a (every
// 50th step: restart from v + dt a); every particle's other density level
// advances by 2 dt drho (step 1 and every 50th step: restart); then both
// levels swap, as the reference's whole-array swaps do (the virtual
// particles' velocity levels swap unchanged).
__global__ void k_verlet_water(DPhase P, double* lvx, double* lvy, double* lrho, double dt, double h2, int step)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const double vx = P.vx[i], vy = P.vy[i];
    double nvx = lvx[i], nvy = lvy[i];
    if (P.id[i] < P.ntotal) {
        const double ax = P.ax[i], ay = P.ay[i];
        P.x[i] = P.x[i] + dt * vx + h2 * ax;
        P.y[i] = P.y[i] + dt * vy + h2 * ay;
        if (step % 50 != 0) {
            nvx = nvx + 2.0 * dt * ax;
            nvy = nvy + 2.0 * dt * ay;
        } else {
            nvx = vx + dt * ax;
            nvy = vy + dt * ay;
        }
    }
    P.vx[i] = nvx;
    P.vy[i] = nvy;
    lvx[i] = vx;
    lvy[i] = vy;
    const double r = P.rho[i], dr = P.drho[i];
    double nr;
    if (step == 1) nr = r + dt * 2.0 * dr;
    else if (step % 50 != 0) nr = lrho[i] + dt * 2.0 * dr;
    else nr = r + dt * dr;
    P.rho[i] = nr;
    lrho[i] = r;
}

// volume_fraction_soil (m_particles.f90:3970-3988, applied by single_step
// :1411 to the pre-update density), then first_step / second_half
// (stress_integration = 1) and the closing Drucker-Prager projection (:583/:601).
// The final update of a soil particle (second_half, plastic_or_not and
// return_mapping, the closing Drucker-Prager projection) as values.
struct SoilEnd {
    double x, y, vx, vy, rho, vof, sxx, sxy, syy, p;
    int fail;
};

__device__ __forceinline__ SoilEnd soil_end(const DPhase& P, const KParams& K, int i, int first)
{
#pragma clang fp contract(off)
    const double dt = K.dt, h = dt / 2.0;
    SoilEnd e;
    e.vof = (K.volume_fraction && K.two_phase) ? P.rho[i] / K.s_rho0 : P.vof[i];
    e.rho = P.rho[i];
    if (!K.summation_density) e.rho = first ? e.rho + h * P.drho[i] : P.rho_min[i] + dt * P.drho[i];
    double sxx, sxy, syy, p;
    if (K.stress_integration == 2) {   // :700-705, :744-749
        if (first) {
            sxx = P.sxx[i] + h * P.dsxx[i];
            sxy = P.sxy[i] + h * P.dsxy[i];
            syy = P.syy[i] + h * P.dsyy[i];
            p = P.p[i] + h * P.dp[i];
        } else {
            sxx = P.sxx_min[i] + dt * P.dsxx[i];
            sxy = P.sxy_min[i] + dt * P.dsxy[i];
            syy = P.syy_min[i] + dt * P.dsyy[i];
            p = P.p_min[i] + dt * P.dp[i];
        }
    } else {
        sxx = P.sxx[i] + dt * P.dsxx[i];
        sxy = P.sxy[i] + dt * P.dsxy[i];
        syy = P.syy[i] + dt * P.dsyy[i];
        p = P.p[i] + dt * P.dp[i];
    }
    e.x = P.x[i];
    e.y = P.y[i];
    const int it = P.itype[i];
    if (it > 0 && !first) {   // (the current level is not read, as for water)
        e.vx = P.vx_min[i] + dt * P.ax[i] + P.avx[i];
        e.vy = P.vy_min[i] + dt * P.ay[i] + P.avy[i];
    } else {
        e.vx = P.vx[i];
        e.vy = P.vy[i];
        if (it > 0) {
            e.vx = e.vx + h * P.ax[i] + P.avx[i];
            e.vy = e.vy + h * P.ay[i] + P.avy[i];
        }
    }
    if (it > 0) {
        e.x = e.x + dt * e.vx;
        e.y = e.y + dt * e.vy;
    }
    if (K.plasticity == 2 && K.two_phase) {
        // plastic_or_not + return_mapping (time_integration :576-579, :594-597,
        // :764-789) on the updated stress
        const double I1 = 3.0 * p;
        const double j2r = sxx * sxx + 2.0 * (sxy * sxy) + syy * syy + (sxx + syy) * (sxx + syy);
        const double J2 = __builtin_sqrt(j2r / 2.0) + 1.0e-6;
        if (fabs(J2 - K.s_alpha1 * I1) < SPH_YIELD_TOL_F) {
            if (K.stress_integration == 1) {
                sxx = sxx + dt * P.dsxx2[i];
                sxy = sxy + dt * P.dsxy2[i];
                syy = syy + dt * P.dsyy2[i];
                p = p + dt * P.dp2[i];
            } else {
                sxx = P.sxx_min[i] + dt * P.dsxx[i];
                sxy = P.sxy_min[i] + dt * P.dsxy[i];
                syy = P.syy_min[i] + dt * P.dsyy[i];
                p = P.p_min[i] + dt * P.dp[i];
            }
        }
    }
    e.fail = dp_project(K.s_alpha1, p, sxx, sxy, syy);
    e.sxx = sxx;
    e.sxy = sxy;
    e.syy = syy;
    e.p = p;
    return e;
}

__device__ __forceinline__ void final_soil(int blk, DPhase P, KParams K, int first)
{
    const int i = blk * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const SoilEnd e = soil_end(P, K, i, first);
    if (K.volume_fraction && K.two_phase) P.vof[i] = e.vof;
    if (!K.summation_density) P.rho[i] = e.rho;
    if (P.itype[i] > 0) {
        P.vx[i] = e.vx;
        P.vy[i] = e.vy;
        P.x[i] = e.x;
        P.y[i] = e.y;
    }
    P.fail[i] = e.fail;
    P.sxx[i] = e.sxx; P.sxy[i] = e.sxy; P.syy[i] = e.syy; P.p[i] = e.p;
}

__global__ void k_final_soil(DPhase P, KParams K, int first)
{
    final_soil(blockIdx.x, P, K, first);
}

// ------------------------------------------------------------------------
// Stage A, water: one gather over water neighbours and one over soil
// neighbours.  Terms (reference routine -> accumulator):
//   con_density (:2402)        -> con        delta_sph(rho) (:2658) -> drd
//   shear_strain_rate (:2855)  -> txx,txy,tyy
//   av_vel (:2761)             -> avx,avy    repulsive_force (:1605) -> rep
//   darcy_law water side (:3832) -> dg      volume_fraction_water2 (:3940) -> dvx_cross
//   volume_fraction_water (:3912) -> vof2
// then pressure() (:2133) per particle: p and the new c.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_water_a(DPhase W, DPhase S, Grid g, KParams K, int mask)
{
    const int t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= W.n) return;
    const CellList cw{W.cell_start, W.order};
    const int a = cw.order[t];
    const double xa = W.x[a], ya = W.y[a], vxa = W.vx[a], vya = W.vy[a];
    const double ha = W.hsml[a], rhoa = W.rho[a], ca = W.c[a];
    const int ita = W.itype[a];
    const bool link = K.nnps == 2;
    const double sk = K.scale_k;
    double con = 0.0, drd = 0.0, txx = 0.0, txy = 0.0, tyy = 0.0, avx = 0.0, avy = 0.0, repx = 0.0, repy = 0.0;
    int cnt = 0;
    for_each_candidate(g, cw, xa, ya, [&](int b) {
        if (b == a) return;
        const double xb = W.x[b], yb = W.y[b];
        const double mh = link ? K.h1_water : sph_mhsml(ha, W.hsml[b]);
        double dx, dy, r;
        if (!sph_pair_test(xa, ya, xb, yb, sk * mh, dx, dy, r)) return;
        if (link && !ll_reach(g, xa, ya, xb, yb)) return;
        ++cnt;
        double w, dwx, dwy;
        sph_kernel(K.skf, r, dx, dy, mh, w, dwx, dwy);
        const double mb = W.mass[b], rhob = W.rho[b], vxb = W.vx[b], vyb = W.vy[b];
        if (mask & M_CON) {
            double vcc = (vxa - vxb) * dwx;
            vcc = vcc + (vya - vyb) * dwy;
            con += mb * vcc;
        }
        if (mask & M_DRHO_DELTA) {
            const double rr = dx * dx + dy * dy;
            const double muv = 2.0 * (rhoa - rhob) / rr;
            double h = dx * muv * dwx;
            h = h + dy * muv * dwy;
            drd += K.delta * ha * ca * mb * h / rhob;
        }
        if (mask & M_STRAIN) {
            const double dvx = vxb - vxa, dvy = vyb - vya;
            double hxx = 2.0 * dvx * dwx - dvy * dwy;
            const double hxy = dvx * dwy + dvy * dwx;
            double hyy = 2.0 * dvy * dwy - dvx * dwx;
            hxx = SPH_C23_F * hxx;
            hyy = SPH_C23_F * hyy;
            txx += mb * hxx / rhob;
            txy += mb * hxy / rhob;
            tyy += mb * hyy / rhob;
        }
        if (mask & M_XSPH) {
            const double mrho = (rhoa + rhob) / 2.0;
            avx -= mb * (vxa - vxb) / mrho * w;
            avy -= mb * (vya - vyb) / mrho * w;
        }
        if ((mask & M_REP) && ita > 0 && W.itype[b] < 0 && !(r > W.dspp)) {
            const double q = W.dspp / r;
            const double f = (pow(q, K.p1) - pow(q, K.p2)) / (r * r);
            repx += K.dd * dx * f;
            repy += K.dd * dy * f;
        }
    });
    double dg_x = 0.0, dg_y = 0.0, dvof_x = 0.0, vof2 = 0.0;
    int cntx = 0;
    if (mask & M_CROSS) {
        const CellList cs{S.cell_start, S.order};
        const double vofa = W.vof[a];
        const double cf0 = -K.numeric_gravity;
        for_each_candidate(g, cs, xa, ya, [&](int b) {
            const double xb = S.x[b], yb = S.y[b];
            const double mh = sph_mhsml(ha, S.hsml[b]);
            double dx, dy, r;
            if (!sph_pair_test(xa, ya, xb, yb, sk * mh, dx, dy, r)) return;
            ++cntx;
            double w, dwx, dwy;
            sph_kernel(K.skf, r, dx, dy, mh, w, dwx, dwy);
            const double mb = S.mass[b], rhob = S.rho[b];
            const double dvx = vxa - S.vx[b], dvy = vya - S.vy[b];
            const double rrw = w / (rhoa * rhob);
            const double cf = vofa * S.vof[b] * rhoa * cf0 / K.s_perm;
            dg_x += mb * (cf * dvx * rrw);
            dg_y += mb * (cf * dvy * rrw);
            const double tmp = dvx * dwx + dvy * dwy;
            dvof_x += mb * tmp / K.s_rho0;
            vof2 += mb * w;
        });
    }
    // outputs
    if (mask & (M_CON | M_DRHO_DELTA)) {
        double base = (mask & M_CON) ? con : ((mask & M_DRHO_ACCUM) ? W.drho[a] : 0.0);
        W.drho[a] = base + drd;
    }
    if (mask & M_STRAIN) { W.txx[a] = txx; W.txy[a] = txy; W.tyy[a] = tyy; }
    if (mask & M_XSPH) {
        if (ita > 0) { avx = K.epsilon * avx; avy = K.epsilon * avy; }
        W.avx[a] = avx; W.avy[a] = avy;
    }
    if (mask & M_REP) { W.repx[a] = repx; W.repy[a] = repy; }
    if (mask & M_CROSS) {
        W.drgx[a] = dg_x; W.drgy[a] = dg_y; W.dvof_x[a] = dvof_x;
        W.vof2[a] = 1.0 - vof2 / K.s_rho0;
        W.countiac[a] = cntx;
    } else {
        W.countiac[a] = cnt;
    }
    if (mask & M_EOS) {
        const double rr = rhoa / (K.w_rho0 * W.vof[a]);
        W.p[a] = K.w_b * (pow(rr, K.w_gamma) - 1.0);
        W.c_next[a] = K.w_c * pow(rr, 3.0);
    }
}

// ------------------------------------------------------------------------
// Stage A, soil.  Same-phase terms: con_density, delta_sph(rho),
// shear_strain_rate, velocity_divergence (:1996), the Jaumann spin (:3083),
// av_vel, repulsive_force.  Cross terms over water neighbours: darcy_law soil
// side and pore_water_pressure (:3877, uses the new water p).  Per particle:
// pressure() sets c, then plastic_or_not + plastic_flow_rule3 and Jaumann_rate.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_soil_a(DPhase S, DPhase W, Grid g, KParams K, int mask)
{
    const int t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= S.n) return;
    const CellList cs{S.cell_start, S.order};
    const int a = cs.order[t];
    const double xa = S.x[a], ya = S.y[a], vxa = S.vx[a], vya = S.vy[a];
    const double ha = S.hsml[a], rhoa = S.rho[a], ca = S.c[a];
    const int ita = S.itype[a];
    const bool link = K.nnps == 2;
    const double sk = K.scale_k;
    double con = 0.0, drd = 0.0, txx = 0.0, txy = 0.0, tyy = 0.0, vcc = 0.0, wxy = 0.0;
    double avx = 0.0, avy = 0.0, repx = 0.0, repy = 0.0;
    int cnt = 0;
    for_each_candidate(g, cs, xa, ya, [&](int b) {
        if (b == a) return;
        const double xb = S.x[b], yb = S.y[b];
        const double mh = link ? K.h1_soil : sph_mhsml(ha, S.hsml[b]);
        double dx, dy, r;
        if (!sph_pair_test(xa, ya, xb, yb, sk * mh, dx, dy, r)) return;
        if (link && !ll_reach(g, xa, ya, xb, yb)) return;
        ++cnt;
        double w, dwx, dwy;
        sph_kernel(K.skf, r, dx, dy, mh, w, dwx, dwy);
        const double mb = S.mass[b], rhob = S.rho[b], vxb = S.vx[b], vyb = S.vy[b];
        if (mask & M_CON) {
            double v = (vxa - vxb) * dwx;
            v = v + (vya - vyb) * dwy;
            con += mb * v;
        }
        if (mask & M_DRHO_DELTA) {
            const double rr = dx * dx + dy * dy;
            const double muv = 2.0 * (rhoa - rhob) / rr;
            double h = dx * muv * dwx;
            h = h + dy * muv * dwy;
            drd += K.delta * ha * ca * mb * h / rhob;
        }
        if (mask & M_STRAIN) {
            const double dvx = vxb - vxa, dvy = vyb - vya;
            double hxx = 2.0 * dvx * dwx - dvy * dwy;
            const double hxy = dvx * dwy + dvy * dwx;
            double hyy = 2.0 * dvy * dwy - dvx * dwx;
            hxx = SPH_C23_F * hxx;
            hyy = SPH_C23_F * hyy;
            txx += mb * hxx / rhob;
            txy += mb * hxy / rhob;
            tyy += mb * hyy / rhob;
            double hv = dvx * dwx;
            hv = hv + dvy * dwy;
            vcc += mb * hv / rhob;
        }
        if (mask & M_PLASTIC) {
            const double dvx = vxb - vxa, dvy = vyb - vya;
            const double hxy = 0.5 * (dvx * dwy - dvy * dwx);
            wxy += mb * hxy / rhob;
        }
        if (mask & M_XSPH) {
            const double mrho = (rhoa + rhob) / 2.0;
            avx -= mb * (vxa - vxb) / mrho * w;
            avy -= mb * (vya - vyb) / mrho * w;
        }
        if ((mask & M_REP) && ita > 0 && S.itype[b] < 0 && !(r > S.dspp)) {
            const double q = S.dspp / r;
            const double f = (pow(q, K.p1) - pow(q, K.p2)) / (r * r);
            repx += K.dd * dx * f;
            repy += K.dd * dy * f;
        }
    });
    double dg_x = 0.0, dg_y = 0.0, po_x = 0.0, po_y = 0.0;
    int cntx = 0;
    if (mask & M_CROSS) {
        const CellList cw{W.cell_start, W.order};
        const double vofa = S.vof[a];
        const double cf0 = -K.numeric_gravity;
        for_each_candidate(g, cw, xa, ya, [&](int b) {
            const double xb = W.x[b], yb = W.y[b];
            const double mh = sph_mhsml(W.hsml[b], ha);
            double dx, dy, r;   // oriented water - soil, as direct_find_2
            if (!sph_pair_test(xb, yb, xa, ya, sk * mh, dx, dy, r)) return;
            ++cntx;
            double w, dwx, dwy;
            sph_kernel(K.skf, r, dx, dy, mh, w, dwx, dwy);
            const double mw = W.mass[b], rhow = W.rho[b];
            const double rrw = w / (rhow * rhoa);
            const double cf = W.vof[b] * vofa * rhow * cf0 / K.s_perm;
            dg_x += mw * (cf * (W.vx[b] - vxa) * rrw);
            dg_y += mw * (cf * (W.vy[b] - vya) * rrw);
            const double mprr = mw * W.p[b] / (rhow * rhoa);
            po_x += mprr * dwx * vofa;
            po_y += mprr * dwy * vofa;
        });
    }
    if (mask & (M_CON | M_DRHO_DELTA)) {
        double base = (mask & M_CON) ? con : ((mask & M_DRHO_ACCUM) ? S.drho[a] : 0.0);
        S.drho[a] = base + drd;
    }
    if (mask & M_STRAIN) { S.txx[a] = txx; S.txy[a] = txy; S.tyy[a] = tyy; S.vcc[a] = vcc; }
    if (mask & M_XSPH) {
        if (ita > 0) { avx = K.epsilon * avx; avy = K.epsilon * avy; }
        S.avx[a] = avx; S.avy[a] = avy;
    }
    if (mask & M_REP) { S.repx[a] = repx; S.repy[a] = repy; }
    if (mask & M_CROSS) {
        S.drgx[a] = dg_x; S.drgy[a] = dg_y; S.porx[a] = po_x; S.pory[a] = po_y;
        S.countiac[a] = cntx;
    } else {
        S.countiac[a] = cnt;
    }
    if (mask & M_EOS) S.c_next[a] = K.s_c;
    if (mask & M_PLASTIC) {
#pragma clang fp contract(off)
        double dsxx = 0.0, dsxy = 0.0, dsyy = 0.0, dpv = 0.0;
        if (mask & M_DS_ACCUM) { dsxx = S.dsxx[a]; dsxy = S.dsxy[a]; dsyy = S.dsyy[a]; dpv = S.dp[a]; }
        const double sxx = S.sxx[a], sxy = S.sxy[a], syy = S.syy[a], p = S.p[a];
        const double alpha = K.s_alpha1, G = K.s_G, Kb = K.s_k;
        int fail = 0;
        if (K.plasticity == 3) {
            // plastic_or_not (:3447-3505)
            const double I1 = 3.0 * p;
            double J2 = sxx * sxx + 2.0 * (sxy * sxy) + syy * syy + (sxx + syy) * (sxx + syy);
            J2 = __builtin_sqrt(J2 / 2.0) + 1.0e-6;
            fail = fabs(J2 - alpha * I1) < SPH_YIELD_TOL_F ? 1 : 0;
            // plastic_flow_rule3 (:3700-3762)
            if (fail) {
                const double exx = txx / 2.0 + vcc / 3.0;
                const double exy = txy / 2.0;
                const double eyy = tyy / 2.0 + vcc / 3.0;
                const double sde = sxx * exx + 2.0 * sxy * exy + syy * eyy;
                double j2 = (sxx * sxx + 2.0 * (sxy * sxy) + syy * syy + (sxx + syy) * (sxx + syy)) / 2.0;
                j2 = j2 + 1.0e-10;
                const double sj = __builtin_sqrt(j2);
                const double dl = (3.0 * alpha * Kb * vcc + G / sj * sde) / (9.0 * (alpha * alpha) * Kb + G);
                if (!(dl < 0.0)) {
                    dsxx = dsxx - dl * (G / sj * sxx);
                    dsxy = dsxy - dl * (G / sj * sxy);
                    dsyy = dsyy - dl * (G / sj * syy);
                    dpv = dpv + 3.0 * Kb * alpha * dl;
                    S.eps_p[a] = S.eps_p[a] + dl * sxy / (2.0 * sj) * SPH_EPSP_F;
                }
            }
            S.fail[a] = fail;
        }
        // Jaumann_rate (:3047-3119): Hooke's law, -K vcc, rotation by the spin
        dsxx = dsxx + G * txx;
        dsxy = dsxy + G * txy;
        dsyy = dsyy + G * tyy;
        dpv = dpv - Kb * vcc;
        dsxx = dsxx + 2.0 * sxy * wxy;
        dsxy = dsxy - (sxx - syy) * wxy;
        dsyy = dsyy - 2.0 * sxy * wxy;
        S.dsxx[a] = dsxx; S.dsxy[a] = dsxy; S.dsyy[a] = dsyy; S.dp[a] = dpv;
        S.wxy[a] = wxy;
    }
}

// ------------------------------------------------------------------------
// Stage B, water: internal force -vof*df(p) + df(vof*sigma) (:1257-1265)
// with sigma = mu*t (newtonian_fluid), art_visc (:2504), delta_sph(vof,dvof)
// (:1337); then /rho, + repulsion, + gravity, - Darcy drag, and
// dvof = delta term - cross term.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_water_b(DPhase W, Grid g, KParams K, int mask)
{
    const int t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= W.n) return;
    const CellList cw{W.cell_start, W.order};
    const int a = cw.order[t];
    const double xa = W.x[a], ya = W.y[a], vxa = W.vx[a], vya = W.vy[a];
    const double ha = W.hsml[a], rhoa = W.rho[a], ca = W.c[a], vofa = W.vof[a], pa = W.p[a];
    const double mu = K.w_mu;
    const double fxx_a = vofa * (mu * W.txx[a]), fxy_a = vofa * (mu * W.txy[a]), fyy_a = vofa * (mu * W.tyy[a]);
    const bool link = K.nnps == 2;
    const double sk = K.scale_k;
    double Px = 0.0, Py = 0.0, Sxx = 0.0, Sxyy = 0.0, Sxyx = 0.0, Syy = 0.0;
    double artx = 0.0, arty = 0.0, dvd = 0.0;
    for_each_candidate(g, cw, xa, ya, [&](int b) {
        if (b == a) return;
        const double xb = W.x[b], yb = W.y[b], hb = W.hsml[b];
        const double mh = link ? K.h1_water : sph_mhsml(ha, hb);
        double dx, dy, r;
        if (!sph_pair_test(xa, ya, xb, yb, sk * mh, dx, dy, r)) return;
        if (link && !ll_reach(g, xa, ya, xb, yb)) return;
        double w, dwx, dwy;
        sph_kernel(K.skf, r, dx, dy, mh, w, dwx, dwy);
        const double mb = W.mass[b], rhob = W.rho[b];
        if (mask & M_MOMENTUM) {
            const double mr = mb / rhob;
            const double pb = W.p[b], vofb = W.vof[b];
            Px += mr * ((pa + pb) * dwx);
            Py += mr * ((pa + pb) * dwy);
            const double fxx = vofb * (mu * W.txx[b]), fxy = vofb * (mu * W.txy[b]), fyy = vofb * (mu * W.tyy[b]);
            Sxx += mr * ((fxx_a + fxx) * dwx);
            Sxyy += mr * ((fxy_a + fxy) * dwy);
            Sxyx += mr * ((fxy_a + fxy) * dwx);
            Syy += mr * ((fyy_a + fyy) * dwy);
        }
        if (mask & M_ART) {
            const double vr0 = (vxa - W.vx[b]) * dx;
            const double vr = vr0 + (vya - W.vy[b]) * dy;
            if (vr < 0.0) {
                const double amh = sph_mhsml(ha, hb);
                const double rr = dx * dx + dy * dy;
                const double muv = amh * vr / (rr + amh * amh * K.etq * K.etq);
                const double mc = 0.5 * (ca + W.c[b]);
                const double mrho = 0.5 * (rhoa + rhob);
                const double piv = (K.beta * muv - K.alpha * mc) * muv / mrho;
                artx += mb * (-piv * dwx);
                arty += mb * (-piv * dwy);
            }
        }
        if (mask & M_DELTA_FIELD) {
            const double rr = dx * dx + dy * dy;
            const double muv = 2.0 * (vofa - W.vof[b]) / rr;
            double h = dx * muv * dwx;
            h = h + dy * muv * dwy;
            dvd += K.delta * ha * ca * mb * h / rhob;
        }
    });
    {
#pragma clang fp contract(off)
        double ax, ay;
        if (mask & M_MOMENTUM) {
            ax = -(vofa * Px) + Sxx + Sxyy;
            ay = -(vofa * Py) + Sxyx + Syy;
            if (rhoa > 0.0) { ax = ax / rhoa; ay = ay / rhoa; }
        } else {
            ax = W.ax[a]; ay = W.ay[a];
        }
        if (mask & M_ART) { ax = ax + artx; ay = ay + arty; }
        if (mask & M_REP) { ax = ax + W.repx[a]; ay = ay + W.repy[a]; }
        if (mask & M_GRAV) ay = ay + K.gravity;
        if (mask & M_CROSS) { ax = ax - W.drgx[a]; ay = ay - W.drgy[a]; }
        W.ax[a] = ax;
        W.ay[a] = ay;
        if (mask & M_DELTA_FIELD) {
            double d = dvd;
            if (mask & M_CROSS) d = d - W.dvof_x[a];
            W.dvof[a] = d;
        }
    }
}

// ------------------------------------------------------------------------
// Stage B, soil: -df3(vof*p) + df3(vof*sigma) (:1270-1277), art_visc,
// delta_sph on p, sxx, sxy, syy (:1323-1331); then /rho, + repulsion,
// + gravity, + Darcy drag, + pore-water pressure force.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_soil_b(DPhase S, Grid g, KParams K, int mask)
{
    const int t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= S.n) return;
    const CellList cs{S.cell_start, S.order};
    const int a = cs.order[t];
    const double xa = S.x[a], ya = S.y[a], vxa = S.vx[a], vya = S.vy[a];
    const double ha = S.hsml[a], rhoa = S.rho[a], ca = S.c[a], vofa = S.vof[a];
    const double pa = S.p[a], sxxa = S.sxx[a], sxya = S.sxy[a], syya = S.syy[a];
    const double r2a = rhoa * rhoa;
    const double fpa = (vofa * pa) / r2a, fxxa = (vofa * sxxa) / r2a, fxya = (vofa * sxya) / r2a,
                 fyya = (vofa * syya) / r2a;
    const bool link = K.nnps == 2;
    const double sk = K.scale_k;
    double Ax = 0.0, Ay = 0.0, Bxx = 0.0, Cxyy = 0.0, Cxyx = 0.0, Dyy = 0.0;
    double artx = 0.0, arty = 0.0, dP = 0.0, dXX = 0.0, dXY = 0.0, dYY = 0.0;
    for_each_candidate(g, cs, xa, ya, [&](int b) {
        if (b == a) return;
        const double xb = S.x[b], yb = S.y[b], hb = S.hsml[b];
        const double mh = link ? K.h1_soil : sph_mhsml(ha, hb);
        double dx, dy, r;
        if (!sph_pair_test(xa, ya, xb, yb, sk * mh, dx, dy, r)) return;
        if (link && !ll_reach(g, xa, ya, xb, yb)) return;
        double w, dwx, dwy;
        sph_kernel(K.skf, r, dx, dy, mh, w, dwx, dwy);
        const double mb = S.mass[b], rhob = S.rho[b];
        if (mask & M_MOMENTUM) {
            const double r2b = rhob * rhob, vofb = S.vof[b];
            const double fp = fpa + (vofb * S.p[b]) / r2b;
            const double fxx = fxxa + (vofb * S.sxx[b]) / r2b;
            const double fxy = fxya + (vofb * S.sxy[b]) / r2b;
            const double fyy = fyya + (vofb * S.syy[b]) / r2b;
            Ax += mb * (fp * dwx);
            Ay += mb * (fp * dwy);
            Bxx += mb * (fxx * dwx);
            Cxyy += mb * (fxy * dwy);
            Cxyx += mb * (fxy * dwx);
            Dyy += mb * (fyy * dwy);
        }
        if (mask & M_ART) {
            const double vr0 = (vxa - S.vx[b]) * dx;
            const double vr = vr0 + (vya - S.vy[b]) * dy;
            if (vr < 0.0) {
                const double amh = sph_mhsml(ha, hb);
                const double rr = dx * dx + dy * dy;
                const double muv = amh * vr / (rr + amh * amh * K.etq * K.etq);
                const double mc = 0.5 * (ca + S.c[b]);
                const double mrho = 0.5 * (rhoa + rhob);
                const double piv = (K.beta * muv - K.alpha * mc) * muv / mrho;
                artx += mb * (-piv * dwx);
                arty += mb * (-piv * dwy);
            }
        }
        if (mask & M_DELTA_FIELD) {
            const double rr = dx * dx + dy * dy;
            const double coef = K.delta * ha * ca * mb;
            {
                const double muv = 2.0 * (pa - S.p[b]) / rr;
                double h = dx * muv * dwx; h = h + dy * muv * dwy;
                dP += coef * h / rhob;
            }
            {
                const double muv = 2.0 * (sxxa - S.sxx[b]) / rr;
                double h = dx * muv * dwx; h = h + dy * muv * dwy;
                dXX += coef * h / rhob;
            }
            {
                const double muv = 2.0 * (sxya - S.sxy[b]) / rr;
                double h = dx * muv * dwx; h = h + dy * muv * dwy;
                dXY += coef * h / rhob;
            }
            {
                const double muv = 2.0 * (syya - S.syy[b]) / rr;
                double h = dx * muv * dwx; h = h + dy * muv * dwy;
                dYY += coef * h / rhob;
            }
        }
    });
    {
#pragma clang fp contract(off)
        double ax, ay;
        if (mask & M_MOMENTUM) {
            ax = -(Ax * rhoa) + Bxx * rhoa + Cxyy * rhoa;
            ay = -(Ay * rhoa) + Cxyx * rhoa + Dyy * rhoa;
            if (rhoa > 0.0) { ax = ax / rhoa; ay = ay / rhoa; }
        } else {
            ax = S.ax[a]; ay = S.ay[a];
        }
        if (mask & M_ART) { ax = ax + artx; ay = ay + arty; }
        if (mask & M_REP) { ax = ax + S.repx[a]; ay = ay + S.repy[a]; }
        if (mask & M_GRAV) ay = ay + K.gravity;
        if (mask & M_CROSS) {
            ax = ax + S.drgx[a]; ay = ay + S.drgy[a];
            ax = ax + S.porx[a]; ay = ay + S.pory[a];
        }
        S.ax[a] = ax;
        S.ay[a] = ay;
        if (mask & M_DELTA_FIELD) {
            S.dp[a] = S.dp[a] + dP;
            S.dsxx[a] = S.dsxx[a] + dXX;
            S.dsxy[a] = S.dsxy[a] + dXY;
            S.dsyy[a] = S.dsyy[a] + dYY;
        }
    }
}

// ------------------------------------------------------------------------
// Single-phase water path, single_step_for_water (:1439-1526).
// ------------------------------------------------------------------------

// sum_density (m_particles.f90:2243-2298): Shepard-normalised summation
// density from the current (old) rho; writes rho_next.
// sum_density (m_particles.f90:2243-2298) of one phase; h1 = hsml(1) of the
// phase (the link_list predicate).
__global__ __launch_bounds__(256) void k_sum_density(DPhase W, Grid g, KParams K, double h1)
{
    const int t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= W.n) return;
    const CellList cw{W.cell_start, W.order};
    const int a = cw.order[t];
    const double xa = W.x[a], ya = W.y[a], ha = W.hsml[a], ma = W.mass[a], rhoa = W.rho[a];
    const bool link = K.nnps == 2;
    double self, d0, d1;
    sph_kernel(K.skf, 0.0, 0.0, 0.0, ha, self, d0, d1);
    double wi = self * ma / rhoa;
    double rs = self * ma;
    for_each_candidate(g, cw, xa, ya, [&](int b) {
        if (b == a) return;
        const double mh = link ? h1 : sph_mhsml(ha, W.hsml[b]);
        double dx, dy, r;
        if (!sph_pair_test(xa, ya, W.x[b], W.y[b], K.scale_k * mh, dx, dy, r)) return;
        if (link && !ll_reach(g, xa, ya, W.x[b], W.y[b])) return;
        double w, dwx, dwy;
        sph_kernel(K.skf, r, dx, dy, mh, w, dwx, dwy);
        const double mb = W.mass[b];
        wi += mb / W.rho[b] * w;
        rs += mb * w;
    });
    W.rho_next[a] = rs / wi;
}

// After sum_density: drho = -rho*(df2(vx,'x')+df2(vy,'y')) + delta_sph(rho);
// t via df (:1482-1484); p = b*((rho/rho0)**gamma-1) (:1478); av_vel.
__global__ __launch_bounds__(256) void k_water1_a(DPhase W, Grid g, KParams K, int mask)
{
    const int t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= W.n) return;
    const CellList cw{W.cell_start, W.order};
    const int a = cw.order[t];
    const double xa = W.x[a], ya = W.y[a], vxa = W.vx[a], vya = W.vy[a];
    const double ha = W.hsml[a], rhoa = W.rho[a], ca = W.c[a];
    const int ita = W.itype[a];
    const bool link = K.nnps == 2;
    double d2x = 0.0, d2y = 0.0, drd = 0.0, DXvx = 0.0, DYvy = 0.0, DYvx = 0.0, DXvy = 0.0, avx = 0.0, avy = 0.0;
    int cnt = 0;
    for_each_candidate(g, cw, xa, ya, [&](int b) {
        if (b == a) return;
        const double xb = W.x[b], yb = W.y[b];
        const double mh = link ? K.h1_water : sph_mhsml(ha, W.hsml[b]);
        double dx, dy, r;
        if (!sph_pair_test(xa, ya, xb, yb, K.scale_k * mh, dx, dy, r)) return;
        if (link && !ll_reach(g, xa, ya, xb, yb)) return;
        ++cnt;
        double w, dwx, dwy;
        sph_kernel(K.skf, r, dx, dy, mh, w, dwx, dwy);
        const double mb = W.mass[b], rhob = W.rho[b], vxb = W.vx[b], vyb = W.vy[b];
        d2x += mb * ((vxb - vxa) * dwx);
        d2y += mb * ((vyb - vya) * dwy);
        if (mask & M_DRHO_DELTA) {
            const double rr = dx * dx + dy * dy;
            const double muv = 2.0 * (rhoa - rhob) / rr;
            double h = dx * muv * dwx;
            h = h + dy * muv * dwy;
            drd += K.delta * ha * ca * mb * h / rhob;
        }
        const double mr = mb / rhob;
        DXvx += mr * ((vxa + vxb) * dwx);
        DYvy += mr * ((vya + vyb) * dwy);
        DYvx += mr * ((vxa + vxb) * dwy);
        DXvy += mr * ((vya + vyb) * dwx);
        if (mask & M_XSPH) {
            const double mrho = (rhoa + rhob) / 2.0;
            avx -= mb * (vxa - vxb) / mrho * w;
            avy -= mb * (vya - vyb) / mrho * w;
        }
    });
    {
#pragma clang fp contract(off)
        const double a2x = d2x / rhoa, a2y = d2y / rhoa;
        W.drho[a] = -(rhoa * (a2x + a2y)) + drd;
        W.txx[a] = SPH_C23_F * (2.0 * DXvx - DYvy);
        W.txy[a] = DYvx + DXvy;
        W.tyy[a] = SPH_C23_F * (2.0 * DYvy - DXvx);
        if (rhoa > 0.0) W.p[a] = K.w_b * (pow(rhoa / K.w_rho0, K.w_gamma) - 1.0);
        if (mask & M_XSPH) {
            if (ita > 0) { avx = K.epsilon * avx; avy = K.epsilon * avy; }
            W.avx[a] = avx; W.avy[a] = avy;
        }
        W.countiac[a] = cnt;
    }
}

// dvx = -df(p,'x') + df(sxx,'x') + df(sxy,'y') (:1494-1498), /rho,
// art_visc, gravity.
__global__ __launch_bounds__(256) void k_water1_b(DPhase W, Grid g, KParams K, int mask)
{
    const int t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= W.n) return;
    const CellList cw{W.cell_start, W.order};
    const int a = cw.order[t];
    const double xa = W.x[a], ya = W.y[a], vxa = W.vx[a], vya = W.vy[a];
    const double ha = W.hsml[a], rhoa = W.rho[a], ca = W.c[a], pa = W.p[a];
    const double mu = K.w_mu;
    const double sxxa = mu * W.txx[a], sxya = mu * W.txy[a], syya = mu * W.tyy[a];
    const bool link = K.nnps == 2;
    double Px = 0.0, Py = 0.0, Sxx = 0.0, Sxyy = 0.0, Sxyx = 0.0, Syy = 0.0, artx = 0.0, arty = 0.0;
    for_each_candidate(g, cw, xa, ya, [&](int b) {
        if (b == a) return;
        const double xb = W.x[b], yb = W.y[b], hb = W.hsml[b];
        const double mh = link ? K.h1_water : sph_mhsml(ha, hb);
        double dx, dy, r;
        if (!sph_pair_test(xa, ya, xb, yb, K.scale_k * mh, dx, dy, r)) return;
        if (link && !ll_reach(g, xa, ya, xb, yb)) return;
        double w, dwx, dwy;
        sph_kernel(K.skf, r, dx, dy, mh, w, dwx, dwy);
        const double mb = W.mass[b], rhob = W.rho[b];
        const double mr = mb / rhob;
        const double pb = W.p[b];
        Px += mr * ((pa + pb) * dwx);
        Py += mr * ((pa + pb) * dwy);
        const double sxxb = mu * W.txx[b], sxyb = mu * W.txy[b], syyb = mu * W.tyy[b];
        Sxx += mr * ((sxxa + sxxb) * dwx);
        Sxyy += mr * ((sxya + sxyb) * dwy);
        Sxyx += mr * ((sxya + sxyb) * dwx);
        Syy += mr * ((syya + syyb) * dwy);
        if (mask & M_ART) {
            const double vr0 = (vxa - W.vx[b]) * dx;
            const double vr = vr0 + (vya - W.vy[b]) * dy;
            if (vr < 0.0) {
                const double amh = sph_mhsml(ha, hb);
                const double rr = dx * dx + dy * dy;
                const double muv = amh * vr / (rr + amh * amh * K.etq * K.etq);
                const double mc = 0.5 * (ca + W.c[b]);
                const double mrho = 0.5 * (rhoa + rhob);
                const double piv = (K.beta * muv - K.alpha * mc) * muv / mrho;
                artx += mb * (-piv * dwx);
                arty += mb * (-piv * dwy);
            }
        }
    });
    {
#pragma clang fp contract(off)
        double ax = -Px + Sxx + Sxyy;
        double ay = -Py + Sxyx + Syy;
        if (rhoa > 0.0) { ax = ax / rhoa; ay = ay / rhoa; }
        if (mask & M_ART) { ax = ax + artx; ay = ay + arty; }
        ay = ay + K.gravity;
        W.ax[a] = ax;
        W.ay[a] = ay;
    }
}

// ------------------------------------------------------------------------
// Materialised pair lists (direct_find / link_list / direct_find_2).
// Count pass: partners with a larger particle number; write pass: those
// partners in ascending particle number, i.e. the reference's i-major /
// j-ascending order, with W and grad W.
// ------------------------------------------------------------------------
__device__ __forceinline__ bool pair_pred(const DPhase& A, int a, const DPhase& B, int b, const Grid& g,
                                          const KParams& K, int link, double h1, double& dx, double& dy,
                                          double& r, double& mh)
{
    mh = link ? h1 : sph_mhsml(A.hsml[a], B.hsml[b]);
    if (!sph_pair_test(A.x[a], A.y[a], B.x[b], B.y[b], K.scale_k * mh, dx, dy, r)) return false;
    if (link && !ll_reach(g, A.x[a], A.y[a], B.x[b], B.y[b])) return false;
    return true;
}

// same == 1: partners b > a of the same phase (and countiac over all partners);
// same == 0: cross pairs, A = water, B = soil, all partners counted.
__global__ void k_pair_count(DPhase A, DPhase B, Grid g, KParams K, int same, int link, double h1,
                             int* fwd, int* countiac)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A.n) return;
    const CellList cl{B.cell_start, B.order};
    const int ida = A.id[a];
    int nf = 0, nall = 0;
    for_each_candidate(g, cl, A.x[a], A.y[a], [&](int b) {
        if (same && b == a) return;
        double dx, dy, r, mh;
        if (!pair_pred(A, a, B, b, g, K, link, h1, dx, dy, r, mh)) return;
        ++nall;
        if (!same || B.id[b] > ida) ++nf;
    });
    fwd[ida] = nf;
    if (countiac) countiac[a] = nall;
}

__global__ void k_pair_write(DPhase A, DPhase B, Grid g, KParams K, int same, int link, double h1,
                             const int* off, int* pi, int* pj, double* w, double* dwdx)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A.n) return;
    const CellList cl{B.cell_start, B.order};
    const int ida = A.id[a];
    int o = off[ida];
    const int cnt = off[ida + 1] - o;
    int prev = same ? ida : -1;
    for (int s = 0; s < cnt; ++s) {
        int best = 0x7fffffff, bestb = 0;
        for_each_candidate(g, cl, A.x[a], A.y[a], [&](int b) {
            const int idb = B.id[b];
            if (idb <= prev || idb >= best) return;
            double dx, dy, r, mh;
            if (!pair_pred(A, a, B, b, g, K, link, h1, dx, dy, r, mh)) return;
            best = idb;
            bestb = b;
        });
        double dx, dy, r, mh, ww, gx, gy;
        pair_pred(A, a, B, bestb, g, K, link, h1, dx, dy, r, mh);
        sph_kernel(K.skf, r, dx, dy, mh, ww, gx, gy);
        pi[o] = ida + 1;
        pj[o] = best + 1;
        w[o] = ww;
        dwdx[2 * o] = gx;
        dwdx[2 * o + 1] = gy;
        ++o;
        prev = best;
    }
}
