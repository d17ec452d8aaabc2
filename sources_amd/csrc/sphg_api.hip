// sphg_api.hip -- host runtime behind include/sphg.h: context, device
// buffers, cell-list build, step orchestration, granular entry points,
// host<->device sync and per-kernel event timing.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sphg.h"
#include "sph_fused.h"
#include "sph_mixed.h"
#include "sph_gnn.h"
#include "sph_batch.h"

#include <new>
#include <unordered_map>

namespace {

constexpr int TPB = 256;

// A launch held back while a batch records a step: the kernel, its grid and
// block, its arguments as the bytes of an ArgTup of its parameters, and a
// launcher of the kernel itself from those bytes.
struct LaunchRec {
    const void* kern;
    int grid, block;
    std::vector<unsigned char> args;
    hipError_t (*solo)(const void*, int, int, hipStream_t, const unsigned char*);
};

template <class... T, size_t... I>
hipError_t solo_launch_at(void (*kern)(T...), int grid, int block, hipStream_t s, const ArgTup<T...>& a,
                          std::index_sequence<I...>)
{
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, s, tup_get<I>(a)...);
    return hipGetLastError();
}

template <class... T>
hipError_t solo_launch(const void* kern, int grid, int block, hipStream_t s, const unsigned char* args)
{
    if (grid <= 0) return hipSuccess;
    return solo_launch_at(reinterpret_cast<void (*)(T...)>(const_cast<void*>(kern)), grid, block, s,
                          *reinterpret_cast<const ArgTup<T...>*>(args), std::index_sequence_for<T...>{});
}

// Skin of the neighbour lists as a fraction of the largest support radius R:
// lists hold candidates within R + skin, a particle that has moved skin / 4
// since the rebuild triggers the next one, and one that has moved skin / 2
// (when a pair could be missing) is an error.
constexpr double NL_SKIN = 0.1;    // neighbour-list skin / largest support radius
constexpr int NL_EVERY = 64;       // steps between periodic list rebuilds
constexpr double NL_LOOK = 4.0;    // steps of look-ahead of the displacement trigger
// a displacement-triggered rebuild within NL_MIN_RUN steps of the previous
// one turns the lists off for NL_OFF steps (the row scans are then cheaper)
constexpr int NL_MIN_RUN = 8;
constexpr int NL_OFF = 256;

// HIP error of the last failing call of the context-free entry point
// sphg_direct_find_noalloc (sphg_last_hip_error(NULL) returns it).
std::atomic<int> noalloc_hip_error{0};

inline int nblk(int n, int b = TPB) { return (n + b - 1) / b; }

struct ProfStat {
    int64_t launches = 0;
    double ms = 0.0;
};

struct PendingEvent {
    std::string name;
    hipEvent_t a, b;
};

// numbers of double and int arrays the reorder permutes
constexpr int REORDER_ND = 24, REORDER_NI = 3;

struct PhaseBuf {
    bool bound = false;
    sphg_phase_view view{};
    sphg_phase_view* user = nullptr;   // the caller's bound view: receives niac / nfail
    int n = 0;
    DPhase d{};
    std::vector<void*> allocs;
    bool cells_valid = false;
    bool cells_nl = false;      // the valid cell list is on the neighbour-list grid
    bool counts_zero = false;   // d.cell_count known to be all zero
    int* cell_part = nullptr;   // tile-local exclusive prefix of the cell counts
    int* tile_sums = nullptr;   // the scan's tile sums (per phase: the two builds may run concurrently)
    // the other velocity and density levels of the Verlet scheme (water)
    double *lvx = nullptr, *lvy = nullptr, *lrho = nullptr;
    // materialised pairs
    int *fwd = nullptr, *off = nullptr;
    int *pi = nullptr, *pj = nullptr;
    double *w = nullptr, *dwdx = nullptr;
    int pair_cap = 0;
    // staging for interleaved host vectors
    double* stage = nullptr;
    // fused path: cell-sorted neighbour records and same-phase masks
    double* rec = nullptr;
    Masks masks{};               // same-phase partners of water stage A, reused by stage B
    // neighbour lists of the FP64 fused step: candidate masks (same phase,
    // cross), cell key per record and position per particle at the rebuild
    Masks nl_same{}, nl_cross{};
    int* nl_cell = nullptr;
    double *nl_x = nullptr, *nl_y = nullptr;
    // second buffers of the arrays the reorder permutes: the permuted copy is
    // written there and the pointers swap
    double* shadow_d[PERM_MAXD] = {};
    int* shadow_i[PERM_MAXI] = {};
    double hmin = 0.0, hmax = 0.0;
};

}  // namespace

struct sphg_ctx {
    sphg_params P{};
    KParams K{};
    Grid g{};          // cell grid (link fields filled per phase)
    Grid gl[2]{};      // per phase copy with the link_list sorting-grid fields
    int ncell = 0;
    int itimestep = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    // side stream: the soil stage runs on it concurrently with the water
    // stages (fork and join events on the main stream)
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_wx = nullptr, ev_sx = nullptr;   // cross kernels done (water on main, soil on side)
    // SPHG_SIDE_STREAM=0 keeps every kernel on the main stream; so does the
    // per-kernel profile mode, whose durations are then not shared
    bool use_side = true;
    hipStream_t lstream = nullptr;   // stream of the next launches when set
    PhaseBuf ph[2];
    int* d_err = nullptr;       // [0] out of range
    int* d_scan = nullptr;      // tile sums
    int scan_cap = 0;
    int last_hip = 0;
    bool prof = false;
    std::map<std::string, ProfStat> stats;
    std::vector<PendingEvent> pending;
    std::vector<hipEvent_t> event_pool;
    int64_t pairs[3] = {0, 0, 0};
    FusedK F{};
    MixK X{};                   // mixed-precision constants
    double cs = 0.0;            // cell size
    double csv = 0.0;           // cell size of the neighbour-list grid
    bool legacy = false;        // SPHG_LEGACY=1: unfused reference-order kernels
    int reorder_every = 16;     // steps between permutations of the state into cell order
    int since_reorder = 0;
    // Graph mode: the step keeps every device pointer of the state where it
    // is (the reorder copies back instead of swapping buffers), so that a
    // captured sequence of steps can be replayed.
    bool stable_ptrs = false;
    // Neighbour lists of the FP64 fused two-phase step (rebuilt with the cell
    // list every reorder period, when a particle has moved skin / 4 from its
    // rebuild position, and whenever the state changed outside a step)
    Grid gv{};                  // list grid: cell size >= R + skin
    Grid gvl[2]{};              // per phase, with the link_list fields
    int ncell_v = 0;
    double skin = 0.0;
    double nl_thr[3] = {0.0, 0.0, 0.0};   // (R + skin)^2: water, soil, cross pairs
    bool nl_valid = false;      // lists, cell list and records belong to the state
    bool nl_due = false;        // the displacement check asked for a rebuild
    int nl_seq = 0;             // list-mode steps issued (coarse steps of K dt count once)
    int nl_rebuilt_at = 0;      // nl_seq of the step of the last rebuild
    int64_t nl_rebuilds = 0;
    bool use_nl = true;         // SPHG_NLIST=0: scan the cell rows every step instead
    double nl_skin = NL_SKIN;   // skin / largest support radius (SPHG_NL_SKIN)
    int nl_every = NL_EVERY;    // steps between periodic rebuilds (SPHG_NL_EVERY)
    // steps [nl_off_from, nl_off_until) run without lists (fast-moving
    // particles); a spell starts at the step after the one that decides it,
    // so that a step's kernels all see one mode
    int nl_off_from = 0, nl_off_until = 0;
    int64_t nl_off_spells = 0;
    bool capturing = false;     // steps are being captured into a graph (no host reads)
    // while set, launch() records the launches here instead of issuing them
    std::vector<LaunchRec>* recording = nullptr;
    int last_k = 1;             // the last step was of size last_k dt (a learned coarse step: K)
    // the final kernels of the step before also packed this step (first_half
    // applied to the state, records written): the step starts at its stages
    bool packed = false;
    int* d_nlflag = nullptr;    // bit 0: rebuild due; bit 1: a particle moved past skin / 2
    int* h_nlflag = nullptr;    // pinned: two slots for the lagged check of sphg_time_integration
    int h_nlstep[2] = {-1, -1}; // nl_seq of the step whose flag each slot holds (-1: none)
    hipEvent_t ev_nl[2] = {nullptr, nullptr};
};

namespace {

#define HIPCHK(ctx, expr)                                  \
    do {                                                   \
        hipError_t e_ = (expr);                            \
        if (e_ != hipSuccess) {                            \
            (ctx)->last_hip = (int)e_;                     \
            return SPHG_ERR_HIP;                           \
        }                                                  \
    } while (0)

#define RC(expr)                  \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != SPHG_OK) return rc_; \
    } while (0)

hipEvent_t get_event(sphg_ctx* c)
{
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void drain_events(sphg_ctx* c)
{
    if (c->pending.empty()) return;
    hipStreamSynchronize(c->stream);
    for (auto& pe : c->pending) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, pe.a, pe.b);
        auto& s = c->stats[pe.name];
        s.launches += 1;
        s.ms += ms;
        c->event_pool.push_back(pe.a);
        c->event_pool.push_back(pe.b);
    }
    c->pending.clear();
}

// Launch helper: brackets the launch with HIP events when profiling is on.
template <typename... KArgs, typename... Args>
int launch(sphg_ctx* c, const char* name, void (*kern)(KArgs...), int grid, int block, Args... args)
{
    if (c->recording) {
        // (empty grids too: every case of a batch records the same sequence)
        LaunchRec r{reinterpret_cast<const void*>(kern), grid, block, {}, &solo_launch<KArgs...>};
        r.args.assign(sizeof(ArgTup<KArgs...>), 0);
        new (r.args.data()) ArgTup<KArgs...>(args...);
        c->recording->push_back(std::move(r));
        return SPHG_OK;
    }
    if (grid <= 0) return SPHG_OK;
    hipEvent_t ea = nullptr, eb = nullptr;
    hipStream_t st = c->lstream ? c->lstream : c->stream;
    if (c->prof) {
        ea = get_event(c);
        eb = get_event(c);
        hipEventRecord(ea, st);
    }
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, st, args...);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        c->last_hip = (int)e;
        return SPHG_ERR_HIP;
    }
    if (c->prof) {
        hipEventRecord(eb, st);
        c->pending.push_back({name, ea, eb});
        if (c->pending.size() > 4096) drain_events(c);
    }
    return SPHG_OK;
}

// The argument tuple of kernel kern (arguments converted to its parameters).
template <class... K, class... A>
ArgTup<K...> args_of(void (*)(K...), const A&... a)
{
    return ArgTup<K...>(a...);
}

// Kernel body F over two argument sets in one launch (the two phases): nb0
// blocks of the first set, then nb1 of the second.
template <class F, class... K>
int launch_pair(sphg_ctx* c, const char* name, int nb0, const ArgTup<K...>& a0, int nb1, const ArgTup<K...>& a1)
{
    return launch(c, name, k_pair<F, K...>, nb0 + nb1, TPB, a0, a1, nb0);
}

template <class T>
int dalloc(sphg_ctx* c, PhaseBuf& b, T** p, size_t count)
{
    void* q = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) {
        c->last_hip = (int)e;
        return SPHG_ERR_OUT_OF_MEMORY;
    }
    b.allocs.push_back(q);
    *p = (T*)q;
    return SPHG_OK;
}

void free_phase(PhaseBuf& b)
{
    for (void* p : b.allocs) hipFree(p);
    b.allocs.clear();
    b = PhaseBuf{};
}

int fill_dbl(sphg_ctx* c, double* a, int n, double v)
{
    return launch(c, "fill", k_fill_dbl, nblk(n), TPB, a, n, v);
}

int fill_int(sphg_ctx* c, int* a, int n, int v)
{
    return launch(c, "fill", k_fill_int, nblk(n), TPB, a, n, v);
}

// Exclusive scan of n ints into out[0..n], total at out[n]; clear (the
// input array or null) is zeroed as it is read.
int exclusive_scan(sphg_ctx* c, const int* in, int* out, int n, int* clear = nullptr)
{
    const int ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (ntiles > SCAN_TILE) return SPHG_ERR_UNSUPPORTED;
    RC(launch(c, "scan", k_scan_tile, ntiles > 0 ? ntiles : 1, SCAN_BLOCK, in, out, n, c->d_scan, clear));
    RC(launch(c, "scan", k_scan_sums, 1, SCAN_BLOCK, c->d_scan, ntiles > 0 ? ntiles : 1));
    RC(launch(c, "scan", k_scan_add, nblk(n > 0 ? n : 1), TPB, out, n, (const int*)c->d_scan, ntiles > 0 ? ntiles : 1));
    return SPHG_OK;
}

int phase_index(int phase) { return phase == SPHG_WATER ? 0 : (phase == SPHG_SOIL ? 1 : -1); }

// Cell list of one phase for the current positions, on the per-step grid
// (nl = false) or on the neighbour-list grid.  A per-step build replaces the
// order the neighbour lists index, so it retires them.
int build_cells(sphg_ctx* c, int k, bool nl = false)
{
    PhaseBuf& b = c->ph[k];
    if (b.cells_valid && b.cells_nl == nl) return SPHG_OK;
    if (!nl) c->nl_valid = false;
    const Grid& g = nl ? c->gvl[k] : c->gl[k];
    const int ncell = nl ? c->ncell_v : c->ncell;
    // cell_count is all zero between builds: the scan clears what it reads
    if (!b.counts_zero) RC(fill_int(c, b.d.cell_count, std::max(c->ncell, c->ncell_v), 0));
    b.counts_zero = false;
    RC(launch(c, "cell_key", k_cell_key, nblk(b.n), TPB, b.d, g, c->d_err));
    // one scan pass: tile-local prefixes and tile sums; the scatter and the
    // per-cell sort add the prefix of the tile sums themselves
    const int ntiles = std::max((ncell + SCAN_TILE - 1) / SCAN_TILE, 1);
    if (ntiles > SCAN_TILE) return SPHG_ERR_UNSUPPORTED;
    RC(launch(c, "scan", k_scan_tile, ntiles, SCAN_BLOCK, (const int*)b.d.cell_count, b.cell_part, ncell, b.tile_sums,
              b.d.cell_count));
    b.counts_zero = true;
    RC(launch(c, "cell_scatter", k_cell_scatter_t, nblk(b.n, SCAN_BLOCK), SCAN_BLOCK, b.d, (const int*)b.cell_part,
              (const int*)b.tile_sums, ntiles));
    RC(launch(c, "cell_sort", k_cell_sort_t, nblk(ncell, SCAN_BLOCK), SCAN_BLOCK, b.d, (const int*)b.cell_part,
              (const int*)b.tile_sums, ntiles, ncell));
    b.cells_valid = true;
    b.cells_nl = nl;
    return SPHG_OK;
}

// The state changed outside a step (upload, restore, correction): cell lists
// and neighbour lists no longer describe it.
void state_changed(sphg_ctx* c)
{
    for (auto& pb : c->ph) pb.cells_valid = false;
    c->nl_valid = false;
    c->packed = false;
    c->last_k = 1;   // the rates the state holds count as those of a plain step
}

// Permute the persistent state of phase k into the current cell order, so
// that slot t holds the t-th particle of the order (order becomes the
// identity and the cell list stays valid).
int reorder(sphg_ctx* c, int k)
{
    PhaseBuf& b = c->ph[k];
    DPhase& d = b.d;
    const int* ord = d.order;
    int** fi[] = {&d.id, &d.itype, &d.fail};
    // av persists across steps when average_velocity is off; the stress
    // levels of stress_integration = 2 persist too
    double** f[] = {&d.x, &d.y, &d.vx, &d.vy, &d.mass, &d.hsml, &d.rho, &d.vof, &d.p, &d.c,
                    &d.drho, &d.ax, &d.ay, &d.dvof, &d.sxx, &d.sxy, &d.syy, &d.eps_p, &d.avx, &d.avy,
                    &d.sxx_min, &d.sxy_min, &d.syy_min, &d.p_min};
    constexpr int ND = (int)(sizeof(f) / sizeof(f[0])), NI = (int)(sizeof(fi) / sizeof(fi[0]));
    static_assert(ND == REORDER_ND && NI == REORDER_NI, "reorder array lists and shadow counts differ");
    // one pass gathers every array into its shadow
    PermArgs A{};
    A.nd = ND;
    A.ni = NI;
    for (int a = 0; a < ND; ++a) { A.sd[a] = *f[a]; A.dd[a] = b.shadow_d[a]; }
    for (int a = 0; a < NI; ++a) { A.si[a] = *fi[a]; A.di[a] = b.shadow_i[a]; }
    RC(launch(c, "permute", k_permute_multi, nblk(b.n), TPB, A, ord, b.n));
    if (c->stable_ptrs) {
        // captured graphs hold the arrays' addresses: copy back instead
        PermArgs B{};
        B.nd = ND;
        B.ni = NI;
        for (int a = 0; a < ND; ++a) { B.sd[a] = b.shadow_d[a]; B.dd[a] = *f[a]; }
        for (int a = 0; a < NI; ++a) { B.si[a] = b.shadow_i[a]; B.di[a] = *fi[a]; }
        RC(launch(c, "permute", k_permute_multi, nblk(b.n), TPB, B, (const int*)nullptr, b.n));
    } else {
        for (int a = 0; a < ND; ++a) std::swap(*f[a], b.shadow_d[a]);
        for (int a = 0; a < NI; ++a) std::swap(*fi[a], b.shadow_i[a]);
    }
    RC(launch(c, "iota", k_iota, nblk(b.n), TPB, d.order, b.n));
    return SPHG_OK;
}

// The device flags after the work issued so far: d_err[0] (out of range)
// and d_nlflag = d_err[1] (neighbour-list displacement).  A raised error
// flag is cleared, so that the next call starts clean.
int check_flags(sphg_ctx* c)
{
    int f[2] = {0, 0};
    HIPCHK(c, hipMemcpyAsync(f, c->d_err, sizeof(f), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->h_nlstep[0] = c->h_nlstep[1] = -1;   // the lagged copies are superseded
    if (f[0] & 1) {
        const int z = 0;
        HIPCHK(c, hipMemcpy(c->d_err, &z, sizeof(int), hipMemcpyHostToDevice));
        return SPHG_ERR_OUT_OF_RANGE;
    }
    if (f[1] & 2) {
        const int z = 0;
        HIPCHK(c, hipMemcpy(c->d_nlflag, &z, sizeof(int), hipMemcpyHostToDevice));
        c->nl_valid = false;
        return SPHG_ERR_MOVED_TOO_FAR;
    }
    return SPHG_OK;
}

void set_kparams(sphg_ctx* c)
{
    const sphg_params& P = c->P;
    KParams& K = c->K;
    K.skf = P.skf;
    K.nnps = P.nnps;
    const int sk = (P.skf == 1 || P.skf == 4) ? 2 : ((P.skf == 2 || P.skf == 3) ? 3 : 0);
    K.scale_k = (double)sk;
    K.alpha = P.alpha; K.beta = P.beta; K.etq = P.etq;
    K.dd = P.dd; K.p1 = P.p1; K.p2 = P.p2;
    auto small_int = [](double v) { return (v >= 0.0 && v <= 64.0 && v == (double)(int)v) ? (int)v : -1; };
    K.mc_ccos = P.soil_cohesion * std::cos(P.soil_phi);
    K.mc_sin = std::sin(P.soil_phi);
    K.mc_pcut = -P.soil_cohesion * std::pow(std::tan(P.soil_phi), -1.0);
    K.stress_integration = P.stress_integration;
    K.summation_density = P.summation_density;
    K.ip1 = small_int(P.p1);
    K.ip2 = small_int(P.p2);
    K.delta = P.delta; K.epsilon = P.epsilon;
    K.gravity = P.gravity; K.numeric_gravity = P.numeric_gravity;
    K.w_rho0 = P.water_rho0; K.w_b = P.water_b; K.w_gamma = P.water_gamma; K.w_c = P.water_c;
    K.w_mu = P.water_viscosity;
    K.s_rho0 = P.soil_rho0; K.s_k = P.soil_k; K.s_c = P.soil_c;
    K.s_G = P.soil_E / (2.0 * (1.0 + P.soil_niu));                       // m_particles.f90:3061
    const double t = std::tan(P.soil_phi);
    K.s_alpha1 = t / std::sqrt(9.0 + 12.0 * (t * t));                   // m_particles.f90:3337
    K.s_perm = P.soil_permeability;
    K.dt = P.dt;
    K.artificial_density = P.artificial_density;
    K.visc_artificial = P.visc_artificial;
    K.average_velocity = P.average_velocity;
    K.soil_artificial_stress = P.soil_artificial_stress;
    K.water_artificial_volume = P.water_artificial_volume;
    K.volume_fraction = P.volume_fraction;
    K.plasticity = P.plasticity;
    K.yield_criterion = P.yield_criterion;
    K.two_phase = P.single_phase ? 0 : 1;
}

int check_params(const sphg_params& P)
{
    if (!(P.skf >= 1 && P.skf <= 4)) return SPHG_ERR_BAD_KERNEL;
    if (P.nnps != 1 && P.nnps != 2) return SPHG_ERR_UNSUPPORTED;
    if (!(P.x_maxgeom > P.x_mingeom && P.y_maxgeom > P.y_mingeom)) return SPHG_ERR_INVALID;
    if (P.precision != SPHG_FP64 && P.precision != SPHG_MIXED) return SPHG_ERR_INVALID;
    if (P.integrate_scheme < 0 || P.integrate_scheme > 2) return SPHG_ERR_INVALID;
    return SPHG_OK;
}

// Smallest T with sqrt(T) >= R: for every dr >= 0, dr < T <=> sqrt(dr) < R,
// because the correctly rounded sqrt is monotone.  The pair predicate
// sqrt(dx*dx+dy*dy) < scale_k*h then costs one compare.
double dr_threshold(double R)
{
    double T = R * R;
    while (std::sqrt(T) < R) T = std::nextafter(T, INFINITY);
    for (;;) {
        const double lo = std::nextafter(T, 0.0);
        if (std::sqrt(lo) >= R) T = lo;
        else break;
    }
    return T;
}

// Cell grid over the geometry box, cell size >= the largest support radius.
int setup_grid(sphg_ctx* c)
{
    const sphg_params& P = c->P;
    double hmax = 0.0;
    for (int k = 0; k < 2; ++k) {
        PhaseBuf& b = c->ph[k];
        if (!b.bound) continue;
        for (int i = 0; i < b.n; ++i) hmax = std::fmax(hmax, b.view.hsml[i]);
    }
    if (!(hmax > 0.0)) return SPHG_ERR_INVALID;
    const double cs = c->K.scale_k * hmax * (1.0 + 1.0e-9);
    Grid g{};
    g.x0 = P.x_mingeom; g.y0 = P.y_mingeom; g.x1 = P.x_maxgeom; g.y1 = P.y_maxgeom;
    g.inv_cs = 1.0 / cs;
    const double fx = std::floor((P.x_maxgeom - P.x_mingeom) / cs) + 1.0;
    const double fy = std::floor((P.y_maxgeom - P.y_mingeom) / cs) + 1.0;
    if (fx * fy > 6.0e7) return SPHG_ERR_UNSUPPORTED;
    g.ncx = (int)fx;
    g.ncy = (int)fy;
    c->g = g;
    c->cs = cs;
    const int ncell = g.ncx * g.ncy;
    // neighbour-list grid: cells of at least the largest support radius plus
    // the skin, so that the lists' candidates within R + skin lie in the
    // 3 x 3 block of a particle's cell
    c->skin = c->nl_skin * c->K.scale_k * hmax;
    {
        const double csv = (c->K.scale_k * hmax + c->skin) * (1.0 + 1.0e-9);
        Grid v = g;
        v.inv_cs = 1.0 / csv;
        v.ncx = (int)(std::floor((P.x_maxgeom - P.x_mingeom) / csv) + 1.0);
        v.ncy = (int)(std::floor((P.y_maxgeom - P.y_mingeom) / csv) + 1.0);
        c->gv = v;
        c->csv = csv;
        c->ncell_v = v.ncx * v.ncy;
    }
    // per phase: reference link_list sorting grid (init_grid, m_particles.f90:1069-1113)
    for (int k = 0; k < 2; ++k) {
        Grid gk = g;
        PhaseBuf& b = c->ph[k];
        gk.link_list = (P.nnps == 2) ? 1 : 0;
        if (b.bound && b.n > 0) {
            const double dgx = P.x_maxgeom - P.x_mingeom, dgy = P.y_maxgeom - P.y_mingeom;
            int ngx = (int)std::sqrt(b.n * dgx / (dgy * 3.0)) + 1;
            if (ngx > P.maxngx) ngx = P.maxngx;
            int ngy = (int)(ngx * dgy / dgx) + 1;
            if (ngy > P.maxngy) ngy = P.maxngy;
            const double h1 = b.view.hsml[0];
            gk.lreach_x = (int)((double)(float)ngx * h1 / dgx) + 1;
            gk.lreach_y = (int)((double)(float)ngy * h1 / dgy) + 1;
            gk.lmin_x = P.x_mingeom;
            gk.lmin_y = P.y_mingeom;
            gk.lscale_x = (double)(float)ngx / dgx;
            gk.lscale_y = (double)(float)ngy / dgy;
            if (k == 0) c->K.h1_water = h1; else c->K.h1_soil = h1;
        }
        c->gl[k] = gk;
        Grid gvk = c->gv;   // the same link_list fields on the list grid
        gvk.link_list = gk.link_list;
        gvk.lmin_x = gk.lmin_x; gvk.lmin_y = gk.lmin_y;
        gvk.lscale_x = gk.lscale_x; gvk.lscale_y = gk.lscale_y;
        gvk.lreach_x = gk.lreach_x; gvk.lreach_y = gk.lreach_y;
        c->gvl[k] = gvk;
    }
    {
        // list radii: R + skin per pair type, R from the largest h
        const double sk = c->K.scale_k;
        const double hw = c->ph[0].bound ? c->ph[0].hmax : 0.0, hs = c->ph[1].bound ? c->ph[1].hmax : 0.0;
        const double rw = sk * hw + c->skin, rs = sk * hs + c->skin, rx = sk * 0.5 * (hw + hs) + c->skin;
        c->nl_thr[0] = rw * rw;
        c->nl_thr[1] = rs * rs;
        c->nl_thr[2] = rx * rx;
    }
    c->nl_valid = false;
    // sqrt-free pair predicate thresholds (uniform smoothing length)
    {
        FusedK& F = c->F;
        const PhaseBuf& W = c->ph[0];
        const PhaseBuf& S = c->ph[1];
        const double sk = c->K.scale_k;
        F.uni_w = (P.nnps == 2) || (W.bound && W.hmin == W.hmax);
        F.uni_s = (P.nnps == 2) || (S.bound && S.hmin == S.hmax);
        F.uni_x = W.bound && S.bound && W.hmin == W.hmax && S.hmin == S.hmax;
        F.thr_w = W.bound ? dr_threshold(sk * W.view.hsml[0]) : 0.0;
        F.thr_s = S.bound ? dr_threshold(sk * S.view.hsml[0]) : 0.0;
        F.thr_x = (W.bound && S.bound) ? dr_threshold(sk * ((W.view.hsml[0] + S.view.hsml[0]) / 2.0)) : 0.0;
        MixK& X = c->X;
        X.cs = (float)cs;
        X.thr_w = (float)F.thr_w;
        X.thr_s = (float)F.thr_s;
        X.thr_x = (float)F.thr_x;
        X.scale_k = (float)sk;
        X.rho0_w = (float)c->K.w_rho0;
        X.rho0_s = (float)c->K.s_rho0;
    }
    // (re)allocate cell arrays
    if (ncell != c->ncell) {
        c->ncell = ncell;
        for (int k = 0; k < 2; ++k) {
            PhaseBuf& b = c->ph[k];
            if (!b.bound) continue;
            RC(dalloc(c, b, &b.d.cell_count, (size_t)ncell));
            RC(dalloc(c, b, &b.cell_part, (size_t)ncell));
            RC(dalloc(c, b, &b.tile_sums, (size_t)(ncell + SCAN_TILE - 1) / SCAN_TILE + 2));
            b.counts_zero = false;
            RC(dalloc(c, b, &b.d.cell_start, (size_t)ncell + 1));
            b.cells_valid = false;
        }
        const int ntiles = (ncell + SCAN_TILE - 1) / SCAN_TILE;
        int need = ntiles + 1;
        for (int k = 0; k < 2; ++k) need = std::max(need, (c->ph[k].n + SCAN_TILE - 1) / SCAN_TILE + 1);
        if (need > c->scan_cap) {
            if (c->d_scan) hipFree(c->d_scan);
            HIPCHK(c, hipMalloc(&c->d_scan, sizeof(int) * (size_t)(need + 1)));
            c->scan_cap = need;
        }
    }
    return SPHG_OK;
}

// ---- host <-> device field transfer ----------------------------------------

// Host arrays are in reference particle order; device slot t holds particle
// id[t].  Uploads gather by id, downloads scatter by id.
__global__ void k_split2(const double* in, double* a, double* b, const int* id, int n)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) { const int i = id[t]; a[t] = in[2 * i]; b[t] = in[2 * i + 1]; }
}

__global__ void k_join2(double* out, const double* a, const double* b, const int* id, int n)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) { const int i = id[t]; out[2 * i] = a[t]; out[2 * i + 1] = b[t]; }
}

int up_scalar(sphg_ctx* c, PhaseBuf& b, const double* h, double* d)
{
    if (!h) return SPHG_OK;
    HIPCHK(c, hipMemcpyAsync(b.stage, h, sizeof(double) * (size_t)b.n, hipMemcpyHostToDevice, c->stream));
    return launch(c, "gather", k_gather<double>, nblk(b.n), TPB, (const double*)b.stage, d, (const int*)b.d.id, b.n);
}

int up_int(sphg_ctx* c, PhaseBuf& b, const int32_t* h, int* d)
{
    if (!h) return SPHG_OK;
    int* st = reinterpret_cast<int*>(b.stage);
    HIPCHK(c, hipMemcpyAsync(st, h, sizeof(int) * (size_t)b.n, hipMemcpyHostToDevice, c->stream));
    return launch(c, "gather", k_gather<int>, nblk(b.n), TPB, (const int*)st, d, (const int*)b.d.id, b.n);
}

int up_vec(sphg_ctx* c, PhaseBuf& b, const double* h, double* dx, double* dy)
{
    if (!h) return SPHG_OK;
    HIPCHK(c, hipMemcpyAsync(b.stage, h, sizeof(double) * 2 * (size_t)b.n, hipMemcpyHostToDevice, c->stream));
    return launch(c, "split", k_split2, nblk(b.n), TPB, (const double*)b.stage, dx, dy, (const int*)b.d.id, b.n);
}

int down_scalar(sphg_ctx* c, PhaseBuf& b, double* h, const double* d)
{
    if (!h) return SPHG_OK;
    RC(launch(c, "scatter", k_scatter<double>, nblk(b.n), TPB, d, b.stage, (const int*)b.d.id, b.n));
    HIPCHK(c, hipMemcpyAsync(h, b.stage, sizeof(double) * (size_t)b.n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SPHG_OK;
}

int down_int(sphg_ctx* c, PhaseBuf& b, int32_t* h, const int* d)
{
    if (!h) return SPHG_OK;
    int* st = reinterpret_cast<int*>(b.stage);
    RC(launch(c, "scatter", k_scatter<int>, nblk(b.n), TPB, d, st, (const int*)b.d.id, b.n));
    HIPCHK(c, hipMemcpyAsync(h, st, sizeof(int) * (size_t)b.n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SPHG_OK;
}

int down_vec(sphg_ctx* c, PhaseBuf& b, double* h, const double* dx, const double* dy)
{
    if (!h) return SPHG_OK;
    RC(launch(c, "join", k_join2, nblk(b.n), TPB, b.stage, dx, dy, (const int*)b.d.id, b.n));
    HIPCHK(c, hipMemcpyAsync(h, b.stage, sizeof(double) * 2 * (size_t)b.n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SPHG_OK;
}

int upload(sphg_ctx* c, int k, uint64_t m)
{
    PhaseBuf& b = c->ph[k];
    const sphg_phase_view& v = b.view;
    DPhase& d = b.d;
    if (m & SPHG_F_X) { RC(up_vec(c, b, v.x, d.x, d.y)); state_changed(c); }
    if (m & SPHG_F_VX) RC(up_vec(c, b, v.vx, d.vx, d.vy));
    if (m & SPHG_F_RHO) RC(up_scalar(c, b, v.rho, d.rho));
    if (m & SPHG_F_P) RC(up_scalar(c, b, v.p, d.p));
    if (m & SPHG_F_C) RC(up_scalar(c, b, v.c, d.c));
    if (m & SPHG_F_VOF) RC(up_scalar(c, b, v.vof, d.vof));
    if (m & SPHG_F_STRESS) {
        if (b.d.material == SPHG_SOIL) {
            RC(up_scalar(c, b, v.sxx, d.sxx));
            RC(up_scalar(c, b, v.sxy, d.sxy));
            RC(up_scalar(c, b, v.syy, d.syy));
        }
    }
    if (m & SPHG_F_EPSP) RC(up_scalar(c, b, v.epsilon_p, d.eps_p));
    if (m & SPHG_F_DRHO) RC(up_scalar(c, b, v.drho, d.drho));
    if (m & SPHG_F_DVX) RC(up_vec(c, b, v.dvx, d.ax, d.ay));
    if (m & SPHG_F_DVOF) RC(up_scalar(c, b, v.dvof, d.dvof));
    if (m & SPHG_F_DSTRESS) {
        RC(up_scalar(c, b, v.dsxx, d.dsxx));
        RC(up_scalar(c, b, v.dsxy, d.dsxy));
        RC(up_scalar(c, b, v.dsyy, d.dsyy));
        RC(up_scalar(c, b, v.dp, d.dp));
    }
    if (m & SPHG_F_AV) RC(up_vec(c, b, v.av, d.avx, d.avy));
    if (m & SPHG_F_STRAIN) {
        RC(up_scalar(c, b, v.txx, d.txx));
        RC(up_scalar(c, b, v.txy, d.txy));
        RC(up_scalar(c, b, v.tyy, d.tyy));
        RC(up_scalar(c, b, v.vcc, d.vcc));
        RC(up_scalar(c, b, v.wxy, d.wxy));
    }
    if (m & SPHG_F_VOF2) RC(up_scalar(c, b, v.vof2, d.vof2));
    if (m & SPHG_F_FAIL) RC(up_int(c, b, v.fail, d.fail));
    if (m & SPHG_F_MIN) {
        RC(up_scalar(c, b, v.rho_min, d.rho_min));
        RC(up_vec(c, b, v.v_min, d.vx_min, d.vy_min));
        RC(up_scalar(c, b, v.vof_min, d.vof_min));
        if (b.d.material == SPHG_SOIL) {
            RC(up_scalar(c, b, v.sxx_min, d.sxx_min));
            RC(up_scalar(c, b, v.sxy_min, d.sxy_min));
            RC(up_scalar(c, b, v.syy_min, d.syy_min));
            RC(up_scalar(c, b, v.p_min, d.p_min));
        }
    }
    if ((m & SPHG_F_DSTRESS) && b.d.material == SPHG_SOIL) {
        RC(up_scalar(c, b, v.dsxx2, d.dsxx2));
        RC(up_scalar(c, b, v.dsxy2, d.dsxy2));
        RC(up_scalar(c, b, v.dsyy2, d.dsyy2));
        RC(up_scalar(c, b, v.dp2, d.dp2));
    }
    return SPHG_OK;
}

int download(sphg_ctx* c, int k, uint64_t m)
{
    PhaseBuf& b = c->ph[k];
    sphg_phase_view& v = b.view;
    DPhase& d = b.d;
    if (m & SPHG_F_X) RC(down_vec(c, b, v.x, d.x, d.y));
    if (m & SPHG_F_VX) RC(down_vec(c, b, v.vx, d.vx, d.vy));
    if (m & SPHG_F_RHO) RC(down_scalar(c, b, v.rho, d.rho));
    if (m & SPHG_F_P) RC(down_scalar(c, b, v.p, d.p));
    if (m & SPHG_F_C) RC(down_scalar(c, b, v.c, d.c));
    if (m & SPHG_F_VOF) RC(down_scalar(c, b, v.vof, d.vof));
    if ((m & SPHG_F_STRESS) && b.d.material == SPHG_SOIL) {
        RC(down_scalar(c, b, v.sxx, d.sxx));
        RC(down_scalar(c, b, v.sxy, d.sxy));
        RC(down_scalar(c, b, v.syy, d.syy));
    }
    if (m & SPHG_F_EPSP) RC(down_scalar(c, b, v.epsilon_p, d.eps_p));
    if (m & SPHG_F_DRHO) RC(down_scalar(c, b, v.drho, d.drho));
    if (m & SPHG_F_DVX) RC(down_vec(c, b, v.dvx, d.ax, d.ay));
    if (m & SPHG_F_DVOF) RC(down_scalar(c, b, v.dvof, d.dvof));
    if (m & SPHG_F_DSTRESS) {
        RC(down_scalar(c, b, v.dsxx, d.dsxx));
        RC(down_scalar(c, b, v.dsxy, d.dsxy));
        RC(down_scalar(c, b, v.dsyy, d.dsyy));
        RC(down_scalar(c, b, v.dp, d.dp));
    }
    if (m & SPHG_F_AV) RC(down_vec(c, b, v.av, d.avx, d.avy));
    if (m & (SPHG_F_STRAIN | SPHG_F_STRESS)) {
        RC(down_scalar(c, b, v.txx, d.txx));
        RC(down_scalar(c, b, v.txy, d.txy));
        RC(down_scalar(c, b, v.tyy, d.tyy));
        if (m & SPHG_F_STRAIN) {
            RC(down_scalar(c, b, v.vcc, d.vcc));
            RC(down_scalar(c, b, v.wxy, d.wxy));
        }
    }
    if (m & SPHG_F_VOF2) RC(down_scalar(c, b, v.vof2, d.vof2));
    if (m & SPHG_F_FAIL) RC(down_int(c, b, v.fail, d.fail));
    if (m & SPHG_F_COUNTIAC) RC(down_int(c, b, v.countiac, d.countiac));
    if (m & SPHG_F_MIN) {
        RC(down_scalar(c, b, v.rho_min, d.rho_min));
        RC(down_vec(c, b, v.v_min, d.vx_min, d.vy_min));
        RC(down_scalar(c, b, v.vof_min, d.vof_min));
        if (b.d.material == SPHG_SOIL) {
            RC(down_scalar(c, b, v.sxx_min, d.sxx_min));
            RC(down_scalar(c, b, v.sxy_min, d.sxy_min));
            RC(down_scalar(c, b, v.syy_min, d.syy_min));
            RC(down_scalar(c, b, v.p_min, d.p_min));
        }
    }
    if ((m & SPHG_F_DSTRESS) && b.d.material == SPHG_SOIL) {
        RC(down_scalar(c, b, v.dsxx2, d.dsxx2));
        RC(down_scalar(c, b, v.dsxy2, d.dsxy2));
        RC(down_scalar(c, b, v.dsyy2, d.dsyy2));
        RC(down_scalar(c, b, v.dp2, d.dp2));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // Water stress is sigma = mu * t (newtonian_fluid, m_particles.f90:2496-2498).
    if ((m & SPHG_F_STRESS) && b.d.material == SPHG_WATER && v.txx) {
        const double mu = c->P.water_viscosity;
        for (int i = 0; i < b.n; ++i) {
            if (v.sxx) v.sxx[i] = mu * v.txx[i];
            if (v.syy) v.syy[i] = mu * v.tyy[i];
            if (v.sxy) v.sxy[i] = mu * v.txy[i];
        }
    }
    if ((m & SPHG_F_FAIL) && v.fail) {
        int nf = 0;
        for (int i = 0; i < b.n; ++i) nf += v.fail[i] == 1;
        v.nfail = nf;
        if (b.user) b.user->nfail = nf;
    }
    return SPHG_OK;
}

int alloc_phase(sphg_ctx* c, int k, const sphg_phase_view* v)
{
    PhaseBuf& b = c->ph[k];
    free_phase(b);
    b.view = *v;
    b.n = v->ntotal + v->nvirt;
    b.bound = true;
    DPhase& d = b.d;
    d.n = b.n;
    d.ntotal = v->ntotal;
    d.material = v->material;
    d.dspp = v->dspp;
    const size_t n = (size_t)b.n;
    RC(dalloc(c, b, &d.id, n));
    RC(launch(c, "iota", k_iota, nblk(b.n), TPB, d.id, b.n));
    // shadows of the reordered arrays (allocated here, never inside a step, so
    // that a step can be captured into a graph)
    for (int a = 0; a < REORDER_ND; ++a) RC(dalloc(c, b, &b.shadow_d[a], n));
    for (int a = 0; a < REORDER_NI; ++a) RC(dalloc(c, b, &b.shadow_i[a], n));
    RC(dalloc(c, b, &d.itype, n));
    RC(dalloc(c, b, &d.fail, n));
    RC(dalloc(c, b, &d.countiac, n));
    RC(dalloc(c, b, &d.nsame, n));
    RC(dalloc(c, b, &d.key, n));
    RC(dalloc(c, b, &d.slot, n));
    RC(dalloc(c, b, &d.order, n));
    double** dbls[] = {&d.x, &d.y, &d.vx, &d.vy, &d.mass, &d.hsml, &d.rho, &d.rho_min, &d.rho_next,
                       &d.vof, &d.vof_min, &d.vof2, &d.dvof, &d.dvof_x, &d.p, &d.c, &d.c_next,
                       &d.drho, &d.ax, &d.ay, &d.avx, &d.avy, &d.vx_min, &d.vy_min,
                       &d.txx, &d.txy, &d.tyy, &d.vcc, &d.wxy, &d.sxx, &d.sxy, &d.syy, &d.eps_p,
                       &d.dsxx, &d.dsxy, &d.dsyy, &d.dp, &d.repx, &d.repy, &d.drgx, &d.drgy,
                       &d.porx, &d.pory, &d.sxx_min, &d.sxy_min, &d.syy_min, &d.p_min,
                       &d.dsxx2, &d.dsxy2, &d.dsyy2, &d.dp2};
    for (double** p : dbls) {
        RC(dalloc(c, b, p, n));
        RC(fill_dbl(c, *p, b.n, 0.0));
    }
    RC(dalloc(c, b, &b.stage, 2 * n));
    RC(fill_int(c, d.itype, b.n, 0));
    RC(fill_int(c, d.fail, b.n, 0));
    RC(fill_int(c, d.countiac, b.n, 0));
    RC(fill_int(c, d.nsame, b.n, 0));
    // fused-path buffers
    RC(dalloc(c, b, &b.rec, (size_t)R_STRIDE * n));
    if (k == 0) {   // water; soil runs both stages in one pass and stores no masks
        RC(dalloc(c, b, &b.masks.m, 3 * n));
        RC(dalloc(c, b, &b.masks.ovf, n));
    }
    for (Masks* m : {&b.nl_same, &b.nl_cross}) {
        RC(dalloc(c, b, &m->m, 2 * n));
        RC(dalloc(c, b, &m->ovf, n));
    }
    RC(dalloc(c, b, &b.nl_cell, n));
    RC(dalloc(c, b, &b.nl_x, n));
    RC(dalloc(c, b, &b.nl_y, n));
    c->nl_valid = false;
    b.hmin = b.hmax = v->hsml[0];
    for (int i = 1; i < b.n; ++i) {
        b.hmin = std::fmin(b.hmin, v->hsml[i]);
        b.hmax = std::fmax(b.hmax, v->hsml[i]);
    }
    // pair-count scratch
    RC(dalloc(c, b, &b.fwd, n));
    RC(dalloc(c, b, &b.off, n + 1));
    b.cells_valid = false;
    return SPHG_OK;
}

int ensure_pairs(sphg_ctx* c, PhaseBuf& b, int cap)
{
    if (cap <= b.pair_cap) return SPHG_OK;
    RC(dalloc(c, b, &b.pi, (size_t)cap));
    RC(dalloc(c, b, &b.pj, (size_t)cap));
    RC(dalloc(c, b, &b.w, (size_t)cap));
    RC(dalloc(c, b, &b.dwdx, 2 * (size_t)cap));
    b.pair_cap = cap;
    return SPHG_OK;
}

void swap_c(PhaseBuf& b) { std::swap(b.d.c, b.d.c_next); }

// ---- fused passes -------------------------------------------------------

constexpr int MASK_A_W = M_CON | M_DRHO_DELTA | M_STRAIN | M_EOS | M_XSPH | M_REP | M_CROSS;
constexpr int MASK_A_S = M_CON | M_DRHO_DELTA | M_STRAIN | M_EOS | M_XSPH | M_REP | M_CROSS | M_PLASTIC;
constexpr int MASK_B = M_MOMENTUM | M_ART | M_DELTA_FIELD | M_REP | M_GRAV | M_CROSS;

int mask_a(const sphg_ctx* c, int base)
{
    int m = base;
    if (!c->P.artificial_density) m &= ~M_DRHO_DELTA;
    if (!c->P.average_velocity) m &= ~M_XSPH;
    if (c->P.single_phase) m &= ~M_CROSS;
    return m;
}

// Fork the side stream off the main stream (it waits for the work issued so
// far) and join it back (the main stream waits for the side work).
bool side_on(const sphg_ctx* c) { return c->use_side && !c->prof; }

int fork_side(sphg_ctx* c)
{
    HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
    HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
    return SPHG_OK;
}

int join_side(sphg_ctx* c)
{
    HIPCHK(c, hipEventRecord(c->ev_join, c->side));
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
    return SPHG_OK;
}

// Launches issued inside its scope go to the side stream (when enabled).
struct SideScope {
    sphg_ctx* c;
    explicit SideScope(sphg_ctx* ctx) : c(ctx) { if (side_on(c)) c->lstream = c->side; }
    ~SideScope() { c->lstream = nullptr; }
};

// The side stream forked off the main stream for the life of the object and
// joined back by join() or, at the latest, by the destructor -- also when an
// error returns early -- so that every fork is joined.  (In a captured graph
// an unjoined fork invalidates the capture and leaves the side stream in it.)
struct SideFork {
    sphg_ctx* c;
    bool open = false;
    int rc = SPHG_OK;
    explicit SideFork(sphg_ctx* ctx) : c(ctx)
    {
        if (side_on(c)) {
            rc = fork_side(c);
            open = rc == SPHG_OK;
        }
    }
    int join()
    {
        if (!open) return SPHG_OK;
        open = false;
        return join_side(c);
    }
    ~SideFork() { join(); }
    SideFork(const SideFork&) = delete;
    SideFork& operator=(const SideFork&) = delete;
};

int mask_b(const sphg_ctx* c, int phase)
{
    int m = MASK_B;
    if (!c->P.visc_artificial) m &= ~M_ART;
    if (phase == 0 && !c->P.water_artificial_volume) m &= ~M_DELTA_FIELD;
    if (phase == 1 && !c->P.soil_artificial_stress) m &= ~M_DELTA_FIELD;
    if (c->P.single_phase) m &= ~M_CROSS;
    return m;
}

// Fused two-phase steps keep neighbour lists with a skin (nl_config),
// except during a spell of fast motion (nl_mode).
bool nl_config(const sphg_ctx* c)
{
    return c->use_nl && !c->legacy && !c->P.single_phase;
}

bool nl_at(const sphg_ctx* c, int step)
{
    return nl_config(c) && !(step >= c->nl_off_from && step < c->nl_off_until);
}

bool nl_mode(const sphg_ctx* c) { return nl_at(c, c->itimestep); }

// single_step (declarations_sph.f90:1131-1435) after the soil Drucker-Prager
// projection of :1244 has been applied (legacy unfused kernels).
int rhs_two_phase_legacy(sphg_ctx* c)
{
    PhaseBuf& W = c->ph[0];
    PhaseBuf& S = c->ph[1];
    RC(build_cells(c, 0));
    RC(build_cells(c, 1));
    RC(launch(c, "water_a", k_water_a, nblk(W.n), TPB, W.d, S.d, c->gl[0], c->K, mask_a(c, MASK_A_W)));
    RC(launch(c, "soil_a", k_soil_a, nblk(S.n), TPB, S.d, W.d, c->gl[1], c->K, mask_a(c, MASK_A_S)));
    swap_c(W);
    swap_c(S);
    RC(launch(c, "water_b", k_water_b, nblk(W.n), TPB, W.d, c->gl[0], c->K, mask_b(c, 0)));
    RC(launch(c, "soil_b", k_soil_b, nblk(S.n), TPB, S.d, c->gl[1], c->K, mask_b(c, 1)));
    return SPHG_OK;
}

// The displacement reference of phase k for a pack (rebuild: record the
// positions).  A rebuild comes past a quarter of the skin, or when four more
// steps at the current speed would pass 0.45 skin; past half of it the run
// fails.
NRef nref(const sphg_ctx* c, int k, bool rebuild)
{
    const double hard = 0.5 * c->skin;
    return NRef{c->ph[k].nl_x, c->ph[k].nl_y, c->d_nlflag, 0.25 * c->skin, 0.45 * c->skin, hard * hard,
                NL_LOOK * c->K.dt, rebuild ? 1 : 0, c->ph[k].nl_cell};
}

// The final kernels a step's stages launch (FP64 list mode): second_half of
// the leapfrog, fused with the next step's pack when ahead.
struct FinalArgs {
    int first, vf, renorm;
    bool ahead;
};

// The fused step: pack (first_half + soil Drucker-Prager projections +
// records), cross terms, partner masks, stage A, stage B.  The soil kernels
// read only the packed records of both phases (the water pressure comes from
// the pack): they run on the side stream while the water kernels run on the
// main stream.
template <int SKF>
int stages_fused(sphg_ctx* c, const FinalArgs* fin)
{
    PhaseBuf& W = c->ph[0];
    PhaseBuf& S = c->ph[1];
    const int cross = c->P.single_phase ? 0 : 1;
    const bool rep_int = c->K.ip1 >= 0 && c->K.ip2 >= 0;
    const double* rW = W.rec;
    const double* rS = S.rec;
    const bool nl = nl_mode(c);
    const NList nW = nl ? NList{W.nl_same, W.nl_cross, W.nl_cell} : NList{};
    const NList nS = nl ? NList{S.nl_same, S.nl_cross, S.nl_cell} : NList{};
    const Grid& gW = nl ? c->gvl[0] : c->gl[0];
    const Grid& gS = nl ? c->gvl[1] : c->gl[1];
    // U instantiations: h uniform in the phase(s) the kernel reads (then
    // hsml(1) is every h), integer repulsion exponents, Drucker-Prager; the
    // general ones read h per particle and partner
    const bool us = c->ph[1].hmin == c->ph[1].hmax && rep_int && c->K.yield_criterion != 1;
    const bool uw = c->ph[0].hmin == c->ph[0].hmax && rep_int;
    const bool ux = c->F.uni_x != 0;
    SideFork fork(c);
    RC(fork.rc);
    // with the finals on both streams, each phase's final kernel waits only
    // for the other phase's cross kernel (which reads its records)
    const bool split = fin && fork.open;
    {
        SideScope side(c);
        if (cross) {
            if (ux) RC(launch(c, "soil_x", k_soil_x4<SKF, true>, nblk(S.n, STAGE_BLOCK), STAGE_BLOCK, S.d, W.d, rS, rW,
                              nS, gS, c->K, c->F));
            else RC(launch(c, "soil_x", k_soil_x4<SKF, false>, nblk(S.n, STAGE_BLOCK), STAGE_BLOCK, S.d, W.d, rS, rW,
                           nS, gS, c->K, c->F));
        }
        if (split) HIPCHK(c, hipEventRecord(c->ev_sx, c->side));
        if (us) RC(launch(c, "soil_ab", k_soil_ab4<SKF, true, false>, nblk(S.n, STAGE_BLOCK), STAGE_BLOCK, S.d, rS,
                          nS, gS, c->K, c->F, cross, mask_b(c, 1)));
        else RC(launch(c, "soil_ab", k_soil_ab4<SKF, false, false>, nblk(S.n, STAGE_BLOCK), STAGE_BLOCK, S.d, rS,
                       nS, gS, c->K, c->F, cross, mask_b(c, 1)));
    }
    // water stage A, then the water cross terms (before stage B, whose tail
    // adds them): beside the soil side's cross pass and then its stage,
    // each pair of concurrent kernels reads the same phase's records
    if (uw) RC(launch(c, "water_a", k_water_a4<SKF, true>, nblk(W.n, STAGE_BLOCK), STAGE_BLOCK, W.d, rW, W.rec,
                      nW, W.masks, gW, c->K, c->F, cross));
    else RC(launch(c, "water_a", k_water_a4<SKF, false>, nblk(W.n, STAGE_BLOCK), STAGE_BLOCK, W.d, rW, W.rec,
                   nW, W.masks, gW, c->K, c->F, cross));
    if (cross) {
        if (ux) RC(launch(c, "water_x", k_water_x4<SKF, true>, nblk(W.n, STAGE_BLOCK), STAGE_BLOCK, W.d, S.d, rW, rS,
                          nW, gW, c->K, c->F));
        else RC(launch(c, "water_x", k_water_x4<SKF, false>, nblk(W.n, STAGE_BLOCK), STAGE_BLOCK, W.d, S.d, rW, rS,
                       nW, gW, c->K, c->F));
    }
    if (split) HIPCHK(c, hipEventRecord(c->ev_wx, c->stream));
    swap_c(W);
    swap_c(S);
    if (uw) RC(launch(c, "water_b", k_water_b4<SKF, true>, nblk(W.n, STAGE_BLOCK), STAGE_BLOCK, W.d, rW, nW,
                      W.masks, gW, c->K, c->F, mask_b(c, 0)));
    else RC(launch(c, "water_b", k_water_b4<SKF, false>, nblk(W.n, STAGE_BLOCK), STAGE_BLOCK, W.d, rW, nW,
                   W.masks, gW, c->K, c->F, mask_b(c, 0)));
    if (fin) {
        auto final_water = [&]() -> int {
            if (fin->ahead)
                return launch(c, "final_water", k_step_water, nblk(W.n), TPB, W.d, c->K, W.rec, fin->first, fin->vf,
                              fin->renorm, (float*)nullptr, c->gvl[0], c->csv, nref(c, 0, false));
            return launch(c, "final_water", k_final_water, nblk(W.n), TPB, W.d, c->K, fin->first, fin->vf,
                          fin->renorm);
        };
        auto final_soil = [&]() -> int {
            if (fin->ahead)
                return launch(c, "final_soil", k_step_soil, nblk(S.n), TPB, S.d, c->K, S.rec, fin->first,
                              (float*)nullptr, c->gvl[1], c->csv, nref(c, 1, false));
            return launch(c, "final_soil", k_final_soil, nblk(S.n), TPB, S.d, c->K, fin->first);
        };
        if (split) {
            {
                SideScope side(c);
                HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_wx, 0));
                RC(final_soil());
            }
            HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_sx, 0));
            RC(final_water());
        } else {
            RC(fork.join());
            RC(final_water());
            RC(final_soil());
        }
    }
    return fork.join();
}

// Single-phase soil (single_step_for_soil, declarations_sph.f90:1530-1655):
// the one-pass soil kernel in its df form, without water partners.
template <int SKF>
int stages_soil_only(sphg_ctx* c)
{
    PhaseBuf& S = c->ph[1];
    const double* rs = S.rec;
    if (S.hmin == S.hmax && c->K.yield_criterion != 1)
        RC(launch(c, "soil_ab", k_soil_ab4<SKF, true, true>, nblk(S.n, STAGE_BLOCK), STAGE_BLOCK, S.d, rs, NList{},
                  c->gl[1], c->K, c->F, 0, mask_b(c, 1)));
    else
        RC(launch(c, "soil_ab", k_soil_ab4<SKF, false, true>, nblk(S.n, STAGE_BLOCK), STAGE_BLOCK, S.d, rs, NList{},
                  c->gl[1], c->K, c->F, 0, mask_b(c, 1)));
    swap_c(S);
    return SPHG_OK;
}

int rhs_soil_only(sphg_ctx* c, int soil_half, int soil_ndp)
{
    PhaseBuf& S = c->ph[1];
    RC(build_cells(c, 1));
    if (c->P.summation_density) {
        // single_step_for_soil's density approximation (declarations_sph.f90:
        // 1558-1562): rho <- sum_density.  first_half and the Drucker-Prager
        // projection of the pack leave rho and x alone under the switch, so
        // the summation can run before them; the pack then records the new rho.
        RC(launch(c, "sum_density", k_sum_density, nblk(S.n), TPB, S.d, c->gl[1], c->K, c->K.h1_soil));
        std::swap(S.d.rho, S.d.rho_next);
    }
    RC(launch(c, "pack", k_pack_soil, nblk(S.n), TPB, S.d, c->K, S.rec, soil_half, soil_ndp, (float*)nullptr,
              c->gl[1], c->cs, NRef{}));
    switch (c->P.skf) {
    case 1: return stages_soil_only<1>(c);
    case 2: return stages_soil_only<2>(c);
    case 3: return stages_soil_only<3>(c);
    default: return stages_soil_only<4>(c);
    }
}

bool soil_only(const sphg_ctx* c) { return c->P.single_phase && !c->ph[0].bound && c->ph[1].bound; }

// Mixed-precision stages: FP32 records and pair sums, FP64 per-particle tails.
template <int SKF>
int stages_mixed(sphg_ctx* c, const FinalArgs* fin)
{
    PhaseBuf& W = c->ph[0];
    PhaseBuf& S = c->ph[1];
    float* qW = reinterpret_cast<float*>(W.rec);
    float* qS = reinterpret_cast<float*>(S.rec);
    // with neighbour lists: the list grid (records hold positions relative
    // to their list cells) and the lists
    const bool nl = nl_mode(c);
    const NList nW = nl ? NList{W.nl_same, W.nl_cross, W.nl_cell} : NList{};
    const NList nS = nl ? NList{S.nl_same, S.nl_cross, S.nl_cell} : NList{};
    const Grid& gW = nl ? c->gvl[0] : c->gl[0];
    const Grid& gS = nl ? c->gvl[1] : c->gl[1];
    MixK X = c->X;
    if (nl) X.cs = (float)c->csv;
    SideFork fork(c);   // soil stage beside the water stages
    RC(fork.rc);
    {
        // the cross terms in a pass of their own before the same-phase pass
        SideScope side(c);
        RC(launch(c, "soil_x", k_soil_x_mx<SKF>, nblk(S.n, STAGE_BLOCK), STAGE_BLOCK, S.d, W.d, (const float*)qS,
                  (const float*)qW, nS, gS, c->K, c->F, X, 1, mask_b(c, 1)));
        RC(launch(c, "soil_ab", k_soil_ab_mx<SKF>, nblk(S.n, STAGE_BLOCK), STAGE_BLOCK, S.d, W.d, (const float*)qS,
                  (const float*)qW, nS, gS, c->K, c->F, X, 1, mask_b(c, 1)));
    }
    // (stage A before the cross pass: beside the soil side's cross pass and
    // then its stage, each pair of concurrent kernels reads one phase's records)
    RC(launch(c, "water_a", k_water_a_mx<SKF>, nblk(W.n, STAGE_BLOCK), STAGE_BLOCK, W.d, S.d, (const float*)qW, qW,
              (const float*)qS, nW, W.masks, gW, c->K, c->F, X, 1));
    RC(launch(c, "water_x", k_water_x_mx<SKF>, nblk(W.n, STAGE_BLOCK), STAGE_BLOCK, W.d, S.d, (const float*)qW, qW,
              (const float*)qS, nW, W.masks, gW, c->K, c->F, X, 1));
    swap_c(W);
    swap_c(S);
    RC(launch(c, "water_b", k_water_b_mx<SKF>, nblk(W.n, STAGE_BLOCK), STAGE_BLOCK, W.d, (const float*)qW, W.masks,
              gW, c->K, c->F, X, mask_b(c, 0)));
    RC(fork.join());
    if (fin) {
        // (the soil stage reads the water records and the water stages the
        // soil records: both finals after the join)
        if (fin->ahead) {
            RC(launch(c, "final_water", k_step_water, nblk(W.n), TPB, W.d, c->K, W.rec, fin->first, fin->vf,
                      fin->renorm, qW, gW, c->csv, nref(c, 0, false)));
            RC(launch(c, "final_soil", k_step_soil, nblk(S.n), TPB, S.d, c->K, S.rec, fin->first, qS, gS, c->csv,
                      nref(c, 1, false)));
        } else {
            RC(launch(c, "final_water", k_final_water, nblk(W.n), TPB, W.d, c->K, fin->first, fin->vf, fin->renorm));
            RC(launch(c, "final_soil", k_final_soil, nblk(S.n), TPB, S.d, c->K, fin->first));
        }
    }
    return SPHG_OK;
}

// Neighbour lists of the fused step at the rebuild: records written in the
// order of the list-grid cell list; same-phase and cross candidates within
// R + skin per particle.
int build_nlists(sphg_ctx* c)
{
    PhaseBuf& W = c->ph[0];
    PhaseBuf& S = c->ph[1];
    const DPhase& w = W.d;
    const DPhase& sd = S.d;
    // the displacement trigger restarts (a past-the-skin bit stays for the check);
    // the rebuild has just permuted the state into cell order (no order lookups)
    RC(launch(c, "nlist", k_keep_bits, 1, TPB, c->d_nlflag, 2));
    RC(launch(c, "nlist", k_nlist, nblk(W.n, 256), 256, (const double*)w.x, (const double*)w.y,
              (const int*)nullptr, W.n, (const double*)w.x, (const double*)w.y, (const int*)nullptr,
              (const int*)w.cell_start, c->gv, c->nl_thr[0], 1, W.nl_same, W.nl_cell));
    RC(launch(c, "nlist", k_nlist, nblk(S.n, 256), 256, (const double*)sd.x, (const double*)sd.y,
              (const int*)nullptr, S.n, (const double*)sd.x, (const double*)sd.y, (const int*)nullptr,
              (const int*)sd.cell_start, c->gv, c->nl_thr[1], 1, S.nl_same, S.nl_cell));
    RC(launch(c, "nlist", k_nlist, nblk(W.n, 256), 256, (const double*)w.x, (const double*)w.y,
              (const int*)nullptr, W.n, (const double*)sd.x, (const double*)sd.y, (const int*)nullptr,
              (const int*)sd.cell_start, c->gv, c->nl_thr[2], 0, W.nl_cross, (int*)nullptr));
    RC(launch(c, "nlist", k_nlist, nblk(S.n, 256), 256, (const double*)sd.x, (const double*)sd.y,
              (const int*)nullptr, S.n, (const double*)w.x, (const double*)w.y, (const int*)nullptr,
              (const int*)w.cell_start, c->gv, c->nl_thr[2], 0, S.nl_cross, (int*)nullptr));
    c->nl_valid = true;
    c->nl_due = false;
    c->nl_rebuilt_at = c->nl_seq + 1;   // the step being issued
    ++c->nl_rebuilds;
    return SPHG_OK;
}

// The stage kernels of the fused two-phase step (records packed), and with
// fin the step's final kernels (FP64 only).
int stages_of(sphg_ctx* c, const FinalArgs* fin = nullptr)
{
    if (c->P.precision == SPHG_MIXED) {
        switch (c->P.skf) {
        case 1: return stages_mixed<1>(c, fin);
        case 2: return stages_mixed<2>(c, fin);
        case 3: return stages_mixed<3>(c, fin);
        default: return stages_mixed<4>(c, fin);
        }
    }
    switch (c->P.skf) {
    case 1: return stages_fused<1>(c, fin);
    case 2: return stages_fused<2>(c, fin);
    case 3: return stages_fused<3>(c, fin);
    default: return stages_fused<4>(c, fin);
    }
}

// Everything of the fused two-phase right-hand side before its stages: cell
// lists, reorder, packs and (list mode) the neighbour lists.
int pack_two_phase(sphg_ctx* c, int water_half, int soil_half, int soil_ndp)
{
    PhaseBuf& W = c->ph[0];
    PhaseBuf& S = c->ph[1];
    const bool mixed = c->P.precision == SPHG_MIXED;
    float* qW = mixed ? reinterpret_cast<float*>(W.rec) : nullptr;
    float* qS = mixed ? reinterpret_cast<float*>(S.rec) : nullptr;
    // FP64 with lists: a rebuild sorts both phases on the list grid (cell
    // list, then the state permuted into its order), the packs record the
    // positions and the lists follow the packs; between rebuilds the cell
    // lists and the record order stay and the packs track the displacement.
    const bool nl = nl_mode(c);
    // (cell lists invalidated since -- a restore, a correction, an upload --
    // mean a rebuild as well: the lists index their order)
    const bool rebuild = nl && (!c->nl_valid || !W.cells_valid || !S.cells_valid || !W.cells_nl || !S.cells_nl);
    if (c->packed) {
        // the final kernels of the step before packed this one
        c->packed = false;
        if (!nl || rebuild || !water_half || !soil_half || soil_ndp != 2) return SPHG_ERR_INVALID;
        return SPHG_OK;
    }
    const NRef nrW = nl ? nref(c, 0, rebuild) : NRef{}, nrS = nl ? nref(c, 1, rebuild) : NRef{};
    if (rebuild) {
        W.cells_valid = false;
        S.cells_valid = false;
        c->since_reorder = 0;
    }
    auto soil_chain = [&]() -> int {
        RC(build_cells(c, 1, nl));
        if (rebuild) RC(reorder(c, 1));
        return launch(c, "pack", k_pack_soil, nblk(S.n), TPB, S.d, c->K, S.rec, soil_half, soil_ndp, qS,
                      nl ? c->gvl[1] : c->gl[1], nl ? c->csv : c->cs, nrS);
    };
    // The two phases' cell lists and packs are independent: the soil chain
    // runs on the side stream beside the water chain (the stages after them
    // read both phases' records).
    {
        SideFork fork(c);
        RC(fork.rc);
        if (fork.open) {
            SideScope sc(c);
            RC(soil_chain());
        }
        RC(build_cells(c, 0, nl));
        if (rebuild) RC(reorder(c, 0));
        RC(launch(c, "pack", k_pack_water, nblk(W.n), TPB, W.d, c->K, W.rec, water_half, qW, nl ? c->gvl[0] : c->gl[0],
                  nl ? c->csv : c->cs, nrW));
        if (fork.open) RC(fork.join());
        else RC(soil_chain());
    }
    if (rebuild) RC(build_nlists(c));
    return SPHG_OK;
}

int rhs_two_phase_fused(sphg_ctx* c, int water_half, int soil_half, int soil_ndp)
{
    RC(pack_two_phase(c, water_half, soil_half, soil_ndp));
    return stages_of(c);
}

// single_step_for_water (declarations_sph.f90:1439-1526).
int rhs_water(sphg_ctx* c)
{
    PhaseBuf& W = c->ph[0];
    RC(build_cells(c, 0));
    RC(launch(c, "sum_density", k_sum_density, nblk(W.n), TPB, W.d, c->gl[0], c->K, c->K.h1_water));
    std::swap(W.d.rho, W.d.rho_next);
    int ma = M_DRHO_DELTA | M_XSPH;
    if (!c->P.artificial_density) ma &= ~M_DRHO_DELTA;
    if (!c->P.average_velocity) ma &= ~M_XSPH;
    RC(launch(c, "water1_a", k_water1_a, nblk(W.n), TPB, W.d, c->gl[0], c->K, ma));
    int mb = c->P.visc_artificial ? M_ART : 0;
    RC(launch(c, "water1_b", k_water1_b, nblk(W.n), TPB, W.d, c->gl[0], c->K, mb));
    return SPHG_OK;
}

int check_fused_switches(const sphg_ctx* c)
{
    const sphg_params& P = c->P;
    if (P.stress_integration != 1 && P.stress_integration != 2) return SPHG_ERR_UNSUPPORTED;
    // summation_density: the two-phase single_step keeps con_density (its
    // sum_density call is commented out, declarations_sph.f90:1190-1199) and
    // the leapfrog leaves rho alone; single_step_for_soil calls sum_density
    // stress_integration = 2 and plasticity = 2 (return mapping) in the FP64 fused steps
    if ((P.stress_integration == 2 || P.plasticity == 2) && (P.precision != SPHG_FP64 || c->legacy))
        return SPHG_ERR_UNSUPPORTED;
    // the Verlet scheme is the reference's single-phase water integrator
    if (P.integrate_scheme == 2 && !P.single_phase) return SPHG_ERR_UNSUPPORTED;
    if (soil_only(c)) {
        // time_integration_for_soil: leapfrog, FP64; either criterion, flow rules 0-3
        if (P.integrate_scheme == 2 || P.precision != SPHG_FP64) return SPHG_ERR_UNSUPPORTED;
        if ((P.yield_criterion != 1 && P.yield_criterion != 2) || P.plasticity < 0 || P.plasticity > 3)
            return SPHG_ERR_UNSUPPORTED;
        return SPHG_OK;
    }
    // the mixed-precision form exists for the fused two-phase step with the
    // direct_find predicate; its records hold 15-bit cell coordinates
    if (P.precision == SPHG_MIXED &&
        (P.single_phase || P.nnps != 1 || c->legacy || c->g.ncx > 32767 || c->g.ncy > 32767))
        return SPHG_ERR_UNSUPPORTED;
    if (!P.single_phase) {
        if (!c->ph[1].bound) return SPHG_ERR_INVALID;
        // Mohr-Coulomb in the FP64 fused step only
        if (P.yield_criterion != 2 && !(P.yield_criterion == 1 && P.precision == SPHG_FP64 && !c->legacy))
            return SPHG_ERR_UNSUPPORTED;
        // plastic_flow_rule (plasticity 1) in the FP64 fused step only
        if (P.plasticity != 3 && P.plasticity != 0 &&
            !((P.plasticity == 1 || P.plasticity == 2) && P.precision == SPHG_FP64 && !c->legacy))
            return SPHG_ERR_UNSUPPORTED;
    }
    if (!c->ph[0].bound) return SPHG_ERR_INVALID;
    return SPHG_OK;
}

// The water volume-fraction renormalisation of single_step
// (declarations_sph.f90:1414-1420) acts on every 40th step; it changes the
// state only when the leapfrog leaves rho alone (summation_density), since
// second_half otherwise recomputes rho and vof from their saved levels.
bool renorm_active(const sphg_params& P)
{
    return !P.single_phase && P.volume_fraction && P.volume_fraction_renorm && P.summation_density;
}

bool renorm_step(const sphg_ctx* c, int step) { return renorm_active(c->P) && step % 40 == 0; }

// Lagged displacement check of the list mode, in each step issued outside
// a stream capture once its packs are queued: the flag word is copied to
// pinned memory behind them, and the copy of the step before is read (it
// has landed by the time the step's packs are queued, so the queue keeps
// the step's stages and the rest of the step before).  A particle past the
// rebuild trigger schedules a rebuild for the step after this one; one past
// half the skin fails the run (its step may have missed pairs).
int nl_poll(sphg_ctx* c)
{
    const int s = ++c->nl_seq, slot = s & 1, o = slot ^ 1;
    HIPCHK(c, hipMemcpyAsync(&c->h_nlflag[slot], c->d_nlflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipEventRecord(c->ev_nl[slot], c->stream));
    c->h_nlstep[slot] = s;
    if (c->h_nlstep[o] != s - 1) return SPHG_OK;
    HIPCHK(c, hipEventSynchronize(c->ev_nl[o]));
    const int f = c->h_nlflag[o];
    c->h_nlstep[o] = -1;
    if (f & 2) return check_flags(c);
    if ((f & 1) && s - 1 >= c->nl_rebuilt_at) {
        c->nl_due = true;
        // rebuilds this close together cost more than the row scans: lists
        // off for a while (never in a batch, whose captured period has them)
        if (s + 1 - c->nl_rebuilt_at < NL_MIN_RUN && !c->stable_ptrs) {
            c->nl_off_from = c->itimestep + 1;
            c->nl_off_until = c->itimestep + 1 + NL_OFF;
            ++c->nl_off_spells;
        }
    }
    return SPHG_OK;
}

// ahead_ok: another step of the same run follows at once (nothing changes
// the state in between), so its pack may join this step's final update.
int one_step(sphg_ctx* c, bool ahead_ok = false)
{
    c->last_k = 1;
    const int step = ++c->itimestep;
    const bool nl = nl_mode(c);
    PhaseBuf& W = c->ph[0];
    PhaseBuf& S = c->ph[1];
    if (soil_only(c)) {
        // time_integration_for_soil (declarations_sph.f90:978-1127): first_half
        // and a Drucker-Prager projection, single_step_for_soil (its failure
        // criterion is the second projection of the pack), first_step /
        // second_half and the closing projection
        if (++c->since_reorder >= c->reorder_every) {
            RC(build_cells(c, 1));
            RC(reorder(c, 1));
            c->since_reorder = 0;
        }
        if (step != 1) RC(rhs_soil_only(c, 1, 2));
        else RC(rhs_soil_only(c, 0, 1));
        RC(launch(c, "final_soil", k_final_soil, nblk(S.n), TPB, S.d, c->K, step == 1 ? 1 : 0));
        S.cells_valid = false;
        return SPHG_OK;
    }
    if (c->P.single_phase) {
        if (step != 1) RC(launch(c, "first_half", k_first_half_water, nblk(W.n), TPB, W.d, c->K));
        RC(rhs_water(c));
        RC(launch(c, "final_water", k_final_water, nblk(W.n), TPB, W.d, c->K, step == 1 ? 1 : 0, 0, 0));
        W.cells_valid = false;
        return SPHG_OK;
    }
    if (c->legacy) {
        if (step != 1) {
            RC(launch(c, "first_half", k_first_half_water, nblk(W.n), TPB, W.d, c->K));
            RC(launch(c, "first_half", k_first_half_soil, nblk(S.n), TPB, S.d, c->K, 1, 2));
        } else {
            RC(launch(c, "first_half", k_first_half_soil, nblk(S.n), TPB, S.d, c->K, 0, 1));
        }
        RC(rhs_two_phase_legacy(c));
    } else if (nl) {
        // the lists are rebuilt (with the reorder) every nl_every steps and
        // when the displacement check asked for it
        if (++c->since_reorder >= c->nl_every || c->nl_due) c->nl_valid = false;
        if (step != 1) RC(pack_two_phase(c, 1, 1, 2));
        else RC(pack_two_phase(c, 0, 0, 1));
        // the displacement check of the step before (it may ask for a
        // rebuild at the next step, or report a particle past half the skin)
        const int prc = c->capturing ? SPHG_OK : nl_poll(c);
        // when the next step keeps the lists (no rebuild, no spell without
        // them), its pack joins this step's final update
        const bool ahead = ahead_ok && prc == SPHG_OK && c->nl_valid && !c->nl_due &&
                           c->since_reorder + 1 < c->nl_every && nl_at(c, c->itimestep + 1);
        const FinalArgs fa{step == 1 ? 1 : 0, c->P.volume_fraction ? 1 : 0, renorm_step(c, step) ? 1 : 0, ahead};
        RC(stages_of(c, &fa));
        c->packed = ahead;
        return prc;
    } else {
        if (++c->since_reorder >= c->reorder_every) {
            RC(build_cells(c, 0));
            RC(build_cells(c, 1));
            RC(reorder(c, 0));
            RC(reorder(c, 1));
            c->since_reorder = 0;
        }
        if (step != 1) RC(rhs_two_phase_fused(c, 1, 1, 2));
        else RC(rhs_two_phase_fused(c, 0, 0, 1));
    }
    const int vf = c->P.volume_fraction ? 1 : 0;
    RC(launch(c, "final_water", k_final_water, nblk(W.n), TPB, W.d, c->K, step == 1 ? 1 : 0, vf,
              renorm_step(c, step) ? 1 : 0));
    RC(launch(c, "final_soil", k_final_soil, nblk(S.n), TPB, S.d, c->K, step == 1 ? 1 : 0));
    // the positions moved: the next step builds its cell lists again
    W.cells_valid = false;
    S.cells_valid = false;
    return SPHG_OK;
}

// time_integration_for_water_by_verlet (declarations_sph.f90:889-974): the
// routine's lastvx / lastrho start as copies of vx and rho at every call.
int verlet_steps(sphg_ctx* c, int nsteps)
{
    PhaseBuf& W = c->ph[0];
    if (!W.lvx) {
        RC(dalloc(c, W, &W.lvx, (size_t)W.n));
        RC(dalloc(c, W, &W.lvy, (size_t)W.n));
        RC(dalloc(c, W, &W.lrho, (size_t)W.n));
    }
    RC(launch(c, "verlet", k_copy<double>, nblk(W.n), TPB, (const double*)W.d.vx, W.lvx, W.n));
    RC(launch(c, "verlet", k_copy<double>, nblk(W.n), TPB, (const double*)W.d.vy, W.lvy, W.n));
    RC(launch(c, "verlet", k_copy<double>, nblk(W.n), TPB, (const double*)W.d.rho, W.lrho, W.n));
    const double dt = c->P.dt, h2 = std::pow(dt, 2.0) / 2.0;   // dt**2./2.
    for (int it = 0; it < nsteps; ++it) {
        const int step = ++c->itimestep;
        RC(rhs_water(c));
        RC(launch(c, "verlet", k_verlet_water, nblk(W.n), TPB, W.d, W.lvx, W.lvy, W.lrho, dt, h2, step));
        W.cells_valid = false;
    }
    return SPHG_OK;
}

// ---- materialised pair lists -------------------------------------------

int materialise(sphg_ctx* c, int ka, int kb, int same, int link, sphg_phase_view* out_view,
                int* out_niac)
{
    PhaseBuf& A = c->ph[ka];
    PhaseBuf& B = c->ph[kb];
    RC(build_cells(c, kb));
    const double h1 = (ka == 0) ? c->K.h1_water : c->K.h1_soil;
    RC(launch(c, "pair_count", k_pair_count, nblk(A.n), TPB, A.d, B.d, c->gl[ka], c->K, same, link, h1,
              A.fwd, A.d.countiac));
    RC(exclusive_scan(c, A.fwd, A.off, A.n));
    int total = 0;
    HIPCHK(c, hipMemcpyAsync(&total, A.off + A.n, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    RC(check_flags(c));
    *out_niac = total;
    if (total > out_view->max_interaction) return SPHG_ERR_TOO_MANY_INTERACTIONS;
    if (!out_view->pair_i || !out_view->pair_j || !out_view->w || !out_view->dwdx) return SPHG_ERR_INVALID;
    RC(ensure_pairs(c, A, total > 0 ? total : 1));
    RC(launch(c, "pair_write", k_pair_write, nblk(A.n), TPB, A.d, B.d, c->gl[ka], c->K, same, link, h1,
              (const int*)A.off, A.pi, A.pj, A.w, A.dwdx));
    if (total > 0) {
        HIPCHK(c, hipMemcpyAsync(out_view->pair_i, A.pi, sizeof(int) * (size_t)total, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(out_view->pair_j, A.pj, sizeof(int) * (size_t)total, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(out_view->w, A.w, sizeof(double) * (size_t)total, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(out_view->dwdx, A.dwdx, sizeof(double) * 2 * (size_t)total, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SPHG_OK;
}

int bound_index(sphg_ctx* c, int phase)
{
    const int k = phase_index(phase);
    if (k < 0 || !c->ph[k].bound) return -1;
    return k;
}

}  // namespace

// ===========================================================================
// public entry points
// ===========================================================================
extern "C" {

int sphg_create(const sphg_params* params, sphg_ctx** out)
{
    if (!params || !out) return SPHG_ERR_INVALID;
    RC(check_params(*params));
    sphg_ctx* c = new sphg_ctx();
    c->P = *params;
    {
        const char* e = std::getenv("SPHG_LEGACY");
        c->legacy = e && e[0] == '1';
        const char* r = std::getenv("SPHG_REORDER_EVERY");
        if (r && std::atoi(r) > 0) c->reorder_every = std::atoi(r);
        c->since_reorder = c->reorder_every;   // first fused step sorts the state
    }
    set_kparams(c);
    {
        const char* e = std::getenv("SPHG_SIDE_STREAM");
        if (e && e[0] == '0') c->use_side = false;
        const char* l = std::getenv("SPHG_NLIST");
        if (l && l[0] == '0') c->use_nl = false;
        const char* k = std::getenv("SPHG_NL_SKIN");
        if (k && std::atof(k) > 0.0) c->nl_skin = std::atof(k);
        const char* v = std::getenv("SPHG_NL_EVERY");
        if (v && std::atoi(v) > 0) c->nl_every = std::atoi(v);
    }
    if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_wx, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_sx, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_nl[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_nl[1], hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc((void**)&c->h_nlflag, 2 * sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&c->d_err, sizeof(int) * 4) != hipSuccess) {
        delete c;
        return SPHG_ERR_HIP;
    }
    c->d_nlflag = c->d_err + 1;
    c->h_nlflag[0] = c->h_nlflag[1] = 0;
    c->h_nlstep[0] = c->h_nlstep[1] = -1;
    const int zeros[4] = {0, 0, 0, 0};
    if (hipMemcpy(c->d_err, zeros, sizeof(zeros), hipMemcpyHostToDevice) != hipSuccess) {
        sphg_destroy(c);
        return SPHG_ERR_HIP;
    }
    *out = c;
    return SPHG_OK;
}

void sphg_destroy(sphg_ctx* c)
{
    if (!c) return;
    hipStreamSynchronize(c->stream);
    drain_events(c);
    for (auto& pb : c->ph) free_phase(pb);
    for (hipEvent_t e : c->event_pool) hipEventDestroy(e);
    if (c->d_err) hipFree(c->d_err);
    if (c->d_scan) hipFree(c->d_scan);
    if (c->side) {
        hipStreamSynchronize(c->side);
        hipStreamDestroy(c->side);
    }
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    if (c->ev_join) hipEventDestroy(c->ev_join);
    if (c->ev_wx) hipEventDestroy(c->ev_wx);
    if (c->ev_sx) hipEventDestroy(c->ev_sx);
    for (hipEvent_t e : c->ev_nl)
        if (e) hipEventDestroy(e);
    if (c->h_nlflag) hipHostFree(c->h_nlflag);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int sphg_set_params(sphg_ctx* c, const sphg_params* params)
{
    if (!c || !params) return SPHG_ERR_INVALID;
    RC(check_params(*params));
    const bool regrid = params->x_mingeom != c->P.x_mingeom || params->x_maxgeom != c->P.x_maxgeom ||
                        params->y_mingeom != c->P.y_mingeom || params->y_maxgeom != c->P.y_maxgeom ||
                        params->skf != c->P.skf || params->nnps != c->P.nnps ||
                        params->maxngx != c->P.maxngx || params->maxngy != c->P.maxngy;
    c->P = *params;
    set_kparams(c);
    if (regrid && (c->ph[0].bound || c->ph[1].bound)) {
        c->ncell = 0;
        RC(setup_grid(c));
        for (auto& pb : c->ph) pb.cells_valid = false;
    }
    return SPHG_OK;
}

int sphg_bind_phase(sphg_ctx* c, int phase, sphg_phase_view* v)
{
    if (!c || !v) return SPHG_ERR_INVALID;
    const int k = phase_index(phase);
    if (k < 0 || v->material != phase) return SPHG_ERR_INVALID;
    if (v->ntotal < 0 || v->nvirt < 0 || v->ntotal + v->nvirt > v->maxn) return SPHG_ERR_INVALID;
    if (!v->x || !v->vx || !v->mass || !v->hsml || !v->itype || !v->rho) return SPHG_ERR_INVALID;
    RC(alloc_phase(c, k, v));
    PhaseBuf& b = c->ph[k];
    b.user = v;
    RC(up_int(c, b, v->itype, b.d.itype));
    RC(up_scalar(c, b, v->mass, b.d.mass));
    RC(up_scalar(c, b, v->hsml, b.d.hsml));
    RC(upload(c, k, SPHG_F_ALL & ~SPHG_F_COUNTIAC));
    if (!v->c) RC(fill_dbl(c, b.d.c, b.n, phase == SPHG_WATER ? c->P.water_c : c->P.soil_c));
    if (!v->vof) RC(fill_dbl(c, b.d.vof, b.n, 1.0));
    HIPCHK(c, hipMemcpyAsync(b.d.c_next, b.d.c, sizeof(double) * (size_t)b.n, hipMemcpyDeviceToDevice, c->stream));
    c->ncell = 0;
    RC(setup_grid(c));
    for (auto& pb : c->ph) pb.cells_valid = false;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SPHG_OK;
}

int sphg_sync(sphg_ctx* c, int phase, uint64_t mask, int to_host)
{
    if (!c) return SPHG_ERR_INVALID;
    const int k = bound_index(c, phase);
    if (k < 0) return SPHG_ERR_INVALID;
    if (to_host) {
        RC(download(c, k, mask));
        RC(check_flags(c));
    } else {
        RC(upload(c, k, mask));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return SPHG_OK;
}

int sphg_set_itimestep(sphg_ctx* c, int it)
{
    if (!c || it < 0) return SPHG_ERR_INVALID;
    c->itimestep = it;
    return SPHG_OK;
}

int sphg_get_itimestep(const sphg_ctx* c) { return c ? c->itimestep : -1; }

int sphg_direct_find(sphg_ctx* c, int phase)
{
    if (!c) return SPHG_ERR_INVALID;
    const int k = bound_index(c, phase);
    if (k < 0) return SPHG_ERR_INVALID;
    RC(upload(c, k, SPHG_F_X));
    int niac = 0;
    const int rc = materialise(c, k, k, 1, 0, &c->ph[k].view, &niac);
    c->ph[k].view.niac = niac;
    c->ph[k].user->niac = niac;
    if (rc == SPHG_OK) RC(download(c, k, SPHG_F_COUNTIAC));
    return rc;
}

int sphg_link_list(sphg_ctx* c, int phase)
{
    if (!c) return SPHG_ERR_INVALID;
    const int k = bound_index(c, phase);
    if (k < 0) return SPHG_ERR_INVALID;
    RC(upload(c, k, SPHG_F_X));
    int niac = 0;
    const int rc = materialise(c, k, k, 1, 1, &c->ph[k].view, &niac);
    c->ph[k].view.niac = niac;
    c->ph[k].user->niac = niac;
    if (rc == SPHG_OK) RC(download(c, k, SPHG_F_COUNTIAC));
    return rc;
}

int sphg_direct_find_2(sphg_ctx* c)
{
    if (!c || !c->ph[0].bound || !c->ph[1].bound) return SPHG_ERR_INVALID;
    RC(upload(c, 0, SPHG_F_X));
    RC(upload(c, 1, SPHG_F_X));
    int niac = 0;
    const int rc = materialise(c, 0, 1, 0, 0, &c->ph[0].view, &niac);
    c->ph[0].view.niac = niac;
    c->ph[0].user->niac = niac;
    if (rc != SPHG_OK) return rc;
    // countiac of the soil phase: number of water partners
    PhaseBuf& S = c->ph[1];
    RC(build_cells(c, 0));
    RC(launch(c, "pair_count", k_pair_count, nblk(S.n), TPB, S.d, c->ph[0].d, c->gl[1], c->K, 0, 0, 0.0,
              S.fwd, S.d.countiac));
    RC(download(c, 0, SPHG_F_COUNTIAC));
    RC(download(c, 1, SPHG_F_COUNTIAC));
    return SPHG_OK;
}

int sphg_con_density(sphg_ctx* c, int phase)
{
    if (!c) return SPHG_ERR_INVALID;
    const int k = bound_index(c, phase);
    if (k < 0) return SPHG_ERR_INVALID;
    RC(upload(c, k, SPHG_F_X | SPHG_F_VX | SPHG_F_RHO));
    RC(build_cells(c, k));
    PhaseBuf& b = c->ph[k];
    if (k == 0)
        RC(launch(c, "water_a", k_water_a, nblk(b.n), TPB, b.d, c->ph[1].d, c->gl[0], c->K, (int)M_CON));
    else
        RC(launch(c, "soil_a", k_soil_a, nblk(b.n), TPB, b.d, c->ph[0].d, c->gl[1], c->K, (int)M_CON));
    RC(download(c, k, SPHG_F_DRHO));
    return check_flags(c);
}

int sphg_delta_sph_density(sphg_ctx* c, int phase)
{
    if (!c) return SPHG_ERR_INVALID;
    const int k = bound_index(c, phase);
    if (k < 0) return SPHG_ERR_INVALID;
    RC(upload(c, k, SPHG_F_X | SPHG_F_VX | SPHG_F_RHO | SPHG_F_C | SPHG_F_DRHO));
    RC(build_cells(c, k));
    PhaseBuf& b = c->ph[k];
    const int m = M_DRHO_DELTA | M_DRHO_ACCUM;
    if (k == 0)
        RC(launch(c, "water_a", k_water_a, nblk(b.n), TPB, b.d, c->ph[1].d, c->gl[0], c->K, m));
    else
        RC(launch(c, "soil_a", k_soil_a, nblk(b.n), TPB, b.d, c->ph[0].d, c->gl[1], c->K, m));
    RC(download(c, k, SPHG_F_DRHO));
    return check_flags(c);
}

int sphg_sum_density(sphg_ctx* c, int phase)
{
    if (!c) return SPHG_ERR_INVALID;
    const int k = bound_index(c, phase);
    if (k < 0) return SPHG_ERR_INVALID;
    RC(upload(c, k, SPHG_F_X | SPHG_F_RHO));
    RC(build_cells(c, k));
    PhaseBuf& b = c->ph[k];
    RC(launch(c, "sum_density", k_sum_density, nblk(b.n), TPB, b.d, c->gl[k], c->K,
              k == 0 ? c->K.h1_water : c->K.h1_soil));
    std::swap(b.d.rho, b.d.rho_next);
    RC(download(c, k, SPHG_F_RHO));
    return check_flags(c);
}

int sphg_internal_force(sphg_ctx* c, int phase)
{
    if (!c) return SPHG_ERR_INVALID;
    const int k = bound_index(c, phase);
    if (k < 0) return SPHG_ERR_INVALID;
    PhaseBuf& b = c->ph[k];
    if (k == 0) {
        RC(upload(c, k, SPHG_F_X | SPHG_F_VX | SPHG_F_RHO | SPHG_F_VOF | SPHG_F_C));
        RC(build_cells(c, k));
        RC(launch(c, "water_a", k_water_a, nblk(b.n), TPB, b.d, c->ph[1].d, c->gl[0], c->K, (int)(M_STRAIN | M_EOS)));
        swap_c(b);
        RC(launch(c, "water_b", k_water_b, nblk(b.n), TPB, b.d, c->gl[0], c->K, (int)M_MOMENTUM));
        RC(download(c, k, SPHG_F_DVX | SPHG_F_P | SPHG_F_C | SPHG_F_STRAIN | SPHG_F_STRESS));
    } else {
        if (c->P.yield_criterion != 2 || (c->P.plasticity != 3 && c->P.plasticity != 0)) return SPHG_ERR_UNSUPPORTED;
        RC(upload(c, k, SPHG_F_X | SPHG_F_VX | SPHG_F_RHO | SPHG_F_VOF | SPHG_F_C | SPHG_F_P | SPHG_F_STRESS |
                            SPHG_F_DSTRESS | SPHG_F_EPSP));
        RC(build_cells(c, k));
        RC(launch(c, "first_half", k_first_half_soil, nblk(b.n), TPB, b.d, c->K, 0, 1));
        RC(launch(c, "soil_a", k_soil_a, nblk(b.n), TPB, b.d, c->ph[0].d, c->gl[1], c->K,
                  (int)(M_STRAIN | M_EOS | M_PLASTIC | M_DS_ACCUM)));
        swap_c(b);
        RC(launch(c, "soil_b", k_soil_b, nblk(b.n), TPB, b.d, c->gl[1], c->K, (int)M_MOMENTUM));
        RC(download(c, k, SPHG_F_DVX | SPHG_F_P | SPHG_F_C | SPHG_F_STRAIN | SPHG_F_STRESS | SPHG_F_DSTRESS |
                              SPHG_F_EPSP | SPHG_F_FAIL));
    }
    return check_flags(c);
}

int sphg_art_visc(sphg_ctx* c, int phase)
{
    if (!c) return SPHG_ERR_INVALID;
    const int k = bound_index(c, phase);
    if (k < 0) return SPHG_ERR_INVALID;
    RC(upload(c, k, SPHG_F_X | SPHG_F_VX | SPHG_F_RHO | SPHG_F_C | SPHG_F_DVX));
    RC(build_cells(c, k));
    PhaseBuf& b = c->ph[k];
    if (c->P.single_phase && k == 0)
        RC(launch(c, "water1_b", k_water1_b, nblk(b.n), TPB, b.d, c->gl[0], c->K, (int)M_ART));
    else if (k == 0)
        RC(launch(c, "water_b", k_water_b, nblk(b.n), TPB, b.d, c->gl[0], c->K, (int)M_ART));
    else
        RC(launch(c, "soil_b", k_soil_b, nblk(b.n), TPB, b.d, c->gl[1], c->K, (int)M_ART));
    RC(download(c, k, SPHG_F_DVX));
    return check_flags(c);
}

int sphg_single_step(sphg_ctx* c, int sync)
{
    if (!c) return SPHG_ERR_INVALID;
    RC(check_fused_switches(c));
    if (soil_only(c)) {
        RC(rhs_soil_only(c, 0, 1));
    } else if (c->P.single_phase) {
        RC(rhs_water(c));
    } else {
        PhaseBuf& S = c->ph[1];
        if (c->legacy) {
            RC(launch(c, "first_half", k_first_half_soil, nblk(S.n), TPB, S.d, c->K, 0, 1));
            RC(rhs_two_phase_legacy(c));
        } else {
            RC(rhs_two_phase_fused(c, 0, 0, 1));
        }
        const sphg_params& P = c->P;
        if (P.volume_fraction && P.volume_fraction_renorm && c->itimestep % 40 == 0)
            RC(launch(c, "renorm", k_renorm_water, nblk(c->ph[0].n), TPB, c->ph[0].d));
    }
    if (sync) {
        const uint64_t m = SPHG_F_ALL & ~SPHG_F_X & ~SPHG_F_VX & ~SPHG_F_MIN;
        for (int k = 0; k < 2; ++k)
            if (c->ph[k].bound) RC(download(c, k, m));
    }
    return check_flags(c);
}

int sphg_time_integration(sphg_ctx* c, int nsteps, int sync)
{
    if (!c || nsteps < 0) return SPHG_ERR_INVALID;
    RC(check_fused_switches(c));
    if (c->P.single_phase && c->P.integrate_scheme == 2) RC(verlet_steps(c, nsteps));
    else
        for (int it = 0; it < nsteps; ++it) RC(one_step(c, it + 1 < nsteps));
    if (sync) {
        for (int k = 0; k < 2; ++k)
            if (c->ph[k].bound) RC(download(c, k, SPHG_F_ALL));
        return check_flags(c);
    }
    HIPCHK(c, hipGetLastError());
    return SPHG_OK;
}

const char* sphg_strerror(int code)
{
    switch (code) {
    case SPHG_OK: return "ok";
    case SPHG_ERR_TOO_MANY_INTERACTIONS: return ">>> ERROR <<< : Too many interactions";
    case SPHG_ERR_OUT_OF_RANGE: return ">>> ERROR <<< : Particle out of range";
    case SPHG_ERR_BAD_KERNEL: return "get_scale_k: No such kernel function!";
    case SPHG_ERR_INVALID: return "invalid argument or unbound phase";
    case SPHG_ERR_HIP: return "HIP runtime error";
    case SPHG_ERR_UNSUPPORTED: return "switch combination not supported by this path";
    case SPHG_ERR_NONFINITE: return "non-finite value detected";
    case SPHG_ERR_OUT_OF_MEMORY: return "out of device memory";
    case SPHG_ERR_MOVED_TOO_FAR: return "a particle moved farther than the neighbour-list skin allows";
    default: return "unknown error";
    }
}

int sphg_last_hip_error(const sphg_ctx* c) { return c ? c->last_hip : noalloc_hip_error.load(); }

void* sphg_stream(sphg_ctx* c) { return c ? (void*)c->stream : nullptr; }

int sphg_profile_enable(sphg_ctx* c, int enable)
{
    if (!c) return SPHG_ERR_INVALID;
    if (!enable) drain_events(c);
    c->prof = enable != 0;
    return SPHG_OK;
}

int sphg_profile_reset(sphg_ctx* c)
{
    if (!c) return SPHG_ERR_INVALID;
    drain_events(c);
    c->stats.clear();
    return SPHG_OK;
}

int sphg_profile_get(sphg_ctx* c, const char* name, int64_t* launches, double* total_ms)
{
    if (!c || !name) return SPHG_ERR_INVALID;
    drain_events(c);
    auto it = c->stats.find(name);
    if (it == c->stats.end()) {
        if (launches) *launches = 0;
        if (total_ms) *total_ms = 0.0;
        return SPHG_ERR_INVALID;
    }
    if (launches) *launches = it->second.launches;
    if (total_ms) *total_ms = it->second.ms;
    return SPHG_OK;
}

int sphg_nlist_stats(const sphg_ctx* c, int64_t* rebuilds, int64_t* off_spells, double* skin)
{
    if (!c || !rebuilds || !off_spells || !skin) return SPHG_ERR_INVALID;
    *rebuilds = c->nl_rebuilds;
    *off_spells = c->nl_off_spells;
    *skin = c->skin;
    return SPHG_OK;
}

int sphg_pair_counts(sphg_ctx* c, int64_t* wp, int64_t* sp, int64_t* xp)
{
    if (!c) return SPHG_ERR_INVALID;
    // Pairs seen by the most recent fused step: half the same-phase neighbour
    // counts, and the cross counts (countiac of water in two-phase runs).
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int64_t out[3] = {0, 0, 0};
    auto sum = [&](const int* d, int n, int64_t& s) -> int {
        std::vector<int> cnt((size_t)n);
        HIPCHK(c, hipMemcpy(cnt.data(), d, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost));
        s = 0;
        for (int v : cnt) s += v;
        return SPHG_OK;
    };
    for (int k = 0; k < 2; ++k) {
        PhaseBuf& b = c->ph[k];
        if (!b.bound) continue;
        int64_t s = 0;
        if (c->P.single_phase) {
            // single-phase steps keep the same-phase counts in countiac
            RC(sum(b.d.countiac, b.n, s));
        } else {
            RC(sum(b.d.nsame, b.n, s));
            if (k == 0) RC(sum(b.d.countiac, b.n, out[2]));
        }
        out[k] = s / 2;
    }
    if (wp) *wp = out[0];
    if (sp) *sp = out[1];
    if (xp) *xp = out[2];
    return SPHG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Learned multi-step surrogate (no counterpart in the reference).
// ---------------------------------------------------------------------------
namespace {

__global__ void k_fill_f32(float* a, int n, float v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// The FP64 fields of a phase that a later step reads; rho is entry 6.
constexpr int N_STATE_FIELDS = 23;
static_assert(N_STATE_FIELDS <= PERM_MAXD, "a snapshot copies its fields in one kernel");
void state_fields(DPhase& d, double** out[N_STATE_FIELDS])
{
    double** list[N_STATE_FIELDS] = {&d.x, &d.y, &d.vx, &d.vy, &d.mass, &d.hsml, &d.rho, &d.vof, &d.p, &d.c,
                                     &d.drho, &d.ax, &d.ay, &d.dvof, &d.sxx, &d.sxy, &d.syy, &d.eps_p,
                                     &d.dsxx, &d.dsxy, &d.dsyy, &d.dp, &d.c_next};
    for (int i = 0; i < N_STATE_FIELDS; ++i) out[i] = list[i];
}

// Persistent per-particle state of a phase (what a step reads from the past).
struct StateCopy {
    int n = 0;
    int *id = nullptr, *itype = nullptr, *fail = nullptr;
    double* f[N_STATE_FIELDS] = {};
};

}  // namespace

struct sphg_gnn {
    sphg_ctx* c = nullptr;
    int K = 4;
    int maxs = 0;
    std::vector<void*> allocs;
    float *W = nullptr, *gW = nullptr, *m = nullptr, *v = nullptr, *loss = nullptr;
    gnn::bf16x8* Wp = nullptr;      // node-MLP weights packed for inference
    bool wp_stale = true;           // W changed since Wp was packed
    int adam_t = 0;
    struct Graph {
        int n = 0, nw = 0;
        float *nf = nullptr, *raw = nullptr, *tgt = nullptr, *tw = nullptr;
        int *nb = nullptr, *nd = nullptr;
    };
    std::vector<Graph> samples;
    Graph cur;                      // graph of the state being advanced
    float* y = nullptr;             // [n][NO] network output
    int* node_of_id[2] = {nullptr, nullptr};
    int* slot_of_id[2] = {nullptr, nullptr};
    int* overflow = nullptr;
    unsigned int* err = nullptr;
    double* errsq = nullptr;        // [2]: sum of squared relative errors, count
    gnn::Scales S{};
    double tsum[gnn::NO] = {0, 0, 0, 0}, tsq[gnn::NO] = {0, 0, 0, 0};
    double tcount = 0.0;
    // [slot][phase]: 0 the last good state, 1 a judged advance's learned
    // result, 2 the state a judge starts from
    StateCopy snap[3][2];
    int64_t snap_rebuilds[3] = {-1, -1, -1};   // list rebuilds when each slot was taken
    // lagged sanity check of the learned advances: device flag, its pinned
    // copies (two slots) and their events; gain fit sums
    int* broken = nullptr;
    int* h_broken = nullptr;
    hipEvent_t ev_ok[2] = {nullptr, nullptr};
    double* fit = nullptr;
};

namespace {

template <class T>
int galloc(sphg_gnn* g, T** p, size_t count)
{
    void* q = nullptr;
    if (hipMalloc(&q, sizeof(T) * std::max<size_t>(count, 1)) != hipSuccess) return SPHG_ERR_OUT_OF_MEMORY;
    g->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return SPHG_OK;
}

int graph_alloc(sphg_gnn* g, sphg_gnn::Graph& G, int n)
{
    G.n = n;
    RC(galloc(g, &G.nf, (size_t)n * gnn::FN));
    RC(galloc(g, &G.raw, (size_t)n * gnn::RAW));
    RC(galloc(g, &G.tgt, (size_t)n * gnn::NO));
    RC(galloc(g, &G.tw, (size_t)n));
    RC(galloc(g, &G.nb, (size_t)n * gnn::DEG));
    RC(galloc(g, &G.nd, (size_t)n));
    return SPHG_OK;
}

int snap_alloc(sphg_gnn* g, StateCopy& s, int n)
{
    s.n = n;
    RC(galloc(g, &s.id, n));
    RC(galloc(g, &s.itype, n));
    RC(galloc(g, &s.fail, n));
    for (int i = 0; i < N_STATE_FIELDS; ++i) RC(galloc(g, &s.f[i], n));
    return SPHG_OK;
}

// to_snap: device state -> snapshot, else snapshot -> device state; one
// copy kernel per phase for all its fields
int snap_copy(sphg_gnn* g, int slot, bool to_snap)
{
    sphg_ctx* c = g->c;
    // both phases in one launch (an unbound phase takes no blocks)
    auto copy = [&](int k) {
        PhaseBuf& b = c->ph[k];
        PermArgs A{};
        if (!b.bound) return args_of(k_permute_multi, A, (const int*)nullptr, 0);
        StateCopy& sc = g->snap[slot][k];
        double** f[N_STATE_FIELDS];
        state_fields(b.d, f);
        int* st[3] = {b.d.id, b.d.itype, b.d.fail};
        int* sn[3] = {sc.id, sc.itype, sc.fail};
        A.nd = N_STATE_FIELDS;
        A.ni = 3;
        for (int i = 0; i < N_STATE_FIELDS; ++i) {
            A.sd[i] = to_snap ? *f[i] : sc.f[i];
            A.dd[i] = to_snap ? sc.f[i] : *f[i];
        }
        for (int i = 0; i < 3; ++i) {
            A.si[i] = to_snap ? st[i] : sn[i];
            A.di[i] = to_snap ? sn[i] : st[i];
        }
        return args_of(k_permute_multi, A, (const int*)nullptr, b.n);
    };
    auto nb = [&](int k) { return c->ph[k].bound ? nblk(c->ph[k].n) : 0; };
    RC(launch_pair<BodyPermute>(c, "snapshot", nb(0), copy(0), nb(1), copy(1)));
    if (to_snap) {
        g->snap_rebuilds[slot] = c->nl_rebuilds;
    } else if (!(nl_mode(c) && c->nl_valid && c->nl_rebuilds == g->snap_rebuilds[slot])) {
        // the neighbour lists (and their cell lists and record order) stay
        // valid for a state of their own rebuild epoch: the slot layout is the
        // same and the displacement check sees the restored positions
        state_changed(c);
    }
    return SPHG_OK;
}

// The snapshot of a phase seen as a DPhase (id, itype, vx, vy, rho).
DPhase snap_view(sphg_gnn* g, int slot, int k)
{
    DPhase d = g->c->ph[k].d;
    StateCopy& s = g->snap[slot][k];
    d.id = s.id;
    d.itype = s.itype;
    d.vx = s.f[2];
    d.vy = s.f[3];
    d.rho = s.f[6];
    return d;
}

// One leapfrog step of size K dt (the coarse step of a learned advance).
int coarse_step(sphg_ctx* c, int K)
{
    const double dt0 = c->K.dt;
    c->K.dt = dt0 * K;
    const int rc = one_step(c);
    c->K.dt = dt0;
    RC(rc);
    c->itimestep += K - 1;
    c->last_k = K;
    return SPHG_OK;
}

// K plain FP64 steps (targets and the judge), whatever precision the context uses.
int plain_steps(sphg_ctx* c, int n)
{
    const int prec = c->P.precision;
    c->P.precision = SPHG_FP64;
    int rc = SPHG_OK;
    for (int s = 0; s < n && rc == SPHG_OK; ++s) rc = one_step(c);
    c->P.precision = prec;
    return rc;
}

// Graph, features and extrapolation base of the current state into G.
int build_graph(sphg_gnn* g, sphg_gnn::Graph& G)
{
    sphg_ctx* c = g->c;
    PhaseBuf& Wb = c->ph[0];
    PhaseBuf& Sb = c->ph[1];
    // with valid neighbour lists the graph comes from them and their cell
    // lists stay; otherwise from cell lists of the current positions
    const bool lists = nl_mode(c) && c->nl_valid && Wb.cells_valid && Sb.cells_valid && Wb.cells_nl && Sb.cells_nl;
    if (!lists) {
        RC(build_cells(c, 0));
        RC(build_cells(c, 1));
    }
    G.nw = Wb.n;
    RC(fill_int(c, g->overflow, 1, 0));
    auto args = [&](int k) {
        PhaseBuf& b = c->ph[k];
        gnn::GatherOut O{G.nf, G.raw, G.nb, G.nd, G.n, G.tw, g->node_of_id[k], g->overflow};
        return args_of(gnn::k_gnn_gather, b.d, lists ? c->gvl[k] : c->gl[k], c->K, k == 0 ? c->F.thr_w : c->F.thr_s,
                       k == 0 ? c->F.uni_w : c->F.uni_s, k == 0 ? 0 : Wb.n, k, g->S, O,
                       lists ? b.nl_same : Masks{}, lists ? (const int*)b.nl_cell : (const int*)nullptr);
    };
    return launch_pair<BodyGnnGather>(c, "gnn_gather", nblk(Wb.n), args(0), nblk(Sb.n), args(1));
}

int forward(sphg_gnn* g, const sphg_gnn::Graph& G)
{
    if (g->wp_stale) {
        RC(launch(g->c, "gnn_pack", gnn::k_gnn_pack, nblk(gnn::NPACK), TPB, (const float*)g->W, g->Wp));
        g->wp_stale = false;
    }
    gnn::GnnArgs A{};
    A.G = gnn::Graph{G.n, G.nf, G.raw, G.nb, G.nd};
    A.W = g->W;
    A.Wp = g->Wp;
    A.out = g->y;
    for (int o = 0; o < gnn::NO; ++o) A.inv_out[o] = 1.0f / g->S.out[o];
    return launch(g->c, "gnn_fwd", gnn::k_gnn<false>, (G.n + 31) / 32, 64, A);
}

// The rate-change features of graph G once the coarse step has run (the
// state then holds the rates evaluated at the advance's start; gap: size in
// dt of the step that produced the rates the graph build saw).
int rate_features(sphg_gnn* g, sphg_gnn::Graph& G, int gap)
{
    sphg_ctx* c = g->c;
    const double coef = 0.5 * g->K * (g->K - 1) * c->P.dt / gap;
    auto args = [&](int k) {
        return args_of(gnn::k_gnn_delta, c->ph[k].d, (const int*)g->node_of_id[k], G.nf, coef);
    };
    return launch_pair<BodyGnnDelta>(c, "gnn_delta", nblk(c->ph[0].n), args(0), nblk(c->ph[1].n), args(1));
}

}  // namespace

extern "C" {

int sphg_gnn_nparams(void) { return gnn::NPARAM; }

int sphg_gnn_create(sphg_ctx* c, int ksteps, int max_samples, uint64_t seed, sphg_gnn** out)
{
    if (!c || !out || ksteps < 1 || max_samples < 0) return SPHG_ERR_INVALID;
    if (c->P.single_phase || !c->ph[0].bound || !c->ph[1].bound) return SPHG_ERR_UNSUPPORTED;
    if (c->legacy) return SPHG_ERR_UNSUPPORTED;
    sphg_gnn* g = new sphg_gnn();
    g->c = c;
    g->K = ksteps;
    g->maxs = max_samples;
    const int n = c->ph[0].n + c->ph[1].n;
    auto fail = [&](int rc) { sphg_gnn_destroy(g); return rc; };
    int rc;
    if ((rc = galloc(g, &g->W, gnn::NPARAM)) || (rc = galloc(g, &g->gW, gnn::NPARAM)) ||
        (rc = galloc(g, &g->m, gnn::NPARAM)) || (rc = galloc(g, &g->v, gnn::NPARAM)) ||
        (rc = galloc(g, &g->loss, 1)) || (rc = galloc(g, &g->Wp, gnn::NPACK)) ||
        (rc = galloc(g, &g->y, (size_t)n * gnn::NO)) || (rc = galloc(g, &g->overflow, 1)) ||
        (rc = galloc(g, &g->err, 1)) || (rc = galloc(g, &g->errsq, 2)) || (rc = graph_alloc(g, g->cur, n)) ||
        (rc = galloc(g, &g->broken, 1)) || (rc = galloc(g, &g->fit, 2 * gnn::NO)))
        return fail(rc);
    // events and pinned slots of the lagged sanity check
    if (hipEventCreateWithFlags(&g->ev_ok[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->ev_ok[1], hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc((void**)&g->h_broken, 2 * sizeof(int), hipHostMallocDefault) != hipSuccess)
        return fail(SPHG_ERR_HIP);
    g->h_broken[0] = g->h_broken[1] = 0;
    for (int k = 0; k < 2; ++k) {
        if ((rc = galloc(g, &g->node_of_id[k], c->ph[k].n)) || (rc = galloc(g, &g->slot_of_id[k], c->ph[k].n)))
            return fail(rc);
        for (int sl = 0; sl < 3; ++sl)
            if ((rc = snap_alloc(g, g->snap[sl][k], c->ph[k].n))) return fail(rc);
    }
    // scales: velocity sqrt(|g| H), density deviation in percent, lithostatic stress
    const double H = c->P.y_maxgeom - c->P.y_mingeom;
    const double gabs = std::fabs(c->K.gravity);
    const double vs = std::sqrt(gabs * H);
    const double kdt = ksteps * c->P.dt;
    g->S.inv_v = (float)(1.0 / vs);
    g->S.inv_a = (float)(kdt / vs);
    g->S.inv_r = 100.0f;
    g->S.inv_s = (float)(1.0 / (c->K.s_rho0 * gabs * H));
    g->S.kdt = (float)kdt;
    for (int o = 0; o < gnn::NO; ++o) g->S.out[o] = g->S.gain[o] = 1.0f;
    // Xavier-uniform weights from SplitMix64(seed), zero biases
    std::vector<float> w(gnn::NPARAM, 0.0f);
    uint64_t st = seed;
    auto uni = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    auto init = [&](int off, int fin, int fout, int cols, double gain) {
        const double a = gain * std::sqrt(6.0 / (fin + fout));
        for (int i = 0; i < fin; ++i)
            for (int j = 0; j < fout; ++j) w[(size_t)off + (size_t)i * cols + j] = (float)(a * (2.0 * uni() - 1.0));
    };
    init(gnn::O_W1, gnn::FE, gnn::HE, gnn::HE, 1.0);
    init(gnn::O_W2, gnn::HE, gnn::HE, gnn::HE, 1.0);
    init(gnn::O_W3, gnn::FI, gnn::D, gnn::D, 1.0);
    init(gnn::O_W4, gnn::D, gnn::D, gnn::D, 1.0);
    init(gnn::O_W5, gnn::D, gnn::NO, gnn::NOP, 0.1);
    if (hipMemcpy(g->W, w.data(), sizeof(float) * gnn::NPARAM, hipMemcpyHostToDevice) != hipSuccess)
        return fail(SPHG_ERR_HIP);
    if ((rc = launch(c, "fill", k_fill_f32, nblk(gnn::NPARAM), TPB, g->m, gnn::NPARAM, 0.0f)) ||
        (rc = launch(c, "fill", k_fill_f32, nblk(gnn::NPARAM), TPB, g->v, gnn::NPARAM, 0.0f)))
        return fail(rc);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(SPHG_ERR_HIP);
    *out = g;
    return SPHG_OK;
}

void sphg_gnn_destroy(sphg_gnn* g)
{
    if (!g) return;
    if (g->c) hipStreamSynchronize(g->c->stream);
    for (hipEvent_t e : g->ev_ok)
        if (e) hipEventDestroy(e);
    if (g->h_broken) hipHostFree(g->h_broken);
    for (void* p : g->allocs) hipFree(p);
    delete g;
}

int sphg_gnn_weights(sphg_gnn* g, float* host, int to_host)
{
    if (!g || !host) return SPHG_ERR_INVALID;
    sphg_ctx* c = g->c;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (to_host) HIPCHK(c, hipMemcpy(host, g->W, sizeof(float) * gnn::NPARAM, hipMemcpyDeviceToHost));
    else {
        HIPCHK(c, hipMemcpy(g->W, host, sizeof(float) * gnn::NPARAM, hipMemcpyHostToDevice));
        g->wp_stale = true;
    }
    return SPHG_OK;
}

int sphg_gnn_num_samples(const sphg_gnn* g) { return g ? (int)g->samples.size() : 0; }

int sphg_gnn_record(sphg_gnn* g)
{
    if (!g) return SPHG_ERR_INVALID;
    if ((int)g->samples.size() >= g->maxs) return SPHG_ERR_INVALID;
    sphg_ctx* c = g->c;
    RC(check_fused_switches(c));
    sphg_gnn::Graph G;
    RC(graph_alloc(g, G, c->ph[0].n + c->ph[1].n));
    const int step0 = c->itimestep;
    RC(snap_copy(g, 0, true));
    const int gap = c->last_k;
    RC(build_graph(g, G));
    RC(coarse_step(c, g->K));           // coarse state -> snapshot 1
    RC(rate_features(g, G, gap));
    RC(snap_copy(g, 1, true));
    RC(snap_copy(g, 0, false));
    c->itimestep = step0;
    RC(plain_steps(c, g->K));           // fine state: the run continues from here
    for (int k = 0; k < 2; ++k) {
        PhaseBuf& b = c->ph[k];
        const DPhase cv = snap_view(g, 1, k);
        RC(launch(c, "slot_of_id", gnn::k_slot_of_id, nblk(b.n), TPB, cv, g->slot_of_id[k]));
        RC(launch(c, "gnn_target", gnn::k_gnn_target, nblk(b.n), TPB, b.d, (const int*)g->node_of_id[k], cv,
                  (const int*)g->slot_of_id[k], G.tgt));
    }
    // running statistics of the targets (host side, for the output scales)
    std::vector<float> t((size_t)G.n * gnn::NO), w((size_t)G.n);
    HIPCHK(c, hipMemcpyAsync(t.data(), G.tgt, sizeof(float) * t.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(w.data(), G.tw, sizeof(float) * w.size(), hipMemcpyDeviceToHost, c->stream));
    int ovf = 0;
    HIPCHK(c, hipMemcpyAsync(&ovf, g->overflow, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < G.n; ++i) {
        if (w[i] == 0.0f) continue;
        for (int o = 0; o < gnn::NO; ++o) {
            g->tsum[o] += t[(size_t)i * gnn::NO + o];
            g->tsq[o] += (double)t[(size_t)i * gnn::NO + o] * t[(size_t)i * gnn::NO + o];
        }
        g->tcount += 1.0;
    }
    for (int o = 0; o < gnn::NO; ++o) {
        const double mean = g->tsum[o] / std::max(g->tcount, 1.0);
        const double var = g->tsq[o] / std::max(g->tcount, 1.0) - mean * mean;
        g->S.out[o] = (float)std::max(std::sqrt(std::max(var, 0.0)), 1e-12);
    }
    g->samples.push_back(G);
    return check_flags(c);
}

}  // extern "C"

namespace {

// Gain of each correction channel: the least-squares factor of the network's
// prediction against the recorded targets over every sample, within
// [0, 1.5].  A channel the network does not predict gets a gain near zero,
// so the correction never adds more error than it removes on the samples.
int calibrate_gain(sphg_gnn* g)
{
    sphg_ctx* c = g->c;
    RC(fill_dbl(c, g->fit, 2 * gnn::NO, 0.0));
    for (const sphg_gnn::Graph& G : g->samples) {
        RC(forward(g, G));
        RC(launch(c, "gnn_fit", gnn::k_gnn_fit, nblk(G.n, 256), 256, (const float*)g->y, (const float*)G.tgt,
                  (const float*)G.tw, G.n, g->S, g->fit));
    }
    double f[2 * gnn::NO] = {};
    HIPCHK(c, hipMemcpyAsync(f, g->fit, sizeof(f), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int o = 0; o < gnn::NO; ++o) {
        const double gain = f[gnn::NO + o] > 0.0 ? f[o] / f[gnn::NO + o] : 0.0;
        g->S.gain[o] = (float)std::min(std::max(gain, 0.0), 1.5);
    }
    return SPHG_OK;
}

}  // namespace

extern "C" {

int sphg_gnn_train(sphg_gnn* g, int iters, float lr, float* loss_out)
{
    if (!g || iters < 0) return SPHG_ERR_INVALID;
    if (g->samples.empty()) return SPHG_ERR_INVALID;
    sphg_ctx* c = g->c;
    double last = 0.0;
    const int ns = (int)g->samples.size();
    for (int it = 0; it < iters; ++it) {
        const sphg_gnn::Graph& G = g->samples[(size_t)(it % ns)];
        RC(launch(c, "fill", k_fill_f32, nblk(gnn::NPARAM), TPB, g->gW, gnn::NPARAM, 0.0f));
        RC(launch(c, "fill", k_fill_f32, 1, TPB, g->loss, 1, 0.0f));
        gnn::GnnArgs A{};
        A.G = gnn::Graph{G.n, G.nf, G.raw, G.nb, G.nd};
        A.W = g->W;
        A.tgt = G.tgt;
        A.tw = G.tw;
        A.gW = g->gW;
        A.loss = g->loss;
        for (int o = 0; o < gnn::NO; ++o) A.inv_out[o] = 1.0f / g->S.out[o];
        // the density correction is what the judge checks: weight it higher
        A.cw[0] = 1.0f; A.cw[1] = 1.0f; A.cw[2] = 4.0f; A.cw[3] = 0.0f;
        A.inv_count = 1.0f / (float)(std::max(g->tcount / ns, 1.0) * 6.0);
        RC(launch(c, "gnn_train", gnn::k_gnn<true>, (G.n + 31) / 32, 64, A));
        ++g->adam_t;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f / (1.0f - std::pow(b1, (float)g->adam_t));
        const float c2 = 1.0f / (1.0f - std::pow(b2, (float)g->adam_t));
        g->wp_stale = true;
        RC(launch(c, "adam", gnn::k_adam, nblk(gnn::NPARAM), TPB, g->W, (const float*)g->gW, g->m, g->v,
                  gnn::NPARAM, lr, b1, b2, 1e-8f, c1, c2, 10.0f));
        if (it + ns >= iters) {   // mean loss over the last pass through the samples
            float l = 0.0f;
            HIPCHK(c, hipMemcpyAsync(&l, g->loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            last += l * A.inv_count / std::min(ns, iters);
        }
    }
    if (loss_out) *loss_out = (float)last;
    return calibrate_gain(g);
}

int sphg_gnn_step(sphg_gnn* g, int use_net)
{
    if (!g) return SPHG_ERR_INVALID;
    sphg_ctx* c = g->c;
    RC(check_fused_switches(c));
    // graph and features of the state the advance starts from, the coarse
    // step (its rhs evaluates the rates at that state), the rate-change
    // features, the network, the correction
    const int gap = c->last_k;
    if (use_net) RC(build_graph(g, g->cur));
    RC(coarse_step(c, g->K));
    if (use_net) {
        RC(rate_features(g, g->cur, gap));
        RC(forward(g, g->cur));
        // (the moved positions reach the displacement check of the next
        // step's pack; without lists the step left its cell lists invalid)
        auto args = [&](int k) {
            return args_of(gnn::k_gnn_correct, c->ph[k].d, c->K, (const int*)g->node_of_id[k], (const float*)g->y,
                           g->S, k);
        };
        RC(launch_pair<BodyGnnCorrect>(c, "gnn_correct", nblk(c->ph[0].n), args(0), nblk(c->ph[1].n), args(1)));
    }
    return SPHG_OK;
}

// Sanity test of the state after a learned advance, without a host round
// trip: k_sanity raises the device flag, k_restore_if then copies the last
// good state (snapshot 0) back on the device, and the flag is copied to the
// pinned slot `slot` behind them.  The host reads it one advance later.
static int sanity_async(sphg_gnn* g, int slot)
{
    sphg_ctx* c = g->c;
    RC(fill_int(c, g->broken, 1, 0));
    const int nb0 = nblk(c->ph[0].n), nb1 = nblk(c->ph[1].n);
    auto check = [&](int k) { return args_of(gnn::k_sanity, c->ph[k].d, c->g, g->broken); };
    RC(launch_pair<BodySanity>(c, "sanity", nb0, check(0), nb1, check(1)));
    auto restore = [&](int k) {
        PhaseBuf& b = c->ph[k];
        StateCopy& sc = g->snap[0][k];
        double** f[N_STATE_FIELDS];
        state_fields(b.d, f);
        PermArgs A{};
        A.nd = N_STATE_FIELDS;
        A.ni = 3;
        for (int i = 0; i < N_STATE_FIELDS; ++i) { A.sd[i] = sc.f[i]; A.dd[i] = *f[i]; }
        A.si[0] = sc.id; A.si[1] = sc.itype; A.si[2] = sc.fail;
        A.di[0] = b.d.id; A.di[1] = b.d.itype; A.di[2] = b.d.fail;
        return args_of(gnn::k_restore_if, (const int*)g->broken, A, b.n);
    };
    RC(launch_pair<BodyRestoreIf>(c, "sanity", nb0, restore(0), nb1, restore(1)));
    HIPCHK(c, hipMemcpyAsync(&g->h_broken[slot], g->broken, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipEventRecord(g->ev_ok[slot], c->stream));
    return SPHG_OK;
}

int sphg_gnn_rollout(sphg_gnn* g, int nlearned, int judge_every, double tol, int use_net, int* nfallback,
                     double* max_err, double* rms_err)
{
    if (!g || nlearned < 0 || judge_every < 0) return SPHG_ERR_INVALID;
    sphg_ctx* c = g->c;
    int fb = 0;
    float worst = 0.0f;
    double sq = 0.0, cnt = 0.0;
    const int J = judge_every > 0 ? judge_every : nlearned + 1;
    int good_step = c->itimestep;   // itimestep of the last known-good state (snapshot 0)
    int since_good = 0;             // learned advances since that state
    int pending = -1;               // slot of the sanity flag not yet read
    RC(snap_copy(g, 0, true));
    // fallback: back to the last good state, plain steps over the span of
    // the advances since (the one that broke the state and any issued after)
    auto fallback = [&]() -> int {
        RC(snap_copy(g, 0, false));
        state_changed(c);
        RC(launch(c, "nlist", k_keep_bits, 1, 1, c->d_nlflag, 0));   // the broken span's displacement
        c->h_nlstep[0] = c->h_nlstep[1] = -1;
        c->itimestep = good_step;
        RC(plain_steps(c, since_good * g->K));
        fb += since_good;            // advances replaced by plain steps
        worst = std::max(worst, 1.0f);
        good_step = c->itimestep;
        since_good = 0;
        pending = -1;
        return snap_copy(g, 0, true);
    };
    // the sanity flag of the advance before the last (or of the last, at a
    // judge and at the end)
    auto resolve = [&](int slot) -> int {
        HIPCHK(c, hipEventSynchronize(g->ev_ok[slot]));
        if (g->h_broken[slot]) return fallback();
        return SPHG_OK;
    };
    // the step's own checks (a particle past half the list skin, or out of
    // the domain) reject a learned state like the sanity test does
    auto broken = [](int rc) { return rc == SPHG_ERR_MOVED_TOO_FAR || rc == SPHG_ERR_OUT_OF_RANGE; };
    for (int it = 0; it < nlearned; ++it) {
        const bool judge = judge_every > 0 && it % J == J - 1;
        if (!judge) {
            const int src = sphg_gnn_step(g, use_net);
            ++since_good;
            if (broken(src)) {
                RC(fallback());
                continue;
            }
            RC(src);
            const int slot = it & 1;
            RC(sanity_async(g, slot));
            const int prev = pending;
            pending = slot;
            if (prev >= 0) RC(resolve(prev));
            continue;
        }
        if (pending >= 0) {
            const int p = pending;
            pending = -1;
            RC(resolve(p));
        }
        // judge: one learned advance against K plain steps from the same
        // state.  The run goes on from the plain result (it re-anchors there);
        // the learned one is measured, and one over tol counts as a fallback.
        // The last good state (slot 0) stays until the learned step's pack
        // has checked the displacement of the state the judge starts from
        // (slot 2): past the skin there, or in an advance before, the span
        // since slot 0 is replaced by plain steps.
        const int step0 = c->itimestep;
        RC(snap_copy(g, 2, true));
        const int lrc = sphg_gnn_step(g, use_net);
        ++since_good;
        const int frc = broken(lrc) ? lrc : (lrc != SPHG_OK ? lrc : check_flags(c));
        if (broken(frc)) {
            RC(fallback());
            continue;
        }
        RC(frc);
        RC(snap_copy(g, 1, true));
        RC(snap_copy(g, 2, false));
        c->itimestep = step0;
        RC(plain_steps(c, g->K));
        RC(fill_int(c, reinterpret_cast<int*>(g->err), 1, 0));
        RC(fill_dbl(c, g->errsq, 2, 0.0));
        for (int k = 0; k < 2; ++k) {
            PhaseBuf& b = c->ph[k];
            RC(launch(c, "slot_of_id", gnn::k_slot_of_id, nblk(b.n), TPB, b.d, g->slot_of_id[k]));
            RC(launch(c, "rho_err", gnn::k_rho_err, nblk(b.n), TPB, snap_view(g, 1, k), b.d,
                      (const int*)g->slot_of_id[k], g->err, g->errsq));
        }
        unsigned int eb = 0;
        double es[2] = {0.0, 0.0};
        HIPCHK(c, hipMemcpyAsync(&eb, g->err, sizeof(eb), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(es, g->errsq, sizeof(es), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        sq += es[0];
        cnt += es[1];
        float e;
        std::memcpy(&e, &eb, sizeof(e));
        if (!(e <= tol)) ++fb;
        worst = std::max(worst, e == e ? e : 1.0f);
        good_step = c->itimestep;
        since_good = 0;
        RC(snap_copy(g, 0, true));
    }
    if (pending >= 0) {
        const int p = pending;
        pending = -1;
        RC(resolve(p));
    }
    {
        // the displacement checks of the last advances' packs
        const int frc = check_flags(c);
        if (broken(frc) && since_good > 0) RC(fallback());
        else RC(frc);
    }
    if (nfallback) *nfallback = fb;
    if (max_err) *max_err = worst;
    if (rms_err) *rms_err = cnt > 0.0 ? std::sqrt(sq / cnt) : 0.0;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return check_flags(c);
}

int sphg_gnn_gains(const sphg_gnn* g, float* gains)
{
    if (!g || !gains) return SPHG_ERR_INVALID;
    for (int o = 0; o < gnn::NO; ++o) gains[o] = g->S.gain[o];
    return SPHG_OK;
}

int sphg_gnn_sample_size(const sphg_gnn* g, int sample)
{
    if (!g || sample < 0 || sample >= (int)g->samples.size()) return -1;
    return g->samples[(size_t)sample].n;
}

int sphg_gnn_debug(sphg_gnn* g, int sample, float* nf, float* raw, int* nb, int* nd, float* tgt, float* tw,
                   float* y, float* grad, float* loss, float* out_scale)
{
    if (!g || sample < 0 || sample >= (int)g->samples.size()) return SPHG_ERR_INVALID;
    sphg_ctx* c = g->c;
    const sphg_gnn::Graph& G = g->samples[(size_t)sample];
    const size_t n = (size_t)G.n;
    RC(launch(c, "fill", k_fill_f32, nblk(gnn::NPARAM), TPB, g->gW, gnn::NPARAM, 0.0f));
    RC(launch(c, "fill", k_fill_f32, 1, TPB, g->loss, 1, 0.0f));
    RC(forward(g, G));
    gnn::GnnArgs A{};
    A.G = gnn::Graph{G.n, G.nf, G.raw, G.nb, G.nd};
    A.W = g->W;
    A.tgt = G.tgt;
    A.tw = G.tw;
    A.gW = g->gW;
    A.loss = g->loss;
    for (int o = 0; o < gnn::NO; ++o) {
        A.inv_out[o] = 1.0f / g->S.out[o];
        A.cw[o] = 1.0f;
    }
    A.inv_count = 1.0f;
    RC(launch(c, "gnn_train", gnn::k_gnn<true>, (G.n + 31) / 32, 64, A));
    auto dl = [&](void* dst, const void* src, size_t bytes) -> int {
        if (dst) HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
        return SPHG_OK;
    };
    RC(dl(nf, G.nf, n * gnn::FN * sizeof(float)));
    const bool scale_nf = nf != nullptr;
    RC(dl(raw, G.raw, n * gnn::RAW * sizeof(float)));
    RC(dl(nd, G.nd, n * sizeof(int)));
    if (nb) {   // the device table is slot-major; the caller gets [n][DEG]
        // (slots past a node's degree are not written on the device: -1 here)
        std::vector<int> t(n * gnn::DEG), d(n);
        HIPCHK(c, hipMemcpyAsync(t.data(), G.nb, t.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(d.data(), G.nd, d.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < n; ++i)
            for (int s = 0; s < gnn::DEG; ++s)
                nb[i * gnn::DEG + (size_t)s] = s < d[i] ? t[(size_t)s * n + i] : -1;
    }
    RC(dl(tgt, G.tgt, n * gnn::NO * sizeof(float)));
    RC(dl(tw, G.tw, n * sizeof(float)));
    RC(dl(y, g->y, n * gnn::NO * sizeof(float)));
    RC(dl(grad, g->gW, gnn::NPARAM * sizeof(float)));
    RC(dl(loss, g->loss, sizeof(float)));
    if (out_scale) for (int o = 0; o < gnn::NO; ++o) out_scale[o] = g->S.out[o];
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (scale_nf)   // the rate-change features as the network reads them (target units)
        for (size_t i = 0; i < n; ++i)
            for (int f = 0; f < 3; ++f) nf[i * gnn::FN + gnn::FR + (size_t)f] *= 1.0f / g->S.out[f];
    return SPHG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched scenarios: many independent contexts advanced by one hipGraph.
// (No counterpart in the reference: its host runs one case per process.)
// ---------------------------------------------------------------------------
struct sphg_batch {
    std::vector<sphg_ctx*> ctx;
    // One launch per kernel for all cases in every captured step (staged
    // form); off when a capture in that form found no room in the arena.
    bool staged = true;
    void* arena = nullptr;              // device: the batched launches' arguments and tile tables
    size_t arena_cap = 0;
    std::vector<hipStream_t> streams;   // contexts are dealt round-robin, largest first
    std::vector<int> order;             // contexts by decreasing particle count
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    int nstreams = 0;
    int period = 0;                     // steps per graph replay
    bool eager = false;                 // autotuner's choice: per-context launches, no graph
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
};

namespace {

// Batched form of a solo kernel: the size of its ArgTup and a launcher of
// k_batched over a tile table.
struct Batched {
    size_t argsz;
    hipError_t (*go)(hipStream_t, const void*, const int2*, int);
};

template <class F, class... T>
hipError_t go_batched(hipStream_t s, const void* a, const int2* t, int nb)
{
    hipLaunchKernelGGL((k_batched<F, T...>), dim3(nb), dim3(256), 0, s, static_cast<const ArgTup<T...>*>(a), t);
    return hipGetLastError();
}

template <class F, class... T>
void reg_batched(std::map<const void*, Batched>& m, void (*solo)(T...))
{
    m[reinterpret_cast<const void*>(solo)] = Batched{sizeof(ArgTup<T...>), &go_batched<F, T...>};
}

template <int SKF>
void reg_batched_skf(std::map<const void*, Batched>& m)
{
    reg_batched<BodyWaterX<SKF, true>>(m, k_water_x4<SKF, true>);
    reg_batched<BodyWaterX<SKF, false>>(m, k_water_x4<SKF, false>);
    reg_batched<BodySoilX<SKF, true>>(m, k_soil_x4<SKF, true>);
    reg_batched<BodySoilX<SKF, false>>(m, k_soil_x4<SKF, false>);
    reg_batched<BodyWaterA<SKF, true>>(m, k_water_a4<SKF, true>);
    reg_batched<BodyWaterA<SKF, false>>(m, k_water_a4<SKF, false>);
    reg_batched<BodyWaterB<SKF, true>>(m, k_water_b4<SKF, true>);
    reg_batched<BodyWaterB<SKF, false>>(m, k_water_b4<SKF, false>);
    reg_batched<BodySoilAB<SKF, true, false>>(m, k_soil_ab4<SKF, true, false>);
    reg_batched<BodySoilAB<SKF, false, false>>(m, k_soil_ab4<SKF, false, false>);
    reg_batched<BodyWaterAMx<SKF>>(m, k_water_a_mx<SKF>);
    reg_batched<BodyWaterBMx<SKF>>(m, k_water_b_mx<SKF>);
    reg_batched<BodySoilABMx<SKF>>(m, k_soil_ab_mx<SKF>);
    reg_batched<BodyWaterXMx<SKF>>(m, k_water_x_mx<SKF>);
    reg_batched<BodySoilXMx<SKF>>(m, k_soil_x_mx<SKF>);
}

// The kernels of a list-mode FP64 step between rebuilds (stages, finals).
const std::map<const void*, Batched>& batched_registry()
{
    static const std::map<const void*, Batched> m = [] {
        std::map<const void*, Batched> r;
        reg_batched_skf<1>(r);
        reg_batched_skf<2>(r);
        reg_batched_skf<3>(r);
        reg_batched_skf<4>(r);
        reg_batched<BodyStepWater>(r, k_step_water);
        reg_batched<BodyStepSoil>(r, k_step_soil);
        reg_batched<BodyFinalWater>(r, k_final_water);
        reg_batched<BodyFinalSoil>(r, k_final_soil);
        reg_batched<BodyFillInt>(r, k_fill_int);
        reg_batched<BodyKeepBits>(r, k_keep_bits);
        reg_batched<BodyIota>(r, k_iota);
        reg_batched<BodyPermute>(r, k_permute_multi);
        reg_batched<BodyScanTile>(r, k_scan_tile);
        reg_batched<BodyCellKey>(r, k_cell_key);
        reg_batched<BodyCellScatter>(r, k_cell_scatter_t);
        reg_batched<BodyCellSort>(r, k_cell_sort_t);
        reg_batched<BodyPackWater>(r, k_pack_water);
        reg_batched<BodyPackSoil>(r, k_pack_soil);
        reg_batched<BodyNlist>(r, k_nlist);
        return r;
    }();
    return m;
}

const Batched* batched_of(const void* kern)
{
    const auto& m = batched_registry();
    const auto it = m.find(kern);
    return it == m.end() ? nullptr : &it->second;
}

size_t batched_max_argsz()
{
    size_t mx = 0;
    for (const auto& kv : batched_registry()) mx = std::max(mx, kv.second.argsz);
    return mx;
}

// Host image of the batch arena: argument arrays and tile tables, each kept
// once (the arguments of a stage recur every other step: the two sound-speed
// buffers alternate), 256-byte aligned.
struct ArenaImage {
    std::vector<unsigned char> bytes;
    size_t cap = 0;
    std::unordered_map<uint64_t, std::vector<size_t>> index;   // content hash -> offsets
    static uint64_t hash(const unsigned char* p, size_t n)
    {
        uint64_t h = 1469598103934665603ull ^ n;   // FNV-1a
        for (size_t i = 0; i < n; ++i) h = (h ^ p[i]) * 1099511628211ull;
        return h;
    }
    // offset of the blob, or SIZE_MAX when the arena is full
    size_t put(const unsigned char* p, size_t n)
    {
        const uint64_t h = hash(p, n);
        auto& offs = index[h];
        for (size_t o : offs)
            if (std::memcmp(bytes.data() + o, p, n) == 0) return o;
        const size_t off = (bytes.size() + 255) & ~size_t(255);
        if (off + n > cap) return SIZE_MAX;
        bytes.resize(off + n);
        std::memcpy(bytes.data() + off, p, n);
        offs.push_back(off);
        return off;
    }
};

void batch_release_graph(sphg_batch* b)
{
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Streams and events of the batch are released only after their work has
// drained: a replay may still be running on them.
void batch_free_streams(sphg_batch* b)
{
    for (hipStream_t s : b->streams) hipStreamSynchronize(s);
    batch_release_graph(b);
    for (hipStream_t s : b->streams) hipStreamDestroy(s);
    for (hipEvent_t e : b->join) hipEventDestroy(e);
    b->streams.clear();
    b->join.clear();
}

int batch_set_streams(sphg_batch* b, int nstreams)
{
    batch_free_streams(b);
    b->nstreams = std::max(1, std::min(nstreams, (int)b->ctx.size()));
    for (int i = 0; i < b->nstreams; ++i) {
        hipStream_t s = nullptr;
        hipEvent_t e = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return SPHG_ERR_HIP;
        b->streams.push_back(s);
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return SPHG_ERR_HIP;
        b->join.push_back(e);
    }
    return SPHG_OK;
}

// True when a stream of the batch is still in a stream capture.
bool batch_capturing(const sphg_batch* b)
{
    for (hipStream_t s : b->streams) {
        hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &st) != hipSuccess || st != hipStreamCaptureStatusNone) return true;
    }
    return false;
}

// Steps of a captured period: the reorder period, or in the list mode the
// list rebuild period (each period starts with the reorder or rebuild, the
// one step whose launches differ; the host cannot act on the displacement
// check inside a replay, and a particle past half the skin still fails the
// run at its end).
int batch_period(const sphg_ctx* c) { return nl_config(c) ? c->nl_every : c->reorder_every; }

// Eager steps of one context until its next step starts a period.
int batch_align(sphg_ctx* c, int* left)
{
    while (*left > 0 && (c->itimestep < 1 || c->since_reorder != batch_period(c) - 1)) {
        RC(one_step(c));
        --*left;
    }
    return SPHG_OK;
}

// Capture `period` steps of every context in the staged form: each step is
// recorded for every context (the host runs the step as usual, the launches
// are held back) and issued as one batched launch per kernel of the step on
// the first stream, over all contexts.  A kernel without a batched form runs
// context by context there; a step whose kernel sequence differs between
// contexts is replayed context by context over the batch streams.
// SPHG_ERR_UNSUPPORTED: the arena is full; the host steps of the period have
// still run, as after any capture.
int batch_capture_staged(sphg_batch* b)
{
    const size_t n = b->ctx.size();
    if (!b->arena) {
        // room for 96 distinct argument arrays and 32 tile tables
        size_t blocks = 0;
        for (sphg_ctx* c : b->ctx) {
            blocks += (size_t)nblk(c->ph[0].n, 256) + (size_t)nblk(c->ph[1].n, 256);
            blocks += 2 * (size_t)nblk(std::max(c->ncell, c->ncell_v), 256);
        }
        const size_t cap = 96 * (n * batched_max_argsz() + 256) + 32 * (blocks * sizeof(int2) + 256);
        if (hipMalloc(&b->arena, cap) != hipSuccess) {
            b->arena = nullptr;
            return SPHG_ERR_HIP;
        }
        b->arena_cap = cap;
    }
    ArenaImage img;
    img.cap = b->arena_cap;
    struct Saved {
        hipStream_t stream;
        bool side;
        int it0;
    };
    std::vector<Saved> saved(n);
    for (size_t i = 0; i < n; ++i) saved[i] = {b->ctx[i]->stream, b->ctx[i]->use_side, b->ctx[i]->itimestep};
    hipStream_t origin = b->streams[0];
    if (hipStreamBeginCapture(origin, hipStreamCaptureModeThreadLocal) != hipSuccess) return SPHG_ERR_HIP;
    int rc = SPHG_OK;
    bool room = true;
    for (sphg_ctx* c : b->ctx) {
        c->use_side = false;
        c->capturing = true;
    }
    std::vector<std::vector<LaunchRec>> recs(n);
    std::vector<unsigned char> args;
    std::vector<int2> tiles;
    unsigned char* base = static_cast<unsigned char*>(b->arena);
    for (int s = 0; s < b->period && rc == SPHG_OK; ++s) {
        // (the host steps run even once the arena is full: the period ends
        // where a capture's period ends)
        for (size_t i = 0; i < n && rc == SPHG_OK; ++i) {
            recs[i].clear();
            b->ctx[i]->recording = &recs[i];
            rc = one_step(b->ctx[i], true);
            b->ctx[i]->recording = nullptr;
        }
        if (rc != SPHG_OK || !room) continue;
        const size_t nk = recs[0].size();
        bool same = true;
        for (size_t i = 1; i < n && same; ++i) {
            same = recs[i].size() == nk;
            for (size_t p = 0; p < nk && same; ++p)
                same = recs[i][p].kern == recs[0][p].kern && recs[i][p].block == recs[0][p].block;
        }
        if (!same) {
            // context by context, over the streams
            if (hipEventRecord(b->fork, origin) != hipSuccess) rc = SPHG_ERR_HIP;
            for (int k = 1; k < b->nstreams && rc == SPHG_OK; ++k)
                if (hipStreamWaitEvent(b->streams[(size_t)k], b->fork, 0) != hipSuccess) rc = SPHG_ERR_HIP;
            for (size_t r = 0; r < n && rc == SPHG_OK; ++r) {
                const size_t i = (size_t)b->order[r];
                hipStream_t st = b->streams[r % (size_t)b->nstreams];
                for (const LaunchRec& lr : recs[i])
                    if (rc == SPHG_OK && lr.solo(lr.kern, lr.grid, lr.block, st, lr.args.data()) != hipSuccess)
                        rc = SPHG_ERR_HIP;
            }
            for (int k = 1; k < b->nstreams && rc == SPHG_OK; ++k) {
                if (hipEventRecord(b->join[(size_t)k], b->streams[(size_t)k]) != hipSuccess ||
                    hipStreamWaitEvent(origin, b->join[(size_t)k], 0) != hipSuccess)
                    rc = SPHG_ERR_HIP;
            }
            continue;
        }
        for (size_t p = 0; p < nk && room && rc == SPHG_OK; ++p) {
            const Batched* bt = batched_of(recs[0][p].kern);
            if (!bt || recs[0][p].block != 256 || recs[0][p].args.size() != bt->argsz) {
                for (size_t i = 0; i < n && rc == SPHG_OK; ++i) {
                    const LaunchRec& lr = recs[i][p];
                    if (lr.solo(lr.kern, lr.grid, lr.block, origin, lr.args.data()) != hipSuccess) rc = SPHG_ERR_HIP;
                }
                continue;
            }
            args.clear();
            tiles.clear();
            for (size_t i = 0; i < n; ++i) {
                const LaunchRec& lr = recs[i][p];
                args.insert(args.end(), lr.args.begin(), lr.args.end());
                for (int k = 0; k < lr.grid; ++k) tiles.push_back(make_int2((int)i, k));
            }
            if (tiles.empty()) continue;
            const size_t oa = img.put(args.data(), args.size());
            const size_t ot = img.put(reinterpret_cast<const unsigned char*>(tiles.data()), tiles.size() * sizeof(int2));
            room = oa != SIZE_MAX && ot != SIZE_MAX;
            if (room && bt->go(origin, base + oa, reinterpret_cast<const int2*>(base + ot), (int)tiles.size()) !=
                            hipSuccess)
                rc = SPHG_ERR_HIP;
        }
    }
    for (size_t i = 0; i < n; ++i) {
        sphg_ctx* c = b->ctx[i];
        c->capturing = false;
        c->recording = nullptr;
        c->stream = saved[i].stream;
        c->use_side = saved[i].side;
        c->itimestep = saved[i].it0;   // the capture issued no work: the step count advances on replay
    }
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(origin, &g);
    if (rc != SPHG_OK || !room || ec != hipSuccess || !g || batch_capturing(b)) {
        if (g) hipGraphDestroy(g);
        if (batch_capturing(b)) {
            const int rs = batch_set_streams(b, b->nstreams);
            if (rs != SPHG_OK) return rs;
        }
        if (rc != SPHG_OK) return rc;
        return room ? SPHG_ERR_HIP : SPHG_ERR_UNSUPPORTED;
    }
    // the arguments and tile tables the batched launches read (no replay of
    // an earlier graph is in flight: runs end synchronised)
    for (hipStream_t st : b->streams) hipStreamSynchronize(st);
    if (hipMemcpy(b->arena, img.bytes.data(), img.bytes.size(), hipMemcpyHostToDevice) != hipSuccess) {
        hipGraphDestroy(g);
        return SPHG_ERR_HIP;
    }
    b->graph = g;
    if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) return SPHG_ERR_HIP;
    return SPHG_OK;
}

// Capture `period` steps of every context: staged (one launch per kernel for
// all contexts), else one stream per group of contexts.
// In the second form a context's main stream is a batch stream and its side
// stream is off: the graph overlaps whole contexts across the streams.  (With
// each context's side stream in the graph as well, the autotuner's best for
// the 64-case sweep was 8.6 ms per step, on one stream, against 3.8 ms
// without.)  The context's own streams never enter a capture.
int batch_capture(sphg_batch* b)
{
    batch_release_graph(b);
    if (b->staged) {
        const int rs = batch_capture_staged(b);
        if (rs != SPHG_ERR_UNSUPPORTED) return rs;
        b->staged = false;   // the arena was too small: the second form from here on
    }
    hipStream_t origin = b->streams[0];
    if (hipStreamBeginCapture(origin, hipStreamCaptureModeThreadLocal) != hipSuccess) return SPHG_ERR_HIP;
    int rc = SPHG_OK;
    if (hipEventRecord(b->fork, origin) != hipSuccess) rc = SPHG_ERR_HIP;
    for (int i = 1; i < b->nstreams && rc == SPHG_OK; ++i)
        if (hipStreamWaitEvent(b->streams[i], b->fork, 0) != hipSuccess) rc = SPHG_ERR_HIP;
    for (size_t r = 0; r < b->order.size() && rc == SPHG_OK; ++r) {
        sphg_ctx* c = b->ctx[(size_t)b->order[r]];
        hipStream_t own = c->stream;
        const bool own_side = c->use_side;
        const int it0 = c->itimestep;
        c->stream = b->streams[r % (size_t)b->nstreams];
        c->use_side = false;
        c->capturing = true;
        // (the last step of a period is followed by the next period's
        // rebuild, so it does not pack ahead: every replay starts unpacked)
        for (int s = 0; s < b->period && rc == SPHG_OK; ++s) rc = one_step(c, true);
        c->capturing = false;
        c->stream = own;
        c->use_side = own_side;
        c->itimestep = it0;   // the capture issued no work: the step count advances on replay
    }
    for (int i = 1; i < b->nstreams && rc == SPHG_OK; ++i) {
        if (hipEventRecord(b->join[(size_t)i], b->streams[(size_t)i]) != hipSuccess ||
            hipStreamWaitEvent(origin, b->join[(size_t)i], 0) != hipSuccess)
            rc = SPHG_ERR_HIP;
    }
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(origin, &g);
    if (rc != SPHG_OK || ec != hipSuccess || !g || batch_capturing(b)) {
        if (g) hipGraphDestroy(g);
        // streams left in a failed capture are not reused
        if (batch_capturing(b)) {
            const int ns = b->nstreams;
            const int rs = batch_set_streams(b, ns);
            if (rs != SPHG_OK) return rs;
        }
        return rc != SPHG_OK ? rc : SPHG_ERR_HIP;
    }
    b->graph = g;
    if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) return SPHG_ERR_HIP;
    return SPHG_OK;
}

}  // namespace

extern "C" {

int sphg_batch_create(sphg_ctx** ctxs, int n, sphg_batch** out)
{
    if (!ctxs || n < 1 || !out) return SPHG_ERR_INVALID;
    sphg_batch* b = new sphg_batch();
    for (int i = 0; i < n; ++i) {
        sphg_ctx* c = ctxs[i];
        // a captured period replays fixed step numbers, so the every-40-step
        // renormalisation (which persists under summation_density) cannot be
        // part of it
        if (!c || c->legacy || c->P.single_phase || c->prof || renorm_active(c->P)) {
            delete b;
            return c ? SPHG_ERR_UNSUPPORTED : SPHG_ERR_INVALID;
        }
        const int rc = check_fused_switches(c);
        if (rc != SPHG_OK) {
            delete b;
            return rc;
        }
        b->ctx.push_back(c);
    }
    b->order.resize((size_t)n);
    for (int i = 0; i < n; ++i) b->order[(size_t)i] = i;
    std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) {
        return b->ctx[(size_t)x]->ph[0].n + b->ctx[(size_t)x]->ph[1].n >
               b->ctx[(size_t)y]->ph[0].n + b->ctx[(size_t)y]->ph[1].n;
    });
    b->period = batch_period(b->ctx[0]);
    for (sphg_ctx* c : b->ctx) {
        if (batch_period(c) != b->period) {
            delete b;
            return SPHG_ERR_UNSUPPORTED;
        }
        if (hipStreamSynchronize(c->stream) != hipSuccess || hipStreamSynchronize(c->side) != hipSuccess) {
            delete b;
            return SPHG_ERR_HIP;
        }
    }
    for (sphg_ctx* c : b->ctx) c->stable_ptrs = true;
    // SPHG_BATCH_STAGED=0: per-context launches in the captured steps too
    const char* st = std::getenv("SPHG_BATCH_STAGED");
    b->staged = !(st && std::atoi(st) == 0);
    if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) {
        delete b;
        return SPHG_ERR_HIP;
    }
    const int rc = batch_set_streams(b, 4);
    if (rc != SPHG_OK) {
        sphg_batch_destroy(b);
        return rc;
    }
    *out = b;
    return SPHG_OK;
}

void sphg_batch_destroy(sphg_batch* b)
{
    if (!b) return;
    batch_free_streams(b);
    if (b->arena) hipFree(b->arena);
    if (b->fork) hipEventDestroy(b->fork);
    for (sphg_ctx* c : b->ctx) c->stable_ptrs = false;
    delete b;
}

int sphg_batch_staged(const sphg_batch* b) { return b ? (b->staged ? 1 : 0) : -1; }

int sphg_batch_set_streams(sphg_batch* b, int nstreams)
{
    if (!b || nstreams < 1) return SPHG_ERR_INVALID;
    return batch_set_streams(b, nstreams);
}

int sphg_batch_run(sphg_batch* b, int nsteps)
{
    if (!b || nsteps < 0) return SPHG_ERR_INVALID;
    if (b->eager) {
        // step-major over the contexts' own streams
        for (int s = 0; s < nsteps; ++s)
            for (sphg_ctx* c : b->ctx) RC(one_step(c));
        for (sphg_ctx* c : b->ctx) {
            HIPCHK(c, hipStreamSynchronize(c->stream));
            RC(check_flags(c));
        }
        return SPHG_OK;
    }
    std::vector<int> left(b->ctx.size(), nsteps);
    // eager steps up to the start of a reorder period
    for (size_t i = 0; i < b->ctx.size(); ++i) RC(batch_align(b->ctx[i], &left[i]));
    int reps = nsteps;
    for (size_t i = 0; i < b->ctx.size(); ++i) {
        sphg_ctx* c = b->ctx[i];
        const bool aligned = c->itimestep >= 1 && c->since_reorder == batch_period(c) - 1;
        reps = std::min(reps, aligned ? left[i] / b->period : 0);
    }
    if (reps > 0) {
        for (sphg_ctx* c : b->ctx) HIPCHK(c, hipStreamSynchronize(c->stream));
        if (!b->exec) RC(batch_capture(b));
        for (int r = 0; r < reps; ++r)
            if (hipGraphLaunch(b->exec, b->streams[0]) != hipSuccess) return SPHG_ERR_HIP;
        if (hipStreamSynchronize(b->streams[0]) != hipSuccess) return SPHG_ERR_HIP;
        for (size_t i = 0; i < b->ctx.size(); ++i) {
            b->ctx[i]->itimestep += reps * b->period;
            left[i] -= reps * b->period;
        }
    }
    // the remainder eagerly
    for (size_t i = 0; i < b->ctx.size(); ++i) {
        sphg_ctx* c = b->ctx[i];
        for (int s = 0; s < left[i]; ++s) RC(one_step(c));
    }
    for (sphg_ctx* c : b->ctx) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        RC(check_flags(c));
    }
    return SPHG_OK;
}

// Autotuner: times trial_periods graph periods for 1, 2, 4, ... streams (up
// to one per context) and the same steps launched eagerly on the contexts'
// own streams, and keeps the fastest.  Numerics do not depend on the choice.
int sphg_batch_tune(sphg_batch* b, int trial_periods, int* best_streams, double* best_ms_per_step)
{
    if (!b || trial_periods < 1) return SPHG_ERR_INVALID;
    int best = 0;   // 0: eager
    double best_ms = 1e300;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return SPHG_ERR_HIP;
    int rc = SPHG_OK;
    const int nctx = (int)b->ctx.size();
    // eager candidate
    b->eager = true;
    if ((rc = sphg_batch_run(b, 1)) == SPHG_OK) {
        const auto t0 = std::chrono::steady_clock::now();
        rc = sphg_batch_run(b, trial_periods * b->period);
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        best_ms = ms / (trial_periods * b->period);
    }
    b->eager = false;
    for (int ns = 1; ns <= nctx && rc == SPHG_OK; ns *= 2) {
        if ((rc = batch_set_streams(b, ns)) != SPHG_OK) break;
        std::vector<int> left(b->ctx.size(), 1 << 30);
        for (size_t i = 0; i < b->ctx.size() && rc == SPHG_OK; ++i) rc = batch_align(b->ctx[i], &left[i]);
        if (rc != SPHG_OK) break;
        for (sphg_ctx* c : b->ctx) hipStreamSynchronize(c->stream);
        if ((rc = batch_capture(b)) != SPHG_OK) break;
        hipGraphLaunch(b->exec, b->streams[0]);   // warm
        hipEventRecord(e0, b->streams[0]);
        for (int r = 0; r < trial_periods; ++r) hipGraphLaunch(b->exec, b->streams[0]);
        hipEventRecord(e1, b->streams[0]);
        hipEventSynchronize(e1);
        float ms = 0.0f;
        hipEventElapsedTime(&ms, e0, e1);
        for (sphg_ctx* c : b->ctx) c->itimestep += (trial_periods + 1) * b->period;
        const double per = ms / (trial_periods * b->period);
        if (per < best_ms) {
            best_ms = per;
            best = ns;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    RC(rc);
    b->eager = best == 0;
    if (!b->eager) RC(batch_set_streams(b, best));
    if (best_streams) *best_streams = best;
    if (best_ms_per_step) *best_ms_per_step = best_ms;
    return SPHG_OK;
}

}  // extern "C"

// ===========================================================================
// direct_find on caller-owned device arrays without any allocation.
//
// The scratch of this entry point (deinterleaved positions, the cell list,
// the per-particle partner counts and their prefix, the scan's tile sums and
// the flags) lives in one file-scope device array per GPU, `scratch`, of
// NOALLOC_SCRATCH_BYTES.  Host code never passes the array's name to a kernel
// (in host code that name is a stand-in, not a device address): the first
// call on each device resolves the array's address with
// hipGetSymbolAddress, with that device current, and keeps it per device.
// Every call on one GPU shares that GPU's array and overwrites it; calls that
// overlap on one GPU race.  What must start at zero (cell counts, flags) is
// cleared by fill kernels at every call.
// ===========================================================================
namespace {

constexpr int NOALLOC_MAX_N = SPHG_NOALLOC_MAX_N;
constexpr int NOALLOC_MAX_CELLS = SPHG_NOALLOC_MAX_CELLS;
constexpr int NOALLOC_MAX_TILES = (NOALLOC_MAX_CELLS > NOALLOC_MAX_N ? NOALLOC_MAX_CELLS : NOALLOC_MAX_N) / SCAN_TILE + 2;

// offsets (bytes) of the scratch regions; each 256-byte aligned
constexpr size_t na_align(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t NA_X = 0;
constexpr size_t NA_Y = NA_X + na_align(sizeof(double) * NOALLOC_MAX_N);
constexpr size_t NA_HMAX = NA_Y + na_align(sizeof(double) * NOALLOC_MAX_N);
constexpr size_t NA_ID = NA_HMAX + 256;
constexpr size_t NA_KEY = NA_ID + na_align(sizeof(int) * NOALLOC_MAX_N);
constexpr size_t NA_SLOT = NA_KEY + na_align(sizeof(int) * NOALLOC_MAX_N);
constexpr size_t NA_ORDER = NA_SLOT + na_align(sizeof(int) * NOALLOC_MAX_N);
constexpr size_t NA_FWD = NA_ORDER + na_align(sizeof(int) * NOALLOC_MAX_N);
constexpr size_t NA_OFF = NA_FWD + na_align(sizeof(int) * NOALLOC_MAX_N);
constexpr size_t NA_COUNT = NA_OFF + na_align(sizeof(int) * (NOALLOC_MAX_N + 1));
constexpr size_t NA_PART = NA_COUNT + na_align(sizeof(int) * NOALLOC_MAX_CELLS);
constexpr size_t NA_START = NA_PART + na_align(sizeof(int) * NOALLOC_MAX_CELLS);
constexpr size_t NA_TILES = NA_START + na_align(sizeof(int) * (NOALLOC_MAX_CELLS + 1));
constexpr size_t NA_FLAGS = NA_TILES + na_align(sizeof(int) * (NOALLOC_MAX_TILES + 1));
constexpr size_t NOALLOC_SCRATCH_BYTES = NA_FLAGS + 256;
static_assert(NOALLOC_SCRATCH_BYTES < (size_t)16 << 20, "the no-allocation scratch stays modest");

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

int na_fail(hipError_t e)
{
    noalloc_hip_error.store((int)e);
    return SPHG_ERR_HIP;
}

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mu;
unsigned char* noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of this GPU's scratch array (resolved once per device).
int noalloc_scratch(unsigned char** out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) return SPHG_ERR_HIP;
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (!noalloc_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return na_fail(e);
        if (!p) return na_fail(hipErrorInvalidSymbol);
        noalloc_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = noalloc_ptr[dev];
    return SPHG_OK;
}

// Largest of n doubles into out[0]: one workgroup (n <= NOALLOC_MAX_N).
__global__ __launch_bounds__(256) void k_max_dbl(const double* a, int n, double* out)
{
    __shared__ double red[256];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) m = fmax(m, a[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

#define NACHK(expr)                                   \
    do {                                              \
        hipError_t e_ = (expr);                       \
        if (e_ != hipSuccess) return na_fail(e_);     \
    } while (0)

template <typename... KArgs, typename... Args>
int na_launch(hipStream_t st, void (*kern)(KArgs...), int grid, int block, Args... args)
{
    if (grid <= 0) return SPHG_OK;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, st, args...);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? SPHG_OK : na_fail(e);
}

}  // namespace

extern "C" {

int sphg_direct_find_noalloc(const sphg_params* params, int n, const double* x, const double* hsml,
                             int max_interaction, int32_t* pair_i, int32_t* pair_j, double* w, double* dwdx,
                             int32_t* countiac, int32_t* niac, void* stream)
{
    if (!params || !x || !hsml || !countiac || !niac) return SPHG_ERR_INVALID;
    *niac = 0;
    RC(check_params(*params));
    if (n < 0 || n > NOALLOC_MAX_N || max_interaction < 0) return SPHG_ERR_INVALID;
    if (n == 0) return SPHG_OK;
    hipStream_t st = static_cast<hipStream_t>(stream);
    (void)hipGetLastError();   // an earlier call's error is not this call's
    unsigned char* base = nullptr;
    RC(noalloc_scratch(&base));
    double* xs = reinterpret_cast<double*>(base + NA_X);
    double* ys = reinterpret_cast<double*>(base + NA_Y);
    double* hmax_d = reinterpret_cast<double*>(base + NA_HMAX);
    int* id = reinterpret_cast<int*>(base + NA_ID);
    int* fwd = reinterpret_cast<int*>(base + NA_FWD);
    int* off = reinterpret_cast<int*>(base + NA_OFF);
    int* cell_part = reinterpret_cast<int*>(base + NA_PART);
    int* tiles = reinterpret_cast<int*>(base + NA_TILES);
    int* flags = reinterpret_cast<int*>(base + NA_FLAGS);

    // the grid: cells of at least the largest support radius over the
    // geometry box, coarsened when the box would need more cells than the
    // scratch holds (the pair list does not depend on the cell size)
    RC(na_launch(st, k_max_dbl, 1, 256, hsml, n, hmax_d));
    double hmax = 0.0;
    NACHK(hipMemcpyAsync(&hmax, hmax_d, sizeof(double), hipMemcpyDeviceToHost, st));
    NACHK(hipStreamSynchronize(st));
    if (!(hmax > 0.0)) return SPHG_ERR_INVALID;
    const sphg_params& P = *params;
    KParams K{};
    K.skf = P.skf;
    K.nnps = 1;
    K.scale_k = (P.skf == 1 || P.skf == 4) ? 2.0 : 3.0;
    const double dgx = P.x_maxgeom - P.x_mingeom, dgy = P.y_maxgeom - P.y_mingeom;
    double cs = K.scale_k * hmax * (1.0 + 1.0e-9);
    double fx = std::floor(dgx / cs) + 1.0, fy = std::floor(dgy / cs) + 1.0;
    while (fx * fy > (double)NOALLOC_MAX_CELLS) {
        cs *= 1.25;
        fx = std::floor(dgx / cs) + 1.0;
        fy = std::floor(dgy / cs) + 1.0;
    }
    Grid g{};
    g.x0 = P.x_mingeom; g.y0 = P.y_mingeom; g.x1 = P.x_maxgeom; g.y1 = P.y_maxgeom;
    g.inv_cs = 1.0 / cs;
    g.ncx = (int)fx;
    g.ncy = (int)fy;
    g.link_list = 0;
    const int ncell = g.ncx * g.ncy;
    const int ntiles = std::max((ncell + SCAN_TILE - 1) / SCAN_TILE, 1);
    const int ptiles = std::max((n + SCAN_TILE - 1) / SCAN_TILE, 1);

    DPhase d{};
    d.n = n;
    d.ntotal = n;
    d.x = xs;
    d.y = ys;
    d.hsml = const_cast<double*>(hsml);
    d.id = id;
    d.key = reinterpret_cast<int*>(base + NA_KEY);
    d.slot = reinterpret_cast<int*>(base + NA_SLOT);
    d.order = reinterpret_cast<int*>(base + NA_ORDER);
    d.cell_count = reinterpret_cast<int*>(base + NA_COUNT);
    d.cell_start = reinterpret_cast<int*>(base + NA_START);
    d.countiac = countiac;

    // positions in slot order (slot t holds particle t), cleared counts and flags
    RC(na_launch(st, k_iota, nblk(n), TPB, id, n));
    RC(na_launch(st, k_split2, nblk(n), TPB, x, xs, ys, (const int*)id, n));
    RC(na_launch(st, k_fill_int, nblk(ncell), TPB, d.cell_count, ncell, 0));
    RC(na_launch(st, k_fill_int, 1, TPB, flags, 1, 0));
    // cell list (the kernels of the context's build_cells)
    RC(na_launch(st, k_cell_key, nblk(n), TPB, d, g, flags));
    RC(na_launch(st, k_scan_tile, ntiles, SCAN_BLOCK, (const int*)d.cell_count, cell_part, ncell, tiles,
                 d.cell_count));
    RC(na_launch(st, k_cell_scatter_t, nblk(n, SCAN_BLOCK), SCAN_BLOCK, d, (const int*)cell_part,
                 (const int*)tiles, ntiles));
    RC(na_launch(st, k_cell_sort_t, nblk(ncell, SCAN_BLOCK), SCAN_BLOCK, d, (const int*)cell_part,
                 (const int*)tiles, ntiles, ncell));
    // partner counts, their prefix, the total (the kernels of materialise)
    RC(na_launch(st, k_pair_count, nblk(n), TPB, d, d, g, K, 1, 0, 0.0, fwd, countiac));
    RC(na_launch(st, k_scan_tile, ptiles, SCAN_BLOCK, (const int*)fwd, off, n, tiles, (int*)nullptr));
    RC(na_launch(st, k_scan_sums, 1, SCAN_BLOCK, tiles, ptiles));
    RC(na_launch(st, k_scan_add, nblk(n), TPB, off, n, (const int*)tiles, ptiles));
    int host[2] = {0, 0};
    NACHK(hipMemcpyAsync(&host[0], off + n, sizeof(int), hipMemcpyDeviceToHost, st));
    NACHK(hipMemcpyAsync(&host[1], flags, sizeof(int), hipMemcpyDeviceToHost, st));
    NACHK(hipStreamSynchronize(st));
    if (host[1] & 1) return SPHG_ERR_OUT_OF_RANGE;
    *niac = host[0];
    if (host[0] > max_interaction) return SPHG_ERR_TOO_MANY_INTERACTIONS;
    if (host[0] > 0 && (!pair_i || !pair_j || !w || !dwdx)) return SPHG_ERR_INVALID;
    RC(na_launch(st, k_pair_write, nblk(n), TPB, d, d, g, K, 1, 0, 0.0, (const int*)off, (int*)pair_i,
                 (int*)pair_j, w, dwdx));
    NACHK(hipStreamSynchronize(st));
    return SPHG_OK;
}

}  // extern "C"
