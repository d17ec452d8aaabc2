// sph_mixed.h -- mixed-precision form of the fused two-phase step.
//
// The state, the leapfrog update, the Drucker-Prager projections, the EOS,
// the flow rule and the Jaumann rate stay FP64 (pack and final kernels and
// the per-particle tails of the stage kernels).  The pair terms are FP32,
// summed in FP64, over FP32 neighbour records of four float4 chunks, stored
// as chunk planes like the FP64 records:
//   q0: xl, yl, tag, h         xl, yl: position relative to the centre of the
//                              particle's cell; tag: cx << 15 | cy, bit 31 set
//                              for a virtual particle (int bits)
//   q1: vx, vy, rho - rho0, m  (rho0 of the phase: differences keep precision)
//   q2: vof, 1/rho, m/rho, p
//   q3: c, s1, s2, s3          water: c, vof*sigma_xx, vof*sigma_xy, vof*sigma_yy
//                              soil:  c, sxx, sxy, syy
// A pair's separation is (xl_a - xl_b) + (cx_a - cx_b) * cs, exact up to the
// FP32 rounding of the cell-local coordinates (~1e-7 relative).  Pair
// selection uses the FP32 distance; pairs at r ~ 2h, where W and grad W
// vanish, may differ from the FP64 selection.
#pragma once
#include "sph_fused.h"

// Waves per SIMD asked of the FP32 water stage kernels (0: the compiler's
// choice).  With the cross terms in passes of their own, the compiler's
// choice (no spills) is about 1.5% faster on the step than six waves with
// 12-32 bytes of spills, and level with five (measured on the 1M slope).
#ifndef SPHG_MX_WAVES
#define SPHG_MX_WAVES 0
#endif
#if SPHG_MX_WAVES > 0
#define STAGE_ATTR_MX __launch_bounds__(STAGE_BLOCK) __attribute__((amdgpu_waves_per_eu(SPHG_MX_WAVES, 8)))
#else
#define STAGE_ATTR_MX __launch_bounds__(STAGE_BLOCK)
#endif
#ifndef SPHG_MX_AB_WAVES
#define SPHG_MX_AB_WAVES 0   // the one-pass soil kernel: compiler's choice (no spills) beats 4 waves with spills
#endif
#if SPHG_MX_AB_WAVES > 0
#define STAGE_ATTR_MX_AB __launch_bounds__(STAGE_BLOCK) __attribute__((amdgpu_waves_per_eu(SPHG_MX_AB_WAVES, 8)))
#else
#define STAGE_ATTR_MX_AB __launch_bounds__(STAGE_BLOCK)
#endif



// FP32 pair terms summed in FP64: a conversion and an add per term (with
// packed FP32 off, FP32 and FP64 adds issue at the same rate, so this costs
// half of a compensated FP32 sum and rounds only once per term, in FP64).
struct Sum64 {
    double s = 0.0;
    __device__ __forceinline__ void add(float v) { s += (double)v; }
    __device__ __forceinline__ double value() const { return s; }
};

struct KCf {
    float inv_h, f, fh1, fh2;
};

__device__ __forceinline__ KCf make_kcf(int skf, double h)
{
    const KC k = make_kc(skf, h);
    return KCf{(float)k.inv_h, (float)k.f, (float)k.fh1, (float)k.fh2};
}

// W and grad W in FP32 (the forms of kern_fast).
template <int SKF, bool NEED_W>
__device__ __forceinline__ void kern_f(const KCf& k, float r, float inv_r, float dx, float dy, float& w, float& gx,
                                       float& gy)
{
    const float q = r * k.inv_h;
    w = 0.0f;
    float s = 0.0f;
    if constexpr (SKF == 1) {
        const bool in1 = q <= 1.0f, in2 = q <= 2.0f;
        const float t = 2.0f - q;
        const float s1 = k.fh2 * (-2.0f + 1.5f * q);
        const float s2 = -(0.5f * k.fh1) * (t * t) * inv_r;
        s = in1 ? s1 : (in2 ? s2 : 0.0f);
        if (NEED_W) {
            const float w1 = k.f * ((float)SPH_C23_F - q * q + ((q * q) * q) * 0.5f);
            const float w2 = k.f * (1.0f / 6.0f) * ((t * t) * t);
            w = in1 ? w1 : (in2 ? w2 : 0.0f);
        }
    } else if constexpr (SKF == 4) {
        const bool in2 = q <= 2.0f;
        const float t = 1.0f - 0.5f * q;
        const float t2 = t * t;
        if (NEED_W) w = in2 ? k.f * ((t2 * t2) * (1.0f + 2.0f * q)) : 0.0f;
        s = k.fh2 * (-5.0f + 7.5f * q - 3.75f * (q * q) + 0.625f * ((q * q) * q));
        s = in2 ? s : 0.0f;
    } else if constexpr (SKF == 2) {
        if (q <= 3.0f) {
            const float e = k.f * __expf(-q * q);
            if (NEED_W) w = e;
            s = -2.0f * e * (k.inv_h * k.inv_h);
        }
    } else {
        const float a = 3.0f - q, b = 2.0f - q, c = 1.0f - q;
        const float a2 = a * a, b2 = b * b, c2 = c * c;
        const float a4 = a2 * a2, b4 = b2 * b2, c4 = c2 * c2;
        if (q <= 1.0f) {
            if (NEED_W) w = k.f * (a4 * a - 6.0f * (b4 * b) + 15.0f * (c4 * c));
            s = k.fh2 * (-120.0f + 120.0f * q - 50.0f * (q * q));  // reference polynomial
        } else if (q <= 2.0f) {
            if (NEED_W) w = k.f * (a4 * a - 6.0f * (b4 * b));
            s = k.fh1 * (-(5.0f * a4) + 30.0f * b4) * inv_r;
        } else if (q <= 3.0f) {
            if (NEED_W) w = k.f * (a4 * a);
            s = k.fh1 * (-(5.0f * a4)) * inv_r;
        }
    }
    gx = s * dx;
    gy = s * dy;
}

// Mixed-precision constants of a step.
struct MixK {
    float cs;                  // cell size
    float thr_w, thr_s, thr_x; // (scale_k h)^2 of uniform-h phases / cross pairs
    float scale_k;
    float rho0_w, rho0_s;
};

// Candidate rows of the 3 x 3 block around cell (cx, cy).
__device__ __forceinline__ LaneRows rows_of_cell(const Grid& g, const int* start, int cx, int cy)
{
    LaneRows L;
    const int c0 = max(cx - 1, 0), c1 = min(cx + 1, g.ncx - 1);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int row = cy - 1 + r;
        const bool ok = row >= 0 && row < g.ncy;
        L.k0[r] = ok ? start[row * g.ncx + c0] : 0;
        L.k1[r] = ok ? start[row * g.ncx + c1 + 1] : 0;
    }
    return L;
}

// Self view of a particle for the FP32 pair sums.
struct SelfF {
    float x, y, h;
    int cx, cy;
};

__device__ __forceinline__ SelfF self_of(float4 q0)
{
    const int tag = __float_as_int(q0.z);
    return SelfF{q0.x, q0.y, q0.w, tag_cx(tag), tag_cy(tag)};
}

// Separation a - b of a pair from cell-local coordinates.
__device__ __forceinline__ void sep(const SelfF& a, float4 b0, float cs, float& dx, float& dy)
{
    const int tb = __float_as_int(b0.z);
    dx = (a.x - b0.x) + (float)(a.cx - tag_cx(tb)) * cs;
    dy = (a.y - b0.y) + (float)(a.cy - tag_cy(tb)) * cs;
}

// Scan of one candidate row (FP32 predicate).
template <class P>
__device__ __forceinline__ uint64_t scan_row_f(const RecQ& rec, int k0, int k1, int self, P&& accept)
{
    uint64_t m = 0;
    if (k1 - k0 > 64) return 0;
#pragma unroll 4
    for (int k = k0; k < k1; ++k) {
        const bool ok = k != self && accept(rec(k, 0));
        m |= (uint64_t)ok << (k - k0);
    }
    return m;
}

__device__ __forceinline__ bool pred_f(bool uni, float thr, float dr, float ha, float hb, float scale_k)
{
    if (uni) return dr < thr;
    return __builtin_sqrtf(dr) < scale_k * (0.5f * (ha + hb));
}

// Partner iteration over a lane's CatMask, then the direct walk of overflowed rows.
template <class B, class P>
__device__ __forceinline__ void visit_f(const CatMask& cm, const LaneRows& L, int self, P&& accept, B&& body)
{
    CatIter it = cat_iter(cm, L);
#pragma unroll 1
    for (int k = it.next(); k >= 0; k = it.next()) body(k);
    if (cm.ovf) {
        // one pass over the concatenated rows (row bounds as scalars: an
        // indexed LaneRows would be kept in scratch memory)
        const int s1 = L.k1[0] - L.k0[0], s2 = s1 + L.k1[1] - L.k0[1], tot = s2 + L.k1[2] - L.k0[2];
        const int e0 = L.k0[0], e1 = L.k0[1] - s1, e2 = L.k0[2] - s2;
#pragma unroll 1
        for (int c = 0; c < tot; ++c) {
            const int k = c + (c < s1 ? e0 : (c < s2 ? e1 : e2));
            if (k != self && accept(k)) body(k);
        }
    }
}

// The exact FP32 partner mask from a neighbour-list mask (as filter_mask of
// the FP64 records): the listed candidates whose records pass accept(q0).
template <class A>
__device__ __forceinline__ CatMask filter_mask_f(const LaneRows& L, const CatMask& vm, const RecQ& rec, A&& acc)
{
    if (vm.ovf) return vm;
    const int s1 = L.k1[0] - L.k0[0], s2 = s1 + L.k1[1] - L.k0[1];
    const int e0 = L.k0[0], e1 = L.k0[1] - s1, e2 = L.k0[2] - s2;
    uint64_t lo = vm.lo, hi = vm.hi, olo = 0, ohi = 0;
#pragma unroll 1
    for (;;) {
        int c[4], k[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const bool in_lo = lo != 0;
            const uint64_t m = in_lo ? lo : hi;
            c[j] = m ? __builtin_ctzll(m) + (in_lo ? 0 : 64) : -1;
            const uint64_t rest = m & (m - 1);
            lo = in_lo ? rest : lo;
            hi = in_lo ? hi : rest;
            k[j] = c[j] < 0 ? -1 : c[j] + (c[j] < s1 ? e0 : (c[j] < s2 ? e1 : e2));
        }
        if (k[0] < 0) break;
        float4 q[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) q[j] = rec(k[j] >= 0 ? k[j] : k[0], 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (k[j] >= 0 && acc(q[j])) {
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

// Partner mask over the rows L: the list M filtered (neighbour lists on), or
// the scan of the three rows.
template <class A>
__device__ __forceinline__ CatMask cand_mask_f(const LaneRows& L, const Masks& M, int n, int tt, int self,
                                               const RecQ& rec, A&& acc)
{
    if (M.m) return filter_mask_f(L, mask_of(M, n, tt), rec, acc);
    const uint64_t m0 = scan_row_f(rec, L.k0[0], L.k1[0], self, acc);
    const uint64_t m1 = scan_row_f(rec, L.k0[1], L.k1[1], self, acc);
    const uint64_t m2 = scan_row_f(rec, L.k0[2], L.k1[2], self, acc);
    return cat_masks(L, m0, m1, m2);
}

// ------------------------------------------------------------------------
// Stage A, water (mixed): the terms of k_water_a4 with FP32 pair sums.
// ------------------------------------------------------------------------
// XM = 1: the same-phase terms (the cross terms come from the XM = 2 pass,
// which runs first and stores them); XM = 2: the cross terms alone.
template <int SKF, int XM>
__device__ __forceinline__ void water_a_mx(int blk, DPhase W, DPhase S, const float* qW, float* qWw, const float* qS,
                                           NList NL, Masks MW, Grid g, KParams K, FusedK F, MixK X, int cross)
{
    const RecQ QWv = recqv(qW, W.n), QSv = recqv(qS, S.n);
    const WaveCtx wc = wave_ctx(W.n, blk);
    const int t = wc.t;
    auto QA = [&](int c) { return QWv(wc.tt, c); };
    const float4 a0 = QA(0), a1 = QA(1), a2 = QA(2);
    const SelfF me = self_of(a0);
    const float vxa = a1.x, vya = a1.y, dra = a1.z;
    const float rhoa = dra + X.rho0_w;
    const int a = W.order[wc.tt];
    const bool uni = F.uni_w;
    const KCf kc = make_kcf(SKF, me.h);
    if constexpr (XM == 2) {
        Sum64 dg_x, dg_y, dvof_x, vof2;
        int cntx = 0;
        if (cross) {
            const LaneRows LX = rows_of_cell(g, S.cell_start, me.cx, me.cy);
            auto accx_q = [&](float4 b0) {
                float dx, dy;
                sep(me, b0, X.cs, dx, dy);
                return pred_f(F.uni_x, X.thr_x, dx * dx + dy * dy, me.h, b0.w, X.scale_k);
            };
            const CatMask cx3 = cand_mask_f(LX, NL.cross, W.n, wc.tt, -1, QSv, accx_q);
            const double vofa_d = W.vof[a], rhoa_d = W.rho[a];
            const float cfa = (float)(vofa_d * rhoa_d * (-K.numeric_gravity) / K.s_perm);
            const float irhoa = a2.y;
            const KCf kx = make_kcf(SKF, F.uni_x ? 0.5 * ((double)me.h + (double)QSv(0, 0).w) : (double)me.h);
            visit_f(cx3, LX, -1, [&](int k) { return accx_q(QSv(k, 0)); }, [&](int k) {
                auto Q = [&](int c) { return QSv(k, c); };
                const float4 b0 = Q(0), b1 = Q(1), b2 = Q(2);
                float dx, dy;
                sep(me, b0, X.cs, dx, dy);
                const float dr = dx * dx + dy * dy;
                const float ir = dr > 0.0f ? __builtin_amdgcn_rsqf(dr) : 0.0f;
                const float r = dr * ir;
                float ww, gx, gy;
                if (F.uni_x) kern_f<SKF, true>(kx, r, ir, dx, dy, ww, gx, gy);
                else kern_f<SKF, true>(make_kcf(SKF, 0.5 * ((double)me.h + (double)b0.w)), r, ir, dx, dy, ww, gx, gy);
                ++cntx;
                const float mb = b1.w;
                const float dvx = vxa - b1.x, dvy = vya - b1.y;
                const float sc = mb * (cfa * b2.x) * (ww * irhoa * b2.y);
                dg_x.add(sc * dvx);
                dg_y.add(sc * dvy);
                dvof_x.add(mb * (dvx * gx + dvy * gy));
                vof2.add(mb * ww);
            });
        }
        if (!wc.valid) return;
        if (cross) {
            W.drgx[a] = dg_x.value();
            W.drgy[a] = dg_y.value();
            W.dvof_x[a] = dvof_x.value() * (1.0 / K.s_rho0);
            W.vof2[a] = 1.0 - vof2.value() / K.s_rho0;
            W.countiac[a] = cntx;
        }
        return;
    }
    const LaneRows L = rows_of_cell(g, W.cell_start, me.cx, me.cy);
    auto acc_q = [&](float4 b0) {
        float dx, dy;
        sep(me, b0, X.cs, dx, dy);
        return pred_f(uni, X.thr_w, dx * dx + dy * dy, me.h, b0.w, X.scale_k);
    };
    const CatMask cm = cand_mask_f(L, NL.same, W.n, wc.tt, t, QWv, acc_q);
    if (wc.valid) {
        MW.m[t] = cm.lo;
        MW.m[W.n + t] = cm.hi;
        MW.ovf[t] = cm.ovf;
    }
    const bool real_a = W.itype[a] > 0;
    const float dspp = (float)W.dspp;
    Sum64 con, drd, txx, txy, tyy, avx, avy;
    float repx = 0.0f, repy = 0.0f;
    int cnt = 0;
    visit_f(cm, L, t, [&](int k) { return acc_q(QWv(k, 0)); }, [&](int k) {
        auto Q = [&](int c) { return QWv(k, c); };
        const float4 b0 = Q(0), b1 = Q(1);   // 1/rho and m/rho recomputed: two loads per partner
        float dx, dy;
        sep(me, b0, X.cs, dx, dy);
        const float dr = dx * dx + dy * dy;
        const float ir = __builtin_amdgcn_rsqf(dr);
        const float r = dr * ir, idr = ir * ir;
        float ww, gx, gy;
        if (uni) kern_f<SKF, true>(kc, r, ir, dx, dy, ww, gx, gy);
        else kern_f<SKF, true>(make_kcf(SKF, 0.5 * ((double)me.h + (double)b0.w)), r, ir, dx, dy, ww, gx, gy);
        ++cnt;
        const float mb = b1.w, irb = __builtin_amdgcn_rcpf(b1.z + X.rho0_w), Vb = mb * irb;
        const float dvx = vxa - b1.x, dvy = vya - b1.y;
        con.add(mb * (dvx * gx + dvy * gy));
        drd.add(mb * (2.0f * (dra - b1.z) * idr * (dx * gx + dy * gy)) * irb);
        txx.add(Vb * (-2.0f * dvx * gx + dvy * gy));
        txy.add(Vb * (-(dvx * gy) - dvy * gx));
        tyy.add(Vb * (-2.0f * dvy * gy + dvx * gx));
        const float f = mb * ww * (2.0f * __builtin_amdgcn_rcpf(rhoa + b1.z + X.rho0_w));
        avx.add(-(dvx * f));
        avy.add(-(dvy * f));
        if (real_a && !(r > dspp) && __float_as_int(b0.z) < 0) {
            const float q = dspp / r;
            const float fr = (__powf(q, (float)K.p1) - __powf(q, (float)K.p2)) * idr;
            repx += (float)K.dd * dx * fr;
            repy += (float)K.dd * dy * fr;
        }
    });

    if (!wc.valid) return;
    const double vofa_d = W.vof[a], ha_d = W.hsml[a], ca_d = W.c[a];
    const double mu = K.w_mu;
    const double txx_d = txx.value() * SPH_C23_F, tyy_d = tyy.value() * SPH_C23_F, txy_d = txy.value();
    W.drho[a] = K.artificial_density ? con.value() + (K.delta * ha_d * ca_d) * drd.value() : con.value();
    W.txx[a] = txx_d; W.txy[a] = txy_d; W.tyy[a] = tyy_d;
    double avx_d = avx.value(), avy_d = avy.value();
    if (real_a) { avx_d *= K.epsilon; avy_d *= K.epsilon; }
    if (K.average_velocity) { W.avx[a] = avx_d; W.avy[a] = avy_d; }
    W.repx[a] = repx; W.repy[a] = repy;
    if (!cross) W.countiac[a] = cnt;   // (with cross terms: the XM = 2 pass's count)
    W.nsame[a] = cnt;
    // p (q2.w) and the new c (q3.x) come from the pack
    float* Q3 = reinterpret_cast<float*>(reinterpret_cast<float4*>(qWw) + 3 * (size_t)W.n + t);
    Q3[1] = (float)(vofa_d * (mu * txx_d));
    Q3[2] = (float)(vofa_d * (mu * txy_d));
    Q3[3] = (float)(vofa_d * (mu * tyy_d));
}

template <int SKF>
__global__ STAGE_ATTR_MX void k_water_a_mx(DPhase W, DPhase S, const float* qW, float* qWw, const float* qS, NList NL,
                                           Masks MW, Grid g, KParams K, FusedK F, MixK X, int cross)
{
    water_a_mx<SKF, 1>(xcd_block(), W, S, qW, qWw, qS, NL, MW, g, K, F, X, cross);
}

// The water side's cross terms (soil partners), before k_water_a_mx.
template <int SKF>
__global__ STAGE_ATTR void k_water_x_mx(DPhase W, DPhase S, const float* qW, float* qWw, const float* qS, NList NL,
                                        Masks MW, Grid g, KParams K, FusedK F, MixK X, int cross)
{
    water_a_mx<SKF, 2>(xcd_block(), W, S, qW, qWw, qS, NL, MW, g, K, F, X, cross);
}

// ------------------------------------------------------------------------
// Stage B, water (mixed): momentum, art_visc, delta_sph(vof).
// ------------------------------------------------------------------------
template <int SKF>
__device__ __forceinline__ void water_b_mx(int blk, DPhase W, const float* qW, Masks MW, Grid g, KParams K, FusedK F,
                                           MixK X, int mask)
{
    const RecQ QWv = recqv(qW, W.n);
    const WaveCtx wc = wave_ctx(W.n, blk);
    const int t = wc.t;
    auto QA = [&](int c) { return QWv(wc.tt, c); };
    const float4 a0 = QA(0), a1 = QA(1), a2 = QA(2), a3 = QA(3);
    const SelfF me = self_of(a0);
    const float vxa = a1.x, vya = a1.y, rhoa = a1.z + X.rho0_w;
    const float vofa = a2.x, pa = a2.w, ca = a3.x, fxx_a = a3.y, fxy_a = a3.z, fyy_a = a3.w;
    const int a = W.order[wc.tt];
    const bool uni = F.uni_w;
    const KCf kc = make_kcf(SKF, me.h);
    const LaneRows L = rows_of_cell(g, W.cell_start, me.cx, me.cy);
    auto acc_q = [&](float4 b0) {
        float dx, dy;
        sep(me, b0, X.cs, dx, dy);
        return pred_f(uni, X.thr_w, dx * dx + dy * dy, me.h, b0.w, X.scale_k);
    };
    CatMask cm{0, 0, 0};
    if (wc.valid) {
        cm.lo = MW.m[t];
        cm.hi = MW.m[W.n + t];
        cm.ovf = MW.ovf[t];
    }
    const float et2 = (float)(K.etq * K.etq);
    const float beta = (float)K.beta, alpha = (float)K.alpha;
    Sum64 Px, Py, Sxx, Sxyy, Sxyx, Syy, artx, arty, dvd;
    visit_f(cm, L, t, [&](int k) { return acc_q(QWv(k, 0)); }, [&](int k) {
        auto Q = [&](int c) { return QWv(k, c); };
        const float4 b0 = Q(0), b1 = Q(1), b2 = Q(2), b3 = Q(3);
        float dx, dy;
        sep(me, b0, X.cs, dx, dy);
        const float dr = dx * dx + dy * dy;
        const float ir = __builtin_amdgcn_rsqf(dr);
        const float r = dr * ir, idr = ir * ir;
        float ww, gx, gy;
        const float amh = 0.5f * (me.h + b0.w);
        if (uni) kern_f<SKF, false>(kc, r, ir, dx, dy, ww, gx, gy);
        else kern_f<SKF, false>(make_kcf(SKF, amh), r, ir, dx, dy, ww, gx, gy);
        const float Vb = b2.z, mb = b1.w;
        const float ps = pa + b2.w;
        Px.add(Vb * (ps * gx));
        Py.add(Vb * (ps * gy));
        const float sxy = fxy_a + b3.z;
        Sxx.add(Vb * ((fxx_a + b3.y) * gx));
        Sxyy.add(Vb * (sxy * gy));
        Sxyx.add(Vb * (sxy * gx));
        Syy.add(Vb * ((fyy_a + b3.w) * gy));
        const float vr = (vxa - b1.x) * dx + (vya - b1.y) * dy;
        const float muv = amh * vr * __builtin_amdgcn_rcpf(dr + amh * amh * et2);
        const float mc = 0.5f * (ca + b3.x);
        float piv = (beta * muv - alpha * mc) * muv * (2.0f * __builtin_amdgcn_rcpf(rhoa + b1.z + X.rho0_w));
        piv = vr < 0.0f ? piv : 0.0f;
        artx.add(-(mb * piv * gx));
        arty.add(-(mb * piv * gy));
        dvd.add(mb * (2.0f * (vofa - b2.x) * idr * (dx * gx + dy * gy)) * b2.y);
    });
    if (!wc.valid) return;
    const double vofa_d = W.vof[a], rhoa_d = W.rho[a], ha_d = W.hsml[a], ca_d = W.c[a];   // new c
    double ax = -(vofa_d * Px.value()) + Sxx.value() + Sxyy.value();
    double ay = -(vofa_d * Py.value()) + Sxyx.value() + Syy.value();
    if (rhoa_d > 0.0) {
        ax = ax / rhoa_d;
        ay = ay / rhoa_d;
    }
    if (mask & M_ART) { ax = ax + artx.value(); ay = ay + arty.value(); }
    ax = ax + W.repx[a];
    ay = ay + W.repy[a];
    ay = ay + K.gravity;
    double dvof = (mask & M_DELTA_FIELD) ? (K.delta * ha_d * ca_d) * dvd.value() : 0.0;
    if (mask & M_CROSS) {
        ax = ax - W.drgx[a];
        ay = ay - W.drgy[a];
        dvof = dvof - W.dvof_x[a];
    }
    W.ax[a] = ax;
    W.ay[a] = ay;
    W.dvof[a] = dvof;
}

template <int SKF>
__global__ STAGE_ATTR_MX void k_water_b_mx(DPhase W, const float* qW, Masks MW, Grid g, KParams K, FusedK F, MixK X,
                                           int mask)
{
    water_b_mx<SKF>(xcd_block(), W, qW, MW, g, K, F, X, mask);
}

// ------------------------------------------------------------------------
// Soil, stages A and B in one kernel (mixed): soil stage B needs nothing
// from the stage A of other soil particles (its partners' p, sigma, vof come
// from the pack and the new sound speed is the constant s_c), so both stages
// share one scan and one partner loop; the terms are those of k_soil_ab4.
// ------------------------------------------------------------------------
// XM = 1: the same-phase terms and the tail (the cross terms from the XM = 2
// pass, which runs first and stores them); XM = 2: the cross terms alone.
template <int SKF, int XM>
__device__ __forceinline__ void soil_ab_mx(int blk, DPhase S, DPhase W, const float* qS, const float* qW, NList NL,
                                           Grid g, KParams K, FusedK F, MixK X, int cross, int mask)
{
    const RecQ QSv = recqv(qS, S.n), QWv = recqv(qW, W.n);
    const WaveCtx wc = wave_ctx(S.n, blk);
    const int t = wc.t;
    auto QA = [&](int c) { return QSv(wc.tt, c); };
    const float4 a0 = QA(0), a1 = QA(1), a2 = QA(2), a3 = QA(3);
    const SelfF me = self_of(a0);
    const float vxa = a1.x, vya = a1.y, dra = a1.z;
    const float rhoa = dra + X.rho0_s;
    const float vofa = a2.x, irhoa = a2.y, pa = a2.w, sxxa = a3.y, sxya = a3.z, syya = a3.w;
    const float ir2a = irhoa * irhoa;
    const float fpa = vofa * pa * ir2a, fxxa = vofa * sxxa * ir2a, fxya = vofa * sxya * ir2a,
                fyya = vofa * syya * ir2a;
    const float cn = (float)K.s_c;   // sound speed after this step (art_visc)
    const int a = S.order[wc.tt];
    const bool uni = F.uni_s;
    const KCf kc = make_kcf(SKF, me.h);
    const bool real_a = S.itype[a] > 0;
    const float dspp = (float)S.dspp;

    Sum64 dg_x, dg_y, po_x, po_y;
    int cntx = 0;
    const double vofa_d = S.vof[a];
    if (XM == 2 && cross) {
        const LaneRows LX = rows_of_cell(g, W.cell_start, me.cx, me.cy);
        auto accx_q = [&](float4 b0) {
            float dx, dy;
            sep(me, b0, X.cs, dx, dy);
            return pred_f(F.uni_x, X.thr_x, dx * dx + dy * dy, me.h, b0.w, X.scale_k);
        };
        const CatMask cx3 = cand_mask_f(LX, NL.cross, S.n, wc.tt, -1, QWv, accx_q);
        const float cfs = (float)(vofa_d * (-K.numeric_gravity) / K.s_perm);
        const KCf kx = make_kcf(SKF, F.uni_x ? 0.5 * ((double)QWv(0, 0).w + (double)me.h) : (double)me.h);
        visit_f(cx3, LX, -1, [&](int k) { return accx_q(QWv(k, 0)); }, [&](int k) {
            auto Q = [&](int c) { return QWv(k, c); };
            const float4 b0 = Q(0), b1 = Q(1), b2 = Q(2);
            float dx, dy;
            sep(me, b0, X.cs, dx, dy);
            dx = -dx;   // water - soil
            dy = -dy;
            const float dr = dx * dx + dy * dy;
            const float ir = dr > 0.0f ? __builtin_amdgcn_rsqf(dr) : 0.0f;
            const float r = dr * ir;
            float ww, gx, gy;
            if (F.uni_x) kern_f<SKF, true>(kx, r, ir, dx, dy, ww, gx, gy);
            else kern_f<SKF, true>(make_kcf(SKF, 0.5 * ((double)b0.w + (double)me.h)), r, ir, dx, dy, ww, gx, gy);
            ++cntx;
            const float mw = b1.w, rhow = b1.z + X.rho0_w;
            const float sc = mw * (b2.x * rhow * cfs) * (ww * b2.y * irhoa);
            dg_x.add(sc * (b1.x - vxa));
            dg_y.add(sc * (b1.y - vya));
            const float mprr = mw * b2.w * (b2.y * irhoa);
            po_x.add(mprr * gx);
            po_y.add(mprr * gy);
        });
    }

    if constexpr (XM == 2) {
        if (!wc.valid) return;
        if (cross) {
            S.drgx[a] = dg_x.value();
            S.drgy[a] = dg_y.value();
            S.porx[a] = po_x.value() * vofa_d;
            S.pory[a] = po_y.value() * vofa_d;
            S.countiac[a] = cntx;
        }
        return;
    }

    const LaneRows L = rows_of_cell(g, S.cell_start, me.cx, me.cy);
    auto acc_q = [&](float4 b0) {
        float dx, dy;
        sep(me, b0, X.cs, dx, dy);
        return pred_f(uni, X.thr_s, dx * dx + dy * dy, me.h, b0.w, X.scale_k);
    };
    const CatMask cm = cand_mask_f(L, NL.same, S.n, wc.tt, t, QSv, acc_q);
    const float et2 = (float)(K.etq * K.etq);
    const float beta = (float)K.beta, alpha = (float)K.alpha;
    Sum64 con, drd, txx, txy, tyy, vcc, wxy, avx, avy;
    Sum64 Fx, Fy, artx, arty, dP, dXX, dXY, dYY;
    float repx = 0.0f, repy = 0.0f;
    int cnt = 0;
    visit_f(cm, L, t, [&](int k) { return acc_q(QSv(k, 0)); }, [&](int k) {
        auto Q = [&](int c) { return QSv(k, c); };
        const float4 b0 = Q(0), b1 = Q(1), b2 = Q(2), b3 = Q(3);
        float dx, dy;
        sep(me, b0, X.cs, dx, dy);
        const float dr = dx * dx + dy * dy;
        const float ir = __builtin_amdgcn_rsqf(dr);
        const float r = dr * ir, idr = ir * ir;
        float ww, gx, gy;
        const float amh = 0.5f * (me.h + b0.w);
        if (uni) kern_f<SKF, true>(kc, r, ir, dx, dy, ww, gx, gy);
        else kern_f<SKF, true>(make_kcf(SKF, 0.5 * ((double)me.h + (double)b0.w)), r, ir, dx, dy, ww, gx, gy);
        ++cnt;
        const float mb = b1.w, vofb = b2.x, irb = b2.y, Vb = b2.z;
        // stage A
        const float dvx = vxa - b1.x, dvy = vya - b1.y;
        const float dvg = dvx * gx + dvy * gy;
        con.add(mb * dvg);
        drd.add(mb * (2.0f * (dra - b1.z) * idr * (dx * gx + dy * gy)) * irb);
        txx.add(Vb * (-2.0f * dvx * gx + dvy * gy));
        txy.add(Vb * (-(dvx * gy) - dvy * gx));
        tyy.add(Vb * (-2.0f * dvy * gy + dvx * gx));
        vcc.add(-(Vb * dvg));
        wxy.add(Vb * (-(dvx * gy) + dvy * gx));
        const float rs = 2.0f * __builtin_amdgcn_rcpf(rhoa + b1.z + X.rho0_s);
        const float f = mb * ww * rs;
        avx.add(-(dvx * f));
        avy.add(-(dvy * f));
        if (real_a && !(r > dspp) && __float_as_int(b0.z) < 0) {
            const float q = dspp / r;
            const float fr = (__powf(q, (float)K.p1) - __powf(q, (float)K.p2)) * idr;
            repx += (float)K.dd * dx * fr;
            repy += (float)K.dd * dy * fr;
        }
        // stage B
        const float ir2b = irb * irb;
        const float fp = fpa + vofb * b2.w * ir2b;
        const float fxx = fxxa + vofb * b3.y * ir2b;
        const float fxy = fxya + vofb * b3.z * ir2b;
        const float fyy = fyya + vofb * b3.w * ir2b;
        Fx.add(mb * ((fxx - fp) * gx + fxy * gy));
        Fy.add(mb * ((fyy - fp) * gy + fxy * gx));
        const float vr = dvx * dx + dvy * dy;
        const float muv = amh * vr * __builtin_amdgcn_rcpf(dr + amh * amh * et2);
        float piv = (beta * muv - alpha * cn) * muv * rs;
        piv = vr < 0.0f ? piv : 0.0f;
        artx.add(-(mb * piv * gx));
        arty.add(-(mb * piv * gy));
        const float hh = 2.0f * idr * (dx * gx + dy * gy) * mb * irb;
        dP.add(hh * (pa - b2.w));
        dXX.add(hh * (sxxa - b3.y));
        dXY.add(hh * (sxya - b3.z));
        dYY.add(hh * (syya - b3.w));
    });
    if (!wc.valid) return;
    const double ha_d = S.hsml[a], ca_d = S.c[a], rhoa_d = S.rho[a];
    const double txx_d = txx.value() * SPH_C23_F, tyy_d = tyy.value() * SPH_C23_F, txy_d = txy.value();
    const double vcc_d = vcc.value(), wxy_d = wxy.value() * 0.5;
    S.drho[a] = K.artificial_density ? con.value() + (K.delta * ha_d * ca_d) * drd.value() : con.value();
    S.txx[a] = txx_d; S.txy[a] = txy_d; S.tyy[a] = tyy_d; S.vcc[a] = vcc_d; S.wxy[a] = wxy_d;
    double avx_d = avx.value(), avy_d = avy.value();
    if (real_a) { avx_d *= K.epsilon; avy_d *= K.epsilon; }
    if (K.average_velocity) { S.avx[a] = avx_d; S.avy[a] = avy_d; }
    S.repx[a] = repx; S.repy[a] = repy;
    // the cross terms the XM = 2 pass stored
    const double dgx_d = cross ? S.drgx[a] : 0.0, dgy_d = cross ? S.drgy[a] : 0.0;
    const double pox_d = cross ? S.porx[a] : 0.0, poy_d = cross ? S.pory[a] : 0.0;
    if (!cross) S.countiac[a] = cnt;
    S.nsame[a] = cnt;
    S.c_next[a] = K.s_c;
    const double p = S.p[a], sxx = S.sxx[a], sxy = S.sxy[a], syy = S.syy[a];
    const double alpha_d = K.s_alpha1, G = K.s_G, Kb = K.s_k;
    double dsxx = 0.0, dsxy = 0.0, dsyy = 0.0, dpv = 0.0;
    if (K.plasticity == 3) {
        const double I1 = 3.0 * p;
        const double j2r = sxx * sxx + 2.0 * (sxy * sxy) + syy * syy + (sxx + syy) * (sxx + syy);
        const double J2 = __builtin_sqrt(j2r / 2.0) + 1.0e-6;
        const int fail = fabs(J2 - alpha_d * I1) < SPH_YIELD_TOL_F ? 1 : 0;
        if (fail) {
            const double exx = txx_d * 0.5 + vcc_d * (1.0 / 3.0);
            const double exy = txy_d * 0.5;
            const double eyy = tyy_d * 0.5 + vcc_d * (1.0 / 3.0);
            const double sde = sxx * exx + 2.0 * sxy * exy + syy * eyy;
            const double sj = __builtin_sqrt(j2r / 2.0 + 1.0e-10);
            const double gs = G / sj;
            const double dl = (3.0 * alpha_d * Kb * vcc_d + gs * sde) / (9.0 * (alpha_d * alpha_d) * Kb + G);
            if (!(dl < 0.0)) {
                dsxx = -dl * (gs * sxx);
                dsxy = -dl * (gs * sxy);
                dsyy = -dl * (gs * syy);
                dpv = 3.0 * Kb * alpha_d * dl;
                S.eps_p[a] = S.eps_p[a] + dl * sxy / (2.0 * sj) * SPH_EPSP_F;
            }
        }
        S.fail[a] = fail;
    }
    dsxx = dsxx + G * txx_d + 2.0 * sxy * wxy_d;
    dsxy = dsxy + G * txy_d - (sxx - syy) * wxy_d;
    dsyy = dsyy + G * tyy_d - 2.0 * sxy * wxy_d;
    dpv = dpv - Kb * vcc_d;

    double ax = Fx.value() * rhoa_d;
    double ay = Fy.value() * rhoa_d;
    if (rhoa_d > 0.0) { ax = ax / rhoa_d; ay = ay / rhoa_d; }
    if (mask & M_ART) { ax = ax + artx.value(); ay = ay + arty.value(); }
    ax = ax + (double)repx;
    ay = ay + (double)repy;
    ay = ay + K.gravity;
    if (mask & M_CROSS) {
        ax = ax + dgx_d; ay = ay + dgy_d;
        ax = ax + pox_d; ay = ay + poy_d;
    }
    if (mask & M_DELTA_FIELD) {
        const double cf = K.delta * ha_d * K.s_c;   // new c
        dpv = dpv + cf * dP.value();
        dsxx = dsxx + cf * dXX.value();
        dsxy = dsxy + cf * dXY.value();
        dsyy = dsyy + cf * dYY.value();
    }
    S.dsxx[a] = dsxx; S.dsxy[a] = dsxy; S.dsyy[a] = dsyy; S.dp[a] = dpv;
    S.ax[a] = ax;
    S.ay[a] = ay;
}

template <int SKF>
__global__ STAGE_ATTR_MX_AB void k_soil_ab_mx(DPhase S, DPhase W, const float* qS, const float* qW, NList NL, Grid g,
                                              KParams K, FusedK F, MixK X, int cross, int mask)
{
    soil_ab_mx<SKF, 1>(xcd_block(), S, W, qS, qW, NL, g, K, F, X, cross, mask);
}

// The soil side's cross terms (water partners), before k_soil_ab_mx.
template <int SKF>
__global__ STAGE_ATTR void k_soil_x_mx(DPhase S, DPhase W, const float* qS, const float* qW, NList NL, Grid g,
                                       KParams K, FusedK F, MixK X, int cross, int mask)
{
    soil_ab_mx<SKF, 2>(xcd_block(), S, W, qS, qW, NL, g, K, F, X, cross, mask);
}
