// sph_gnn.h -- learned multi-step surrogate on the cell-list graph.
//
// A message-passing network advances the particles K leapfrog steps at once:
// one leapfrog step of size K dt (the coarse step, one right-hand side), plus
// the network's correction of velocity and density towards K steps of size
// dt:  v += c_v,  rho += c_rho,  x += K dt c_v / 2.
// Its graph is the same-phase neighbour graph of the SPH step (pairs with
// r < 2h, at most DEG per node); it sees node features (rates, velocity,
// density, stress) and edge features (relative position / h, velocity,
// density and rate differences) of the state the advance starts from.
//
// Network (weights FP32 master copy, BF16 operands, FP32 accumulation):
//   edge layer  z_ij = e_ij W1            (e_ij: FE features, e_ij[0] = 1 = bias)
//   message     m_i  = (sum_j relu(z_ij) / 16) W2
//   node MLP    y_i  = W5 relu(W4 relu(W3 [n_i, m_i] + b3) + b4) + b5
// Since W2 is linear, the sum over edges is taken before it: the per-edge
// work is one 32 x 32 x 16 MFMA per neighbour slot for 32 particles.
//
// One wavefront handles 32 particles.  Every matrix product is a
// v_mfma_f32_32x32x16_bf16 tile product whose operands are read through
// accessors (LDS or global), so forward and backward (training) share one
// primitive.  MFMA operand layout (gfx950, wave64): lane l holds row l % 32
// of A and column l % 32 of B, k = (l / 32) * 8 + [0, 8); the 16 results
// of lane l are C[(v / 4) * 8 + (l / 32) * 4 + v % 4][l % 32], v = 0..15.
#pragma once
#include "sph_fused.h"

namespace gnn {

constexpr int FE = 16;    // edge input features
constexpr int HE = 32;    // edge hidden width = message width
constexpr int FN = 32;    // node input features
constexpr int FR = 16;    // features FR..FR+2: rate-change estimates of the correction (vx, vy, rho)
constexpr int FI = FN + HE;  // node MLP input width
constexpr int D = 64;     // node hidden width
constexpr int NO = 4;     // outputs used: c_vx, c_vy, c_rho, (spare)
constexpr int NOP = 32;   // output tile width
constexpr int DEG = 48;   // neighbour slots per node
constexpr int RAW = 12;   // per-node floats feeding the edge features
constexpr float EDGE_NORM = 1.0f / 16.0f;

// parameter offsets (floats), row-major [in][out]
constexpr int O_W1 = 0;
constexpr int O_W2 = O_W1 + FE * HE;
constexpr int O_W3 = O_W2 + HE * HE;
constexpr int O_B3 = O_W3 + FI * D;
constexpr int O_W4 = O_B3 + D;
constexpr int O_B4 = O_W4 + D * D;
constexpr int O_W5 = O_B4 + D;
constexpr int O_B5 = O_W5 + D * NOP;
constexpr int NPARAM = O_B5 + NOP;

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ f32x16 zero16()
{
    f32x16 z;
#pragma unroll
    for (int v = 0; v < 16; ++v) z[v] = 0.0f;
    return z;
}

// acc += A (32 x K) B (K x 32); a(row, k), b(k, col) return FP32 values that
// are rounded to BF16 operands.  K is a multiple of 16.
template <class FA, class FB>
__device__ __forceinline__ f32x16 mma(FA&& a, FB&& b, int K, f32x16 acc)
{
    const int l = lane_id(), r = l & 31, h = l >> 5;
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x8 av, bv;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            av[i] = (__bf16)a(r, k0 + h * 8 + i);
            bv[i] = (__bf16)b(k0 + h * 8 + i, r);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    }
    return acc;
}

// Weights of the node MLP as BF16 B operands for inference: for each matrix,
// n-tile and 16-row k-step, 64 lanes x 8 values (lane l: column l & 31 of
// the tile, rows k0 + (l >> 5) * 8 + 0..7) -- the values mma() rounds from the
// FP32 weights, one 16-byte load per lane and step.
constexpr int P_W2 = 0;                // HE x HE: 1 tile x 2 steps
constexpr int P_W3 = P_W2 + 2 * 64;    // FI x D: 2 tiles x 4 steps
constexpr int P_W4 = P_W3 + 8 * 64;    // D x D: 2 tiles x 4 steps
constexpr int P_W5 = P_W4 + 8 * 64;    // D x NOP: 1 tile x 4 steps
constexpr int NPACK = P_W5 + 4 * 64;

// acc += A (32 x K) Bp, Bp: the packed operands of one n-tile (K / 16 steps).
template <class FA>
__device__ __forceinline__ f32x16 mma_p(FA&& a, const bf16x8* bp, int K, f32x16 acc)
{
    const int l = lane_id(), r = l & 31, h = l >> 5;
    for (int k0 = 0, ks = 0; k0 < K; k0 += 16, ++ks) {
        bf16x8 av;
#pragma unroll
        for (int i = 0; i < 8; ++i) av[i] = (__bf16)a(r, k0 + h * 8 + i);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bp[ks * 64 + l], acc, 0, 0, 0);
    }
    return acc;
}

// f(row, col, value&) over the 16 results a lane holds.
template <class F>
__device__ __forceinline__ void tile_each(f32x16& c, F&& f)
{
    const int l = lane_id(), col = l & 31, h = l >> 5;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        float x = c[v];
        f((v >> 2) * 8 + h * 4 + (v & 3), col, x);
        c[v] = x;
    }
}

__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

// Scales of the features and outputs (set from the first recorded sample).
struct Scales {
    float inv_v;     // 1 / velocity scale
    float inv_a;     // K dt / velocity scale (acceleration -> velocity change)
    float inv_r;     // 1 / density-deviation scale (relative)
    float inv_s;     // 1 / stress scale
    float kdt;       // K dt
    float out[NO];   // output scales of the training targets (target = y * out)
    float gain[NO];  // calibrated shrinkage of the applied correction: c = y * out * gain
};

// One graph: nodes [0, n) = water (cell order) then soil.
struct Graph {
    int n;
    const float* nf;   // [n][FN]
    const float* raw;  // [n][RAW]
    const int* nb;     // [DEG][n]: slot-major, so a wave's 32 nodes read one slot contiguously
    const int* nd;     // [n]
};

// Edge features of node i (raw ri) and neighbour j (raw rj); zero when j < 0.
__device__ __forceinline__ void edge_features(const float* ri, const float* rj, bool has, float* e)
{
#pragma unroll
    for (int f = 0; f < FE; ++f) e[f] = 0.0f;
    if (!has) return;
    const float ih = 1.0f / ri[7];
    const float dx = ri[0] - rj[0], dy = ri[1] - rj[1];
    const float r = sqrtf(dx * dx + dy * dy);
    const float q = fminf(r * ih * 0.5f, 1.0f);
    e[0] = 1.0f;
    e[1] = dx * ih;
    e[2] = dy * ih;
    e[3] = r * ih;
    e[4] = (1.0f - q) * (1.0f - q);
    e[5] = ri[2] - rj[2];
    e[6] = ri[3] - rj[3];
    e[7] = ri[4] - rj[4];
    e[8] = ri[5] - rj[5];
    e[9] = ri[6] - rj[6];
    e[10] = rj[8];
}

struct GnnArgs {
    Graph G;
    const float* W;      // [NPARAM]
    const bf16x8* Wp;    // [NPACK] packed node-MLP weights (inference)
    float* out;          // [n][NO] (forward) or nullptr
    // training
    const float* tgt;    // [n][NO] targets (normalised)
    const float* tw;     // [n] loss weight (1 real, 0 virtual)
    float* gW;           // [NPARAM] gradient accumulator
    float* loss;         // [1] accumulator of sum w (y - t)^2
    float inv_count;     // gradient scale: 1 / (number of weighted terms)
    float inv_out[NO];   // targets are physical corrections; y predicts target * inv_out
    float cw[NO];        // loss weight of each output
};

// Forward (TRAIN = false) or forward + backward (TRAIN = true) over 32
// nodes per workgroup (one wavefront).
// LDS tiles of one wavefront.  Inference reuses two regions: X0 then H4 in
// the first, HS then H3 in the second (each dead before the other is
// written); training keeps all of them for the backward pass.
constexpr int TILE_D = 32 * (D + 1);
constexpr int lds_floats(bool train)
{
    return train ? 32 * (FI + 1) + 3 * TILE_D + 32 * (FE + 1) + 32 * (HE + 1) : 2 * TILE_D;
}

template <bool TRAIN>
__global__ __launch_bounds__(64) void k_gnn(GnnArgs A)
{
    __shared__ float lds[lds_floats(TRAIN)];
    typedef float RowFI[FI + 1];
    typedef float RowD[D + 1];
    typedef float RowFE[FE + 1];
    typedef float RowHE[HE + 1];
    RowFI* X0 = reinterpret_cast<RowFI*>(lds);
    RowD* H4 = TRAIN ? reinterpret_cast<RowD*>(lds + 32 * (FI + 1)) : reinterpret_cast<RowD*>(lds);
    RowHE* HS = TRAIN ? reinterpret_cast<RowHE*>(lds + 32 * (FI + 1) + TILE_D) : reinterpret_cast<RowHE*>(lds + TILE_D);
    RowD* H3 = TRAIN ? reinterpret_cast<RowD*>(lds + 32 * (FI + 1) + TILE_D + 32 * (HE + 1))
                     : reinterpret_cast<RowD*>(lds + TILE_D);
    RowD* T = reinterpret_cast<RowD*>(lds + 32 * (FI + 1) + 2 * TILE_D + 32 * (HE + 1));
    RowFE* E = reinterpret_cast<RowFE*>(lds + 32 * (FI + 1) + 3 * TILE_D + 32 * (HE + 1));
    const int l = lane_id(), p = l & 31, h = l >> 5;
    const int base = blockIdx.x * 32;
    const int i = base + p;
    const bool valid = i < A.G.n;
    const float* W = A.W;
    // acc = A x (n-tile nt of the weight matrix at o, ld columns): packed
    // operands for inference, FP32 weights rounded in mma() for training
    auto mmw = [&](auto&& a, int o, int ld, int pbase, int nt, int K) {
        if constexpr (!TRAIN) return mma_p(a, A.Wp + pbase + nt * (K / 16) * 64, K, zero16());
        else return mma(a, [&](int k, int c) { return W[o + k * ld + nt * 32 + c]; }, K, zero16());
    };
    if (h == 0) {
#pragma unroll
        for (int f = 0; f < FN; ++f) {
            float v = valid ? A.G.nf[(size_t)i * FN + f] : 0.0f;
            if (f >= FR && f < FR + 3) v *= A.inv_out[f - FR];   // physical units -> target units
            X0[p][f] = v;
        }
    }
    const int deg = valid ? A.G.nd[i] : 0;
    const int maxdeg = wave_max(deg);
    float ri[RAW];
#pragma unroll
    for (int f = 0; f < RAW; ++f) ri[f] = valid ? A.G.raw[(size_t)i * RAW + f] : 1.0f;
    // W1 as the B operand: lane holds column p, rows h*8 .. h*8+7
    bf16x8 w1;
#pragma unroll
    for (int q = 0; q < 8; ++q) w1[q] = (__bf16)W[O_W1 + (h * 8 + q) * HE + p];

    // neighbour index of slot s (slot-major table), -1 past the degree
    auto nbr = [&](int s) { return (s < deg) ? A.G.nb[(size_t)s * A.G.n + i] : -1; };
    // The two half-waves build different features of the same edge: h = 0
    // (features 0-7) reads raw 0-4 of the neighbour (position and three field
    // values), h = 1 (features 8-15) raw 5, 6 and 8.  Each half loads one
    // aligned float4 (raw 0-3, or 4-7) and one float (raw 4, or 8); the
    // values the other half's features would read stay zero.
    auto load_raw = [&](int j, float* rj) {
#pragma unroll
        for (int f = 0; f < RAW; ++f) rj[f] = 0.0f;
        if (j < 0) return;
        const float* row = A.G.raw + (size_t)j * RAW;
        const float4 q = reinterpret_cast<const float4*>(row)[h];
        const float w = row[4 + 4 * h];
        if (h == 0) {
            rj[0] = q.x; rj[1] = q.y; rj[2] = q.z; rj[3] = q.w; rj[4] = w;
        } else {
            rj[4] = q.x; rj[5] = q.y; rj[6] = q.z; rj[7] = q.w; rj[8] = w;
        }
    };
    auto edge_from = [&](int j, const float* rj) {
        float e[FE];
        edge_features(ri, rj, j >= 0, e);
        bf16x8 av;
#pragma unroll
        for (int q = 0; q < 8; ++q) av[q] = (__bf16)(h ? e[8 + q] : e[q]);
        return av;
    };
    auto edge_operand = [&](int j) {
        float rj[RAW];
        load_raw(j, rj);
        return edge_from(j, rj);
    };

    // ---- edge layer and sum over neighbours.  Two slots per pass: half-wave
    // h builds all 16 features of slot s + h of its node, then lanes r and
    // r + 32 exchange the feature half the other slot's MFMA operand needs
    // (v_permlane32_swap), so each edge's features are computed once.  The
    // sum takes the slots in order, as one slot per pass would.
    f32x16 hsum = zero16();
    int jn = nbr(h);
#pragma unroll 1
    for (int s = 0; s < maxdeg; s += 2) {
        const int j = jn;
        jn = nbr(s + 2 + h);   // the next pair's index in flight during this one
        float rj[RAW];
#pragma unroll
        for (int f = 0; f < RAW; ++f) rj[f] = 0.0f;
        if (j >= 0) {
            const float* row = A.G.raw + (size_t)j * RAW;
            const float4 q0 = reinterpret_cast<const float4*>(row)[0];
            const float4 q1 = reinterpret_cast<const float4*>(row)[1];
            rj[0] = q0.x; rj[1] = q0.y; rj[2] = q0.z; rj[3] = q0.w;
            rj[4] = q1.x; rj[5] = q1.y; rj[6] = q1.z; rj[7] = q1.w;
            rj[8] = row[8];
        }
        float e[FE];
        edge_features(ri, rj, j >= 0, e);
        bf16x8 flo, fhi;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            flo[q] = (__bf16)e[q];
            fhi[q] = (__bf16)e[8 + q];
        }
        const u32x4 lo = __builtin_bit_cast(u32x4, flo), hi = __builtin_bit_cast(u32x4, fhi);
        u32x4 o1, o2;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            // lane r offers features 8-15 of slot s, lane r + 32 features 0-7
            // of slot s + 1; each takes the other's (first result: the lower
            // lanes' values in both halves, second: the upper lanes')
            const unsigned x = h ? lo[q] : hi[q];
            const auto sw = __builtin_amdgcn_permlane32_swap(x, x, false, false);
            const unsigned y = h ? sw[0] : sw[1];
            o1[q] = h ? y : lo[q];   // slot s: features 0-7 | 8-15
            o2[q] = h ? hi[q] : y;   // slot s + 1
        }
        f32x16 z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, o1), w1, zero16(), 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 16; ++v) hsum[v] += fmaxf(z[v], 0.0f);
        if (s + 1 < maxdeg) {
            z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, o2), w1, zero16(), 0, 0, 0);
#pragma unroll
            for (int v = 0; v < 16; ++v) hsum[v] += fmaxf(z[v], 0.0f);
        }
    }
    tile_each(hsum, [&](int row, int col, float& v) { HS[row][col] = v * EDGE_NORM; });
    __syncthreads();
    // ---- message
    {
        f32x16 m = mmw([&](int r, int k) { return HS[r][k]; }, O_W2, HE, P_W2, 0, HE);
        tile_each(m, [&](int row, int col, float& v) { X0[row][FN + col] = v; });
    }
    __syncthreads();
    // ---- node MLP
#pragma unroll 1
    for (int nt = 0; nt < 2; ++nt) {
        f32x16 z = mmw([&](int r, int k) { return X0[r][k]; }, O_W3, D, P_W3, nt, FI);
        tile_each(z, [&](int row, int col, float& v) {
            H3[row][nt * 32 + col] = fmaxf(v + W[O_B3 + nt * 32 + col], 0.0f);
        });
    }
    __syncthreads();
#pragma unroll 1
    for (int nt = 0; nt < 2; ++nt) {
        f32x16 z = mmw([&](int r, int k) { return H3[r][k]; }, O_W4, D, P_W4, nt, D);
        tile_each(z, [&](int row, int col, float& v) {
            H4[row][nt * 32 + col] = fmaxf(v + W[O_B4 + nt * 32 + col], 0.0f);
        });
    }
    __syncthreads();
    f32x16 y = mmw([&](int r, int k) { return H4[r][k]; }, O_W5, NOP, P_W5, 0, D);
    if constexpr (!TRAIN) {
        tile_each(y, [&](int row, int col, float& v) {
            if (col < NO && base + row < A.G.n) A.out[(size_t)(base + row) * NO + col] = v + W[O_B5 + col];
        });
        return;
    } else {
        // ---- loss and dL/dy
        float lsum = 0.0f;
        tile_each(y, [&](int row, int col, float& v) {
            const int node = base + row;
            float d = 0.0f;
            if (col < NO && node < A.G.n) {
                const float wgt = A.tw[node];
                const float err = v + W[O_B5 + col] - A.tgt[(size_t)node * NO + col] * A.inv_out[col];
                lsum += A.cw[col] * wgt * err * err;
                d = 2.0f * A.cw[col] * wgt * err * A.inv_count;
            }
            T[row][col] = d;
        });
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o);
        if (l == 0) unsafeAtomicAdd(A.loss, lsum);
        __syncthreads();
        float* gW = A.gW;
        // dW5 = H4^T dy, db5 = sum dy
#pragma unroll 1
        for (int mt = 0; mt < 2; ++mt) {
            f32x16 g = mma([&](int r, int k) { return H4[k][mt * 32 + r]; }, [&](int k, int c) { return T[k][c]; },
                           32, zero16());
            tile_each(g, [&](int row, int col, float& v) {
                if (col < NO) unsafeAtomicAdd(&gW[O_W5 + (mt * 32 + row) * NOP + col], v);
            });
        }
        if (l < NO) {
            float sb = 0.0f;
            for (int k = 0; k < 32; ++k) sb += T[k][l];
            unsafeAtomicAdd(&gW[O_B5 + l], sb);
        }
        // dz4 = (dy W5^T) * [H4 > 0], into H4
        f32x16 g4[2];
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
            g4[nt] = mma([&](int r, int k) { return T[r][k]; },
                         [&](int k, int c) { return W[O_W5 + (nt * 32 + c) * NOP + k]; }, 32, zero16());
        __syncthreads();
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
            tile_each(g4[nt], [&](int row, int col, float& v) {
                float& hv = H4[row][nt * 32 + col];
                hv = hv > 0.0f ? v : 0.0f;
            });
        __syncthreads();
        // dW4 = H3^T dz4, db4
#pragma unroll 1
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll 1
            for (int nt = 0; nt < 2; ++nt) {
                f32x16 g = mma([&](int r, int k) { return H3[k][mt * 32 + r]; },
                               [&](int k, int c) { return H4[k][nt * 32 + c]; }, 32, zero16());
                tile_each(g, [&](int row, int col, float& v) {
                    unsafeAtomicAdd(&gW[O_W4 + (mt * 32 + row) * D + nt * 32 + col], v);
                });
            }
        {
            float sb = 0.0f;
            for (int k = 0; k < 32; ++k) sb += H4[k][l];
            unsafeAtomicAdd(&gW[O_B4 + l], sb);
        }
        // dz3 = (dz4 W4^T) * [H3 > 0], into T
#pragma unroll 1
        for (int nt = 0; nt < 2; ++nt) {
            f32x16 g = mma([&](int r, int k) { return H4[r][k]; },
                           [&](int k, int c) { return W[O_W4 + (nt * 32 + c) * D + k]; }, D, zero16());
            tile_each(g, [&](int row, int col, float& v) {
                T[row][nt * 32 + col] = H3[row][nt * 32 + col] > 0.0f ? v : 0.0f;
            });
        }
        __syncthreads();
        // dW3 = X0^T dz3 (rows 0..FI-1), db3
#pragma unroll 1
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll 1
            for (int nt = 0; nt < 2; ++nt) {
                f32x16 g = mma([&](int r, int k) { const int f = mt * 32 + r; return f < FI ? X0[k][f] : 0.0f; },
                               [&](int k, int c) { return T[k][nt * 32 + c]; }, 32, zero16());
                tile_each(g, [&](int row, int col, float& v) {
                    const int f = mt * 32 + row;
                    if (f < FI) unsafeAtomicAdd(&gW[O_W3 + f * D + nt * 32 + col], v);
                });
            }
        {
            float sb = 0.0f;
            for (int k = 0; k < 32; ++k) sb += T[k][l];
            unsafeAtomicAdd(&gW[O_B3 + l], sb);
        }
        // dm = dz3 W3[FN:, :]^T, into H3 (32 x HE)
        {
            f32x16 g = mma([&](int r, int k) { return T[r][k]; },
                           [&](int k, int c) { return W[O_W3 + (FN + c) * D + k]; }, D, zero16());
            __syncthreads();
            tile_each(g, [&](int row, int col, float& v) { H3[row][col] = v; });
        }
        __syncthreads();
        // dW2 = HS^T dm
        {
            f32x16 g = mma([&](int r, int k) { return HS[k][r]; }, [&](int k, int c) { return H3[k][c]; }, 32,
                           zero16());
            tile_each(g, [&](int row, int col, float& v) { unsafeAtomicAdd(&gW[O_W2 + row * HE + col], v); });
        }
        // d hsum = (dm W2^T) / 16, in the layout of hsum
        f32x16 dh = mma([&](int r, int k) { return H3[r][k]; }, [&](int k, int c) { return W[O_W2 + c * HE + k]; },
                        HE, zero16());
#pragma unroll
        for (int v = 0; v < 16; ++v) dh[v] *= EDGE_NORM;
        // edge layer: dW1 = sum_s e_s^T (dh * [z_s > 0])
        f32x16 g1 = zero16();
#pragma unroll 1
        for (int s = 0; s < maxdeg; ++s) {
            const bf16x8 av = edge_operand(nbr(s));
#pragma unroll
            for (int q = 0; q < 8; ++q) E[p][h * 8 + q] = (float)av[q];
            f32x16 z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, w1, zero16(), 0, 0, 0);
            int v = 0;
            tile_each(z, [&](int row, int col, float& zv) {
                T[row][col] = zv > 0.0f ? dh[v] : 0.0f;
                ++v;
            });
            __syncthreads();
            g1 = mma([&](int r, int k) { return r < FE ? E[k][r] : 0.0f; }, [&](int k, int c) { return T[k][c]; }, 32,
                     g1);
            __syncthreads();
        }
        tile_each(g1, [&](int row, int col, float& v) {
            if (row < FE) unsafeAtomicAdd(&gW[O_W1 + row * HE + col], v);
        });
    }
}

// Adam (Kingma & Ba 2015) on the FP32 master weights.
// Packed weights (see P_W2): one entry of 8 BF16 values per thread.
__global__ void k_gnn_pack(const float* W, bf16x8* Wp)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NPACK) return;
    int o, ld, base, steps;
    if (e < P_W3) { o = O_W2; ld = HE; base = P_W2; steps = 2; }
    else if (e < P_W4) { o = O_W3; ld = D; base = P_W3; steps = 4; }
    else if (e < P_W5) { o = O_W4; ld = D; base = P_W4; steps = 4; }
    else { o = O_W5; ld = NOP; base = P_W5; steps = 4; }
    const int q = e - base, l = q & 63, ks = (q >> 6) % steps, nt = (q >> 6) / steps;
    const int r = l & 31, h = l >> 5;
    bf16x8 v;
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = (__bf16)W[o + (ks * 16 + h * 8 + i) * ld + nt * 32 + r];
    Wp[e] = v;
}

__global__ void k_adam(float* W, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2, float clip)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = fminf(fmaxf(g[i], -clip), clip);
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    W[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// ------------------------------------------------------------------------
// Graph and features of one phase, in cell order: node = off + t for the
// particle a = order[t].  Neighbours: same-phase pairs with r < 2h (the
// exact predicate of the step), at most DEG (nd counts the kept ones).
// ------------------------------------------------------------------------
struct GatherOut {
    float* nf;
    float* raw;
    int* nb;
    int* nd;
    int n_nodes;       // nodes of both phases: the slot stride of nb
    float* tw;         // [n]
    int* node_of_id;   // [n of the phase] (phase-local ids)
    int* overflow;     // [1]
};

// NLM.m = nullptr: candidates are the rows of the particle's current cell (a
// cell list of the current positions on grid g).  Otherwise the cell list
// and record order are those of the FP64 step's neighbour lists (grid g is
// their grid): the candidates are the particle's listed same-phase
// candidates, tested with the current positions; a list that overflowed
// walks its rows.
__device__ __forceinline__ void gnn_gather_body(int blk, DPhase P, Grid g, KParams K, double thr, int uni, int off,
                                                int phase, Scales S, GatherOut O, Masks NLM, const int* nl_cell)
{
    const int t = blk * blockDim.x + threadIdx.x;
    if (t >= P.n) return;
    const int a = P.order[t];
    const int node = off + t;
    const double x = P.x[a], y = P.y[a], h = P.hsml[a];
    const double rho0 = phase == 0 ? K.w_rho0 : K.s_rho0;
    const bool virt = P.itype[a] < 0;
    const double kdt = S.kdt;
    float nf[FN];
    nf[0] = 1.0f;
    nf[1] = (float)phase;
    nf[2] = virt ? 1.0f : 0.0f;
    nf[3] = (float)(P.ax[a] * S.inv_a);
    nf[4] = (float)(P.ay[a] * S.inv_a);
    nf[5] = (float)(P.vx[a] * S.inv_v);
    nf[6] = (float)(P.vy[a] * S.inv_v);
    nf[7] = (float)((P.rho[a] - rho0) / rho0 * S.inv_r);
    nf[8] = (float)(P.drho[a] * kdt / rho0 * S.inv_r);
    nf[9] = (float)P.vof[a];
    if (phase == 1) {
        nf[10] = (float)(P.sxx[a] * S.inv_s);
        nf[11] = (float)(P.sxy[a] * S.inv_s);
        nf[12] = (float)(P.syy[a] * S.inv_s);
        nf[13] = (float)(P.p[a] * S.inv_s);
    } else {
        nf[10] = nf[11] = nf[12] = nf[13] = 0.0f;
    }
    nf[14] = nf[15] = 0.0f;
    // the rates the state holds (of the step before): k_gnn_delta turns them
    // into the rate-change features once the coarse step has evaluated the
    // rates of this state
    nf[FR] = (float)P.ax[a];
    nf[FR + 1] = (float)P.ay[a];
    nf[FR + 2] = (float)P.drho[a];
#pragma unroll
    for (int f = FR + 3; f < FN; ++f) nf[f] = 0.0f;
    // rows of 16 and 12 floats, written as 16-byte vectors
    float4* nf4 = reinterpret_cast<float4*>(O.nf + (size_t)node * FN);
#pragma unroll
    for (int q = 0; q < FN / 4; ++q) nf4[q] = make_float4(nf[4 * q], nf[4 * q + 1], nf[4 * q + 2], nf[4 * q + 3]);
    float4* rw4 = reinterpret_cast<float4*>(O.raw + (size_t)node * RAW);
    rw4[0] = make_float4((float)x, (float)y, nf[5], nf[6]);
    rw4[1] = make_float4(nf[7], nf[3], nf[4], (float)h);
    rw4[2] = make_float4(nf[2], 0.0f, 0.0f, 0.0f);
    O.tw[node] = virt ? 0.0f : 1.0f;
    O.node_of_id[P.id[a]] = node;
    // neighbours (cell order of the same phase), in ascending k
    int cnt = 0;
    int* nb = O.nb + node;   // slot s at nb[s * n_nodes]
    const size_t ns = (size_t)O.n_nodes;
    // candidates in groups of four: the index loads, then the position loads
    // of all four, then the predicates in candidate order
    auto group = [&](int k0, int k1, int k2, int k3) {
        const int kq[4] = {k0, k1, k2, k3};
        int bq[4];
        double xq[4], yq[4], hq[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) bq[q] = kq[q] >= 0 ? P.order[kq[q]] : -1;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int b = bq[q] >= 0 ? bq[q] : a;
            xq[q] = P.x[b];
            yq[q] = P.y[b];
            hq[q] = uni ? h : P.hsml[b];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = kq[q];
            if (bq[q] < 0 || k == t) continue;
            const double dx = x - xq[q], dy = y - yq[q];
            double dr;
            {
#pragma clang fp contract(off)
                dr = dx * dx + dy * dy;
            }
            if (!pred(uni != 0, thr, dr, h, hq[q], K)) continue;
            if (cnt < DEG) nb[(size_t)cnt * ns] = off + k;
            ++cnt;
        }
    };
    int cx, cy;
    if (nl_cell) {
        const int key = nl_cell[t];
        cy = key / g.ncx;
        cx = key - cy * g.ncx;
    } else {
        clamped_cell(g, x, y, cx, cy);
    }
    const Rows R = block_rows_at(g, P.cell_start, cx, cy);
    const CatMask vm = NLM.m ? mask_of(NLM, P.n, t) : CatMask{0, 0, 7};
    if (!vm.ovf) {
        // listed candidates: the mask's bits over the rows, ascending
        const int s1 = R.k1[0] - R.k0[0], s2 = s1 + R.k1[1] - R.k0[1];
        const int e0 = R.k0[0], e1 = R.k0[1] - s1, e2 = R.k0[2] - s2;
        uint64_t lo = vm.lo, hi = vm.hi;
#pragma unroll 1
        while (lo | hi) {
            int kq[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const bool in_lo = lo != 0;
                const uint64_t m = in_lo ? lo : hi;
                const int c = m ? __builtin_ctzll(m) + (in_lo ? 0 : 64) : -1;
                const uint64_t rest = m & (m - 1);
                lo = in_lo ? rest : lo;
                hi = in_lo ? hi : rest;
                kq[q] = c < 0 ? -1 : c + (c < s1 ? e0 : (c < s2 ? e1 : e2));
            }
            group(kq[0], kq[1], kq[2], kq[3]);
        }
    } else {
#pragma unroll 1
        for (int r = 0; r < 3; ++r) {
            const int k0 = r == 0 ? R.k0[0] : (r == 1 ? R.k0[1] : R.k0[2]);
            const int k1 = r == 0 ? R.k1[0] : (r == 1 ? R.k1[1] : R.k1[2]);
#pragma unroll 1
            for (int kb = k0; kb < k1; kb += 4) {
                int kq[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) kq[q] = kb + q < k1 ? kb + q : -1;
                group(kq[0], kq[1], kq[2], kq[3]);
            }
        }
    }
    if (cnt > DEG) atomicOr(O.overflow, 1);
    O.nd[node] = min(cnt, DEG);   // (slots past the degree are never read)
}

__global__ void k_gnn_gather(DPhase P, Grid g, KParams K, double thr, int uni, int off, int phase, Scales S,
                             GatherOut O, Masks NLM, const int* nl_cell)
{
    gnn_gather_body(blockIdx.x, P, g, K, thr, uni, off, phase, S, O, NLM, nl_cell);
}

// Rate-change features: after the coarse step P holds the rates evaluated at
// the state the advance starts from, and features FR..FR+2 hold the rates of
// the evaluation `gap` steps of dt before it.  Over K steps of dt a rate r
// changing at dr/dt moves the fine result from the coarse one by about
// K (K - 1) / 2 dt^2 dr/dt: with dr/dt ~ (r_new - r_old) / (gap dt) the
// features become coef (r_new - r_old), coef = K (K - 1) / 2 dt / gap, in the
// units of the correction (velocity, density).
__device__ __forceinline__ void gnn_delta_body(int blk, DPhase P, const int* node_of_id, float* nf, double coef)
{
    const int a = blk * blockDim.x + threadIdx.x;
    if (a >= P.n) return;
    float* f = nf + (size_t)node_of_id[P.id[a]] * FN + FR;
    f[0] = (float)((P.ax[a] - (double)f[0]) * coef);
    f[1] = (float)((P.ay[a] - (double)f[1]) * coef);
    f[2] = (float)((P.drho[a] - (double)f[2]) * coef);
}

__global__ void k_gnn_delta(DPhase P, const int* node_of_id, float* nf, double coef)
{
    gnn_delta_body(blockIdx.x, P, node_of_id, nf, coef);
}

// Targets: fine state (K plain steps, current) minus coarse state (one step
// of K dt, snapshot C), by particle id, in physical units.
__global__ void k_gnn_target(DPhase P, const int* node_of_id, DPhase C, const int* slot_of_id_c, float* tgt)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= P.n) return;
    const int id = P.id[a];
    const int node = node_of_id[id];
    const int b = slot_of_id_c[id];
    float* tg = tgt + (size_t)node * NO;
    tg[0] = (float)(P.vx[a] - C.vx[b]);
    tg[1] = (float)(P.vy[a] - C.vy[b]);
    tg[2] = (float)(P.rho[a] - C.rho[b]);
    tg[3] = 0.0f;
}

// The network's correction of the coarse step (real particles).
__device__ __forceinline__ void gnn_correct_body(int blk, DPhase P, KParams K, const int* node_of_id, const float* y,
                                                 Scales S, int phase)
{
    const int a = blk * blockDim.x + threadIdx.x;
    if (a >= P.n) return;
    if (P.itype[a] <= 0) return;
    const int node = node_of_id[P.id[a]];
    const float* yy = y + (size_t)node * NO;
    const double cvx = (double)yy[0] * S.out[0] * S.gain[0];
    const double cvy = (double)yy[1] * S.out[1] * S.gain[1];
    const double crh = (double)yy[2] * S.out[2] * S.gain[2];
    P.vx[a] = P.vx[a] + cvx;
    P.vy[a] = P.vy[a] + cvy;
    P.x[a] = P.x[a] + 0.5 * (double)S.kdt * cvx;
    P.y[a] = P.y[a] + 0.5 * (double)S.kdt * cvy;
    P.rho[a] = P.rho[a] + crh;
    if (phase == 1) P.vof[a] = P.rho[a] / K.s_rho0;
}

__global__ void k_gnn_correct(DPhase P, KParams K, const int* node_of_id, const float* y, Scales S, int phase)
{
    gnn_correct_body(blockIdx.x, P, K, node_of_id, y, S, phase);
}

// out[0]: max (as float bits: non-negative floats order as their bits);
// sq[0] += sum of squares, sq[1] += count.  One atomic per block.
__global__ __launch_bounds__(256) void k_rho_err(DPhase A, DPhase B, const int* slot_of_id_b, unsigned int* out,
                                                 double* sq)
{
    __shared__ float smax[4];
    __shared__ double ssum[4], scnt[4];
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    float e = 0.0f;
    double e2 = 0.0, c = 0.0;
    if (a < A.n && A.itype[a] > 0) {
        const int b = slot_of_id_b[A.id[a]];
        const double ed = fabs(A.rho[a] - B.rho[b]) / fabs(B.rho[b]);
        e = (float)ed;
        e2 = ed * ed;
        c = 1.0;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        e = fmaxf(e, __shfl_xor(e, o));
        e2 += __shfl_xor(e2, o);
        c += __shfl_xor(c, o);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        smax[w] = e;
        ssum[w] = e2;
        scnt[w] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 1; i < nw; ++i) {
            e = fmaxf(e, smax[i]);
            e2 += ssum[i];
            c += scnt[i];
        }
        atomicMax(out, __float_as_uint(e));
        unsafeAtomicAdd(&sq[0], e2);
        unsafeAtomicAdd(&sq[1], c);
    }
}

// Flags a state a learned advance has broken: a non-finite or non-positive
// density, a non-finite velocity, or a particle outside the cell grid.
__device__ __forceinline__ void sanity_body(int blk, DPhase P, Grid g, int* flag)
{
    const int a = blk * blockDim.x + threadIdx.x;
    if (a >= P.n) return;
    const double x = P.x[a], y = P.y[a], r = P.rho[a];
    const bool ok = x >= g.x0 && x <= g.x1 && y >= g.y0 && y <= g.y1 && r > 0.0 && r < 1.0e300 &&
                    P.vx[a] == P.vx[a] && P.vy[a] == P.vy[a];
    if (!ok) atomicOr(flag, 1);
}

__global__ void k_sanity(DPhase P, Grid g, int* flag)
{
    sanity_body(blockIdx.x, P, g, flag);
}

// Least-squares gain of each output channel over one sample (real nodes):
// acc[o] += sum w y_o t_o / out_o, acc[NO + o] += sum w y_o^2.  The gain
// sum(y t) / sum(y^2) scales the applied correction so that it removes
// as much of the target as the network's prediction explains.
__global__ __launch_bounds__(256) void k_gnn_fit(const float* y, const float* tgt, const float* tw, int n,
                                                 Scales S, double* acc)
{
    __shared__ double part[4][2 * NO];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[2 * NO];
#pragma unroll
    for (int o = 0; o < NO; ++o) {
        const bool ok = i < n && tw[i] > 0.0f;
        const double yo = ok ? (double)y[(size_t)i * NO + o] : 0.0;
        const double to = ok ? (double)tgt[(size_t)i * NO + o] / (double)S.out[o] : 0.0;
        v[o] = yo * to;
        v[NO + o] = yo * yo;
    }
#pragma unroll
    for (int q = 0; q < 2 * NO; ++q)
        for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_down(v[q], off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
        for (int q = 0; q < 2 * NO; ++q) part[w][q] = v[q];
    __syncthreads();
    if (threadIdx.x < 2 * NO) {
        const double sum = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x];
        atomicAdd(&acc[threadIdx.x], sum);
    }
}

// The state of a phase copied back from a snapshot (A.s* -> A.d*) when the
// sanity flag is set: a learned advance that broke the state is undone on
// the device, before any later kernel reads it.
__device__ __forceinline__ void restore_if_body(int blk, const int* flag, PermArgs A, int n)
{
    if (*flag == 0) return;
    const int i = blk * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // (constant indices into the argument arrays: a runtime-indexed loop
    // would copy them to scratch memory)
#pragma unroll
    for (int q = 0; q < PERM_MAXD; ++q)
        if (q < A.nd) A.dd[q][i] = A.sd[q][i];
#pragma unroll
    for (int q = 0; q < PERM_MAXI; ++q)
        if (q < A.ni) A.di[q][i] = A.si[q][i];
}

__global__ void k_restore_if(const int* flag, PermArgs A, int n)
{
    restore_if_body(blockIdx.x, flag, A, n);
}

__global__ void k_slot_of_id(DPhase P, int* slot_of_id)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < P.n) slot_of_id[P.id[a]] = a;
}

}  // namespace gnn
