// One launch per stage for a batch of independent cases (no counterpart in
// the reference, whose host runs one case per process).
//
// The host records the kernel launches of one step of every case, stores
// each case's arguments as an ArgTup in device memory, and launches the
// batched form of each kernel of the step once over all cases: a tile table maps each workgroup of the batched grid to (case,
// block of the case).  The batched kernel calls the same device body as the
// solo kernel with that block, so every particle sees exactly the arithmetic
// of a solo launch.
#pragma once

#include <utility>

// Kernel arguments of one case in the order of the kernel's parameters
// (a plain, trivially copyable nest: the host writes it as bytes).
template <class... T>
struct ArgTup;

template <>
struct ArgTup<> {
};

template <class H, class... T>
struct ArgTup<H, T...> {
    H h;
    ArgTup<T...> t;
    ArgTup() = default;
    template <class A0, class... A>
    __host__ __device__ explicit ArgTup(const A0& a0, const A&... a) : h(a0), t(a...)
    {
    }
};

template <size_t I, class H, class... T>
__host__ __device__ __forceinline__ const auto& tup_get(const ArgTup<H, T...>& a)
{
    if constexpr (I == 0) return a.h;
    else return tup_get<I - 1>(a.t);
}

template <class F, class Tup, size_t... I>
__device__ __forceinline__ void tup_call(int blk, const Tup& a, std::index_sequence<I...>)
{
    F{}(blk, tup_get<I>(a)...);
}

// Workgroup b of the batched grid (XCD-aware, so that each XCD works on a
// contiguous run of cases) runs block tiles[b].y of case tiles[b].x.
template <class F, class... T>
__global__ __launch_bounds__(256) void k_batched(const ArgTup<T...>* args, const int2* tiles)
{
    const int2 tl = tiles[xcd_block()];
    tup_call<F>(tl.y, args[tl.x], std::index_sequence_for<T...>{});
}

// The device bodies of the kernels a list-mode FP64 step launches.
template <int SKF, bool U>
struct BodyWaterX {
    template <class... A>
    __device__ __forceinline__ void operator()(int blk, const A&... a) const { water_x4<SKF, U>(blk, a...); }
};
template <int SKF, bool U>
struct BodySoilX {
    template <class... A>
    __device__ __forceinline__ void operator()(int blk, const A&... a) const { soil_x4<SKF, U>(blk, a...); }
};
template <int SKF, bool U>
struct BodyWaterA {
    template <class... A>
    __device__ __forceinline__ void operator()(int blk, const A&... a) const { water_a4<SKF, U>(blk, a...); }
};
template <int SKF, bool U>
struct BodyWaterB {
    template <class... A>
    __device__ __forceinline__ void operator()(int blk, const A&... a) const { water_b4<SKF, U>(blk, a...); }
};
template <int SKF, bool U, bool SOLO>
struct BodySoilAB {
    template <class... A>
    __device__ __forceinline__ void operator()(int blk, const A&... a) const { soil_ab4<SKF, U, SOLO>(blk, a...); }
};
struct BodyStepWater {
    template <class... A>
    __device__ __forceinline__ void operator()(int blk, const A&... a) const { step_water(blk, a...); }
};
struct BodyStepSoil {
    template <class... A>
    __device__ __forceinline__ void operator()(int blk, const A&... a) const { step_soil(blk, a...); }
};
struct BodyFinalWater {
    template <class... A>
    __device__ __forceinline__ void operator()(int blk, const A&... a) const { final_water(blk, a...); }
};
struct BodyFinalSoil {
    template <class... A>
    __device__ __forceinline__ void operator()(int blk, const A&... a) const { final_soil(blk, a...); }
};

// ... and of the kernels of its rebuild step (cell lists, reorder, packs,
// neighbour lists).
#define SPHG_BATCH_BODY(NAME, FN)                                                        \
    struct NAME {                                                                        \
        template <class... A>                                                            \
        __device__ __forceinline__ void operator()(int blk, const A&... a) const { FN(blk, a...); } \
    };
SPHG_BATCH_BODY(BodyFillInt, fill_int_body)
SPHG_BATCH_BODY(BodyKeepBits, keep_bits_body)
SPHG_BATCH_BODY(BodyIota, iota_body)
SPHG_BATCH_BODY(BodyPermute, permute_multi_body)
SPHG_BATCH_BODY(BodyScanTile, scan_tile_body)
SPHG_BATCH_BODY(BodyCellKey, cell_key_body)
SPHG_BATCH_BODY(BodyCellScatter, cell_scatter_t_body)
SPHG_BATCH_BODY(BodyCellSort, cell_sort_t_body)
SPHG_BATCH_BODY(BodyPackWater, pack_water_body)
SPHG_BATCH_BODY(BodyPackSoil, pack_soil_body)
SPHG_BATCH_BODY(BodyNlist, nlist_body)

// ... and of the mixed-precision stage kernels.
template <int SKF>
struct BodyWaterAMx {
    template <class... A>
    __device__ __forceinline__ void operator()(int blk, const A&... a) const { water_a_mx<SKF, 1>(blk, a...); }
};
template <int SKF>
struct BodyWaterXMx {
    template <class... A>
    __device__ __forceinline__ void operator()(int blk, const A&... a) const { water_a_mx<SKF, 2>(blk, a...); }
};
template <int SKF>
struct BodyWaterBMx {
    template <class... A>
    __device__ __forceinline__ void operator()(int blk, const A&... a) const { water_b_mx<SKF>(blk, a...); }
};
template <int SKF>
struct BodySoilABMx {
    template <class... A>
    __device__ __forceinline__ void operator()(int blk, const A&... a) const { soil_ab_mx<SKF, 1>(blk, a...); }
};
template <int SKF>
struct BodySoilXMx {
    template <class... A>
    __device__ __forceinline__ void operator()(int blk, const A&... a) const { soil_ab_mx<SKF, 2>(blk, a...); }
};
#undef SPHG_BATCH_BODY

// Two argument sets of one body in one launch (the two phases of a context):
// blocks [0, nb0) run the first set, the rest the second.
template <class F, class... T>
__global__ __launch_bounds__(256) void k_pair(ArgTup<T...> a0, ArgTup<T...> a1, int nb0)
{
    const int b = blockIdx.x;
    if (b < nb0) tup_call<F>(b, a0, std::index_sequence_for<T...>{});
    else tup_call<F>(b - nb0, a1, std::index_sequence_for<T...>{});
}

// The surrogate's per-phase kernels.
#define SPHG_PAIR_BODY(NAME, FN)                                                         \
    struct NAME {                                                                        \
        template <class... A>                                                            \
        __device__ __forceinline__ void operator()(int blk, const A&... a) const { FN(blk, a...); } \
    };
SPHG_PAIR_BODY(BodyGnnGather, gnn::gnn_gather_body)
SPHG_PAIR_BODY(BodyGnnDelta, gnn::gnn_delta_body)
SPHG_PAIR_BODY(BodyGnnCorrect, gnn::gnn_correct_body)
SPHG_PAIR_BODY(BodySanity, gnn::sanity_body)
SPHG_PAIR_BODY(BodyRestoreIf, gnn::restore_if_body)
#undef SPHG_PAIR_BODY
