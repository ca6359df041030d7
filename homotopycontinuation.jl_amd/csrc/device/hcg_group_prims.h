// Cross-lane primitives of the lane-group tracker: one path per group of GL consecutive lanes
// of a wave, lane r of the group holding entry r of every vector and row r of the Jacobian.
// GL = 16 (a DPP row; systems of up to 16 variables) unless the including file defines
// GL 32 (half a wave; up to 24 variables, the library's largest system size).
//
// Reductions that the reference writes as sequential loops over the entries (sums, maxima
// with its NaN conventions, first-index searches) are done as the same sequential loop over
// broadcasts of each lane's term, so every lane obtains exactly the sequential result.
#pragma once
#include "hcg_arith.h"

#ifndef GL
#define GL 16
#endif
static_assert(GL == 16 || GL == 32, "lane groups of 16 or 32 lanes");
#define GL_MASK ((GL == 32) ? 0xFFFFFFFFull : 0xFFFFull)

__device__ __forceinline__ int g_lane() { return (int)(__lane_id() & (unsigned)(GL - 1)); }
__device__ __forceinline__ int g_base() { return (int)(__lane_id() & ~(unsigned)(GL - 1)); }

// broadcast of lane J (compile time) of the group: DPP row_newbcast within a 16-lane row;
// across the two rows of a 32-lane group, a permute through the LDS crossbar
template <int J> __device__ __forceinline__ int gb_i(int v) {
    if constexpr (GL == 16) {
        return __builtin_amdgcn_update_dpp(0, v, 0x150 + J, 0xF, 0xF, false);
    } else {
        return __shfl(v, g_base() + J, 64);
    }
}
template <int J> __device__ __forceinline__ double gb_d(double v) {
    int2 u = __builtin_bit_cast(int2, v);
    u.x = gb_i<J>(u.x);
    u.y = gb_i<J>(u.y);
    return __builtin_bit_cast(double, u);
}
template <int J> __device__ __forceinline__ cplx gb_c(cplx v) { return mkc(gb_d<J>(v.re), gb_d<J>(v.im)); }

// broadcast of lane j (run time, the same for the whole group): LDS crossbar
__device__ __forceinline__ int gs_i(int v, int j) { return __shfl(v, g_base() + j, 64); }
__device__ __forceinline__ double gs_d(double v, int j) { return __shfl(v, g_base() + j, 64); }
__device__ __forceinline__ cplx gs_c(cplx v, int j) { return mkc(gs_d(v.re, j), gs_d(v.im, j)); }

// any / all over the group's first n lanes
__device__ __forceinline__ bool g_any(bool p, int n) {
    const unsigned long long m = __ballot(p && g_lane() < n ? 1 : 0);
    return ((m >> g_base()) & GL_MASK) != 0ull;
}
__device__ __forceinline__ bool g_all(bool p, int n) { return !g_any(!p, n); }

// Sequential folds over lanes 0..n-1 (the reference's loop order); NN: compile-time bound.
// F(acc, term) with acc starting at lane 0's term.
template <int NN, class Fn> __device__ __forceinline__ double g_fold(double v, int n, Fn f) {
    double acc = gb_d<0>(v);
#define HCG_GFOLD_STEP(J) if (J < NN && J < n) acc = f(acc, gb_d<(J < GL ? J : 0)>(v));
    HCG_GFOLD_STEP(1) HCG_GFOLD_STEP(2) HCG_GFOLD_STEP(3) HCG_GFOLD_STEP(4) HCG_GFOLD_STEP(5)
    HCG_GFOLD_STEP(6) HCG_GFOLD_STEP(7) HCG_GFOLD_STEP(8) HCG_GFOLD_STEP(9) HCG_GFOLD_STEP(10)
    HCG_GFOLD_STEP(11) HCG_GFOLD_STEP(12) HCG_GFOLD_STEP(13) HCG_GFOLD_STEP(14) HCG_GFOLD_STEP(15)
    HCG_GFOLD_STEP(16) HCG_GFOLD_STEP(17) HCG_GFOLD_STEP(18) HCG_GFOLD_STEP(19) HCG_GFOLD_STEP(20)
    HCG_GFOLD_STEP(21) HCG_GFOLD_STEP(22) HCG_GFOLD_STEP(23) HCG_GFOLD_STEP(24) HCG_GFOLD_STEP(25)
    HCG_GFOLD_STEP(26) HCG_GFOLD_STEP(27) HCG_GFOLD_STEP(28) HCG_GFOLD_STEP(29) HCG_GFOLD_STEP(30)
    HCG_GFOLD_STEP(31)
#undef HCG_GFOLD_STEP
    return acc;
}
// sum starting from `init` over lanes 0..n-1 in order
template <int NN> __device__ __forceinline__ double g_sum(double v, int n, double init) {
    double acc = init;
#define HCG_GSUM_STEP(J) if (J < NN && J < n) acc += gb_d<(J < GL ? J : 0)>(v);
    HCG_GSUM_STEP(0) HCG_GSUM_STEP(1) HCG_GSUM_STEP(2) HCG_GSUM_STEP(3) HCG_GSUM_STEP(4) HCG_GSUM_STEP(5)
    HCG_GSUM_STEP(6) HCG_GSUM_STEP(7) HCG_GSUM_STEP(8) HCG_GSUM_STEP(9) HCG_GSUM_STEP(10)
    HCG_GSUM_STEP(11) HCG_GSUM_STEP(12) HCG_GSUM_STEP(13) HCG_GSUM_STEP(14) HCG_GSUM_STEP(15)
    HCG_GSUM_STEP(16) HCG_GSUM_STEP(17) HCG_GSUM_STEP(18) HCG_GSUM_STEP(19) HCG_GSUM_STEP(20)
    HCG_GSUM_STEP(21) HCG_GSUM_STEP(22) HCG_GSUM_STEP(23) HCG_GSUM_STEP(24) HCG_GSUM_STEP(25)
    HCG_GSUM_STEP(26) HCG_GSUM_STEP(27) HCG_GSUM_STEP(28) HCG_GSUM_STEP(29) HCG_GSUM_STEP(30)
    HCG_GSUM_STEP(31)
#undef HCG_GSUM_STEP
    return acc;
}
// index of the first lane (0..n-1) whose predicate holds, or n
__device__ __forceinline__ int g_first(bool p, int n) {
    const unsigned long long m = (__ballot(p && g_lane() < n ? 1 : 0) >> g_base()) & GL_MASK;
    return m ? __builtin_ctzll(m) : n;
}

// compile-time loop over 0..NN-1 (unrolled; the body gets an integral constant)
template <int I, int N> struct GFor {
    template <class Fn> __device__ __forceinline__ static void run(Fn&& f) {
        if constexpr (I < N) {
            f(std::integral_constant<int, I>());
            GFor<I + 1, N>::run(f);
        }
    }
};
#define G_FOR(NN, J, ...) GFor<0, NN>::run([&](auto J##_c) { constexpr int J = decltype(J##_c)::value; __VA_ARGS__ })
