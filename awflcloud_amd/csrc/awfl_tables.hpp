// awfl_tables.hpp -- compile-time views of the generated transform tables.
//
// Each accessor exposes a namespace-scope constexpr array of awfl_tables.h as
// a constexpr reference, so that kernels indexing it with compile-time
// indices (awfl::sfor) fold every entry into an instruction immediate and can
// drop zero entries with `if constexpr`.
#pragma once
#include <type_traits>
#include <utility>

#include "awfl_tables.h"

namespace awfl {

// ---------------------------------------------------------------------------
// compile-time loop: f(std::integral_constant<int, I>) for I = 0..N-1
// ---------------------------------------------------------------------------
template <class F, int... I>
__host__ __device__ __forceinline__ constexpr void sfor_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__host__ __device__ __forceinline__ constexpr void sfor(F&& f) {
  if constexpr (N > 0) sfor_impl(f, std::make_integer_sequence<int, N>{});
}

// ---------------------------------------------------------------------------
// GLL nodes/weights on [-1/2, 1/2] and the GLL collocation derivative
// (derivMatrix of Spatial::init, S:298-309), stored [in point][out point].
// ---------------------------------------------------------------------------
template <int N> struct Gll;
template <int N> struct Deriv;
template <int N> struct Tv;   // truncated TV quadratic form (TransformMatrices coefs_to_tv)

#define AWFL_GLL_SPEC(N)                                         \
  template <> struct Gll<N> {                                    \
    static constexpr const auto& pts = awfl_gll_pts_##N;         \
    static constexpr const auto& wts = awfl_gll_wts_##N;         \
  };
AWFL_GLL_SPEC(1) AWFL_GLL_SPEC(2) AWFL_GLL_SPEC(3) AWFL_GLL_SPEC(4) AWFL_GLL_SPEC(5)
AWFL_GLL_SPEC(6) AWFL_GLL_SPEC(7) AWFL_GLL_SPEC(8) AWFL_GLL_SPEC(9)
#undef AWFL_GLL_SPEC

#define AWFL_DT_SPEC(N)                                          \
  template <> struct Deriv<N> {                                  \
    static constexpr const auto& m = awfl_deriv_##N;             \
  };                                                             \
  template <> struct Tv<N> {                                     \
    static constexpr int n = AWFL_TV_##N##_NTERMS;               \
    static constexpr const auto& j = awfl_tv_##N##_j;            \
    static constexpr const auto& k = awfl_tv_##N##_k;            \
    static constexpr const auto& c = awfl_tv_##N##_c;            \
  };
AWFL_DT_SPEC(2) AWFL_DT_SPEC(3) AWFL_DT_SPEC(4) AWFL_DT_SPEC(5) AWFL_DT_SPEC(6)
AWFL_DT_SPEC(7) AWFL_DT_SPEC(8) AWFL_DT_SPEC(9)
#undef AWFL_DT_SPEC

// ---------------------------------------------------------------------------
// WENO: hs+2 polynomials [poly][cell][coef], convexified ideal weights and
// sigma (WenoLimiter.h:28-90).
// ---------------------------------------------------------------------------
template <int ORD> struct Weno;
#define AWFL_WENO_SPEC(O)                                        \
  template <> struct Weno<O> {                                   \
    static constexpr int hs = (O - 1) / 2;                       \
    static constexpr const auto& recon = awfl_weno_recon_##O;    \
    static constexpr const auto& idl = awfl_idl_##O;             \
    static constexpr double sigma = awfl_sigma_##O;              \
  };
AWFL_WENO_SPEC(3) AWFL_WENO_SPEC(5) AWFL_WENO_SPEC(7) AWFL_WENO_SPEC(9)
#undef AWFL_WENO_SPEC

// ---------------------------------------------------------------------------
// projections to NGLL GLL points: coefs_to_gll (= coefs_to_gll_lower) and
// sten_to_gll (= c2g_lower * sten_to_coefs), S:291-294, [in][out].
// ---------------------------------------------------------------------------
template <int ORD, int NGLL> struct Proj;
#define AWFL_PROJ_SPEC(O, M)                                              \
  template <> struct Proj<O, M> {                                         \
    static constexpr const auto& c2g = awfl_coefs_to_gll_lower_##O##_##M; \
    static constexpr const auto& s2g = awfl_sten_to_gll_##O##_##M;        \
  };
AWFL_PROJ_SPEC(3, 2) AWFL_PROJ_SPEC(3, 3)
AWFL_PROJ_SPEC(5, 2) AWFL_PROJ_SPEC(5, 3) AWFL_PROJ_SPEC(5, 4) AWFL_PROJ_SPEC(5, 5)
AWFL_PROJ_SPEC(7, 2) AWFL_PROJ_SPEC(7, 3) AWFL_PROJ_SPEC(7, 4) AWFL_PROJ_SPEC(7, 5)
AWFL_PROJ_SPEC(7, 6) AWFL_PROJ_SPEC(7, 7)
AWFL_PROJ_SPEC(9, 2) AWFL_PROJ_SPEC(9, 3) AWFL_PROJ_SPEC(9, 4) AWFL_PROJ_SPEC(9, 5)
AWFL_PROJ_SPEC(9, 6) AWFL_PROJ_SPEC(9, 7) AWFL_PROJ_SPEC(9, 8) AWFL_PROJ_SPEC(9, 9)
#undef AWFL_PROJ_SPEC

}  // namespace awfl
