// Finite-difference operators over tap accessors t(k), k = -M..M, shared by
// the acoustic paths.  C is any coefficient block with the fixed-order
// weights slap[]/slap_c0/sgrad[] and the caller's tables lap[]/grad[].
#pragma once

#include <hip/hip_runtime.h>

namespace swp {
namespace fd {

// The scale of a stencil sum by 1/h^2 or 1/h.  Kernels multiply by the
// reciprocal the host computed once (a float scale).  A Scale also carries h
// itself: built with SWP_EXACT_DIV (the parity build, which turns
// contraction off too) its scaling divides by h, one correctly rounded
// operation where the oracle divides, so that build's records can be held to
// the oracle at the last bits.
struct Scale {
    float inv, h;
};
__host__ __device__ inline Scale scale_of(float h) { return {1.0f / h, h}; }

__device__ __forceinline__ float scaled(float a, float inv) { return a * inv; }
__device__ __forceinline__ float scaled(float a, Scale s)
{
#ifdef SWP_EXACT_DIV
    return a / s.h;
#else
    return a * s.inv;
#endif
}
// a / (2 h): the first-order centred difference.
__device__ __forceinline__ float scaled_half(float a, float inv) { return a * (0.5f * inv); }
__device__ __forceinline__ float scaled_half(float a, Scale s)
{
#ifdef SWP_EXACT_DIV
    return a / (2.f * s.h);
#else
    return a * (0.5f * s.inv);
#endif
}

// Forward second derivative: order 4 in summed form, the other fixed orders
// in difference form (csrc/cuda/operators/laplace.cuh:68-387).
template <int M, class C, class T, class S>
__device__ __forceinline__ float lap_taps(const C &c, T t, S inv_h2)
{
    const float u0 = t(0);
    if constexpr (M == 1) {
        return scaled((t(-1) - u0) + (t(1) - u0), inv_h2);
    } else if constexpr (M == 2) {
        return scaled(c.slap[0] * (t(-1) + t(1)) + c.slap[1] * (t(-2) + t(2)) - c.slap_c0 * u0, inv_h2);
    } else {
        float acc = 0.f;
#pragma unroll
        for (int k = 1; k <= M; ++k) acc += c.slap[k - 1] * ((t(-k) - u0) + (t(k) - u0));
        return scaled(acc, inv_h2);
    }
}

// Runtime order from the table lap[1..M] (difference form).
template <class C, class T, class S>
__device__ __forceinline__ float lap_taps_rt(const C &c, T t, S inv_h2, int M)
{
    const float u0 = t(0);
    float acc = 0.f;
    for (int k = 1; k <= M; ++k) acc += c.lap[k] * ((t(k) - u0) + (t(-k) - u0));
    return scaled(acc, inv_h2);
}

// Centred first derivative (csrc/cuda/operators/gradient.cuh:47-88).
template <int M, class C, class T, class S>
__device__ __forceinline__ float d1_taps(const C &c, T t, S inv_h)
{
    if constexpr (M == 1) {
        return scaled_half(t(1) - t(-1), inv_h);
    } else {
        float acc = 0.f;
#pragma unroll
        for (int k = 1; k <= M; ++k) acc += c.sgrad[k - 1] * (t(k) - t(-k));
        return scaled(acc, inv_h);
    }
}

template <class C, class T, class S>
__device__ __forceinline__ float d1_taps_rt(const C &c, T t, S inv_h, int M)
{
    float acc = 0.f;
    for (int k = 1; k <= M; ++k) acc += c.grad[k] * (t(k) - t(-k));
    return scaled(acc, inv_h);
}

// MS > 0: fixed order MS; MS == 0: the runtime order M from the tables.
template <int MS, class C, class T, class S>
__device__ __forceinline__ float lap(const C &c, T t, S inv_h2, int M)
{
    if constexpr (MS > 0) return lap_taps<MS>(c, t, inv_h2);
    else return lap_taps_rt(c, t, inv_h2, M);
}

template <int MS, class C, class T, class S>
__device__ __forceinline__ float d1(const C &c, T t, S inv_h, int M)
{
    if constexpr (MS > 0) return d1_taps<MS>(c, t, inv_h);
    else return d1_taps_rt(c, t, inv_h, M);
}

}  // namespace fd
}  // namespace swp
