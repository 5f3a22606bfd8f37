// Device-only helpers shared by the HIP kernel files of libfdfd (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace fdfd {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

template <typename R> __device__ inline R shfl_x(R v, int o) { return __shfl_xor(v, o, 64); }

// Pointers read from job tables are generic; accesses through them compile to
// flat instructions, which also count against the LDS wait counter and so
// serialize LDS reads behind outstanding global loads.  These helpers access
// complex values through the global address space (global_load/store).
#define GLOBAL_AS __attribute__((address_space(1)))
template <typename R> struct ev2;
template <> struct ev2<float> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct ev2<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <typename C> __device__ inline C gload(const C* p) {
    using V = typename ev2<decltype(C{}.x)>::type;
    const V v = *(const GLOBAL_AS V*)p;
    C c;
    c.x = v.x;
    c.y = v.y;
    return c;
}
template <typename C> __device__ inline void gstore(C* p, C c) {
    using V = typename ev2<decltype(C{}.x)>::type;
    V v;
    v.x = c.x;
    v.y = c.y;
    *(GLOBAL_AS V*)p = v;
}
template <typename T> __device__ inline T gload_s(const T* p) { return *(const GLOBAL_AS T*)p; }
// two adjacent complex floats as one 16-byte access (p 16-byte aligned)
__device__ inline f4v gload2c(const float2* p) { return *(const GLOBAL_AS f4v*)p; }
__device__ inline void gstore2c(float2* p, f4v v) { *(GLOBAL_AS f4v*)p = v; }

// Split operands of the three-product (fp16x3) matrix products: a row of
// FP32 values with largest |re|, |im| = mx is scaled by 2^split_exp(mx), which
// brings mx to at most 2^14, and each scaled value x is split into
// hi = f16(x), lo = f16(x - hi) (x - hi is exact in FP32).
typedef _Float16 h8v __attribute__((ext_vector_type(8)));

constexpr int SP_EMAX = 60;  // row exponents are clamped to [-SP_EMAX, SP_EMAX]

// 2^e as a float, |e| <= 126
__device__ inline float exp2i(int e) { return __int_as_float((127 + e) << 23); }

__device__ inline int split_exp(float mx) {
    if (!(mx > 0.f)) return 0;
    int ex;
    const float f = frexpf(mx, &ex);            // mx = f 2^ex, f in [0.5, 1)
    const int cl = (f == 0.5f) ? ex - 1 : ex;  // ceil(log2 mx)
    return min(SP_EMAX, max(-SP_EMAX, 14 - cl));
}

__device__ inline void split_f16(float x, _Float16& h, _Float16& l) {
    h = (_Float16)x;
    l = (_Float16)(x - (float)h);
}

// 1 / z for a complex pivot with one hardware reciprocal: z = (a + ib) 2^e
// with max(|a|, |b|) in [0.5, 1) (exact power-of-two scaling, so no
// overflow or underflow in |.|^2), 1 / z = (a - ib) r 2^-e with r = 1 / d,
// d = a^2 + b^2 in [0.25, 2), r from v_rcp refined by Newton steps
// r <- r + r (1 - d r) (one for FP32, two for FP64: the FP64 estimate has
// about half the bits).  Relative error a few units in the last place, the
// accuracy of the three-division (Smith) form it replaces on short chains.
// z = 0 gives non-finite values; callers select them away.
__device__ inline float rcp_hw(float d) { return __builtin_amdgcn_rcpf(d); }
__device__ inline double rcp_hw(double d) { return __builtin_amdgcn_rcp(d); }
template <typename C> __device__ inline C crcp_scaled(C z) {
    using R = decltype(C{}.x);
    int e;
    (void)frexp(fmax(fabs(z.x), fabs(z.y)), &e);
    const R a = ldexp(z.x, -e), b = ldexp(z.y, -e);
    const R d = fma(a, a, b * b);
    R r = rcp_hw(d);
    r = fma(r, fma(-d, r, R(1)), r);
    if constexpr (sizeof(R) == 8) r = fma(r, fma(-d, r, R(1)), r);
    return mk<C>(ldexp(a * r, -e), ldexp(-b * r, -e));
}

// compile-time loop: f(std::integral_constant<int, I>) for I in [B, E)
template <int B, int E, typename F> __device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// 16x16x4 MFMA.  Operands: A[i = lane % 16][k = lane / 16],
// B[k = lane / 16][j = lane % 16]; result element v of a lane is
// D[row(lane, v)][lane % 16].  The f32 form keeps four consecutive rows per
// lane group, the f64 form interleaves them.
template <typename R> struct mfma_t;
template <> struct mfma_t<double> {
    using acc = d4v;
    static __device__ inline acc run(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct mfma_t<float> {
    using acc = f4v;
    static __device__ inline acc run(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

}  // namespace fdfd
