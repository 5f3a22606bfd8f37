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
