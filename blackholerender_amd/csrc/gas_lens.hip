// gas_lens.hip — the gas lens table: a fan of rays from a static camera on +x, each marched with
// law S's float setup and RK4, tabulating log r at uniformly spaced swept angles
// (shaders/gas_lens.comp:11-44, shaders/gas_lens.glsl:13-23; built by GasParticles,
// src/GasParticles.cpp:164-166, with uAlphaCrit = shadowAngularRadius(camR)).
//
// One thread per fan ray (64-lane blocks, as the compute shader's local size): the table is
// latency-bound by its longest ray (up to a full turn of swept angle at steps of 0.4% of r near
// the photon sphere), not by throughput. Arithmetic as the CPU restatement in oracle/: the setup
// and RK4 of the march, the IEEE-only atan2 for the swept angle of a step, log / cos / sin in
// double rounded to float, the fan exponent 3 as an exact cube.

#include "gas_lens.hpp"
#include "lawS_device.hpp"

namespace bhs {
namespace {

constexpr int kLensBlock = 64;

__device__ __forceinline__ float lens_alpha(float u, float ac) {   // gas_lens.glsl:19-23
    const float pi = 3.14159265358979f;
    const float uc = ac / pi;
    if (u < uc) {
        const float t = (uc - u) / uc;
        return ac - ac * (t * t * t);
    }
    const float t = (u - uc) / (1.0f - uc);
    return ac + (pi - ac) * (t * t * t);
}

__device__ __forceinline__ float log_d(float x) { return (float)::log((double)x); }

// Fan ray i of the table of a (every row j of column i).
__device__ __forceinline__ void lens_ray(const LensArgs& a, int32_t i, float* __restrict__ table) {
    const int64_t n = a.n_rays;
    const float al = lens_alpha(((float)i + 0.5f) / (float)a.n_rays, a.alpha_crit);   // :14
    const V3<float> P{a.cam_r, 0.0f, 0.0f};                                            // :16
    const V3<float> d{-(float)::cos((double)al), (float)::sin((double)al), 0.0f};
    const Setup<float> st = lawS_setup(P, d);                                          // :17-19
    V3<float> x = P, v = st.v;
    const float dphi_node = 6.28318530717959f / (float)(a.n_phis - 1);                // :20
    const float log_h = log_d(2.0f);                                                   // :21
    table[i] = log_d(a.cam_r);                                                         // :22
    int32_t j = 1;
    float phi = 0.0f, r = a.cam_r, r2 = dot3(x, x);
    for (int32_t s = 0; s < a.max_steps && j < a.n_phis; ++s) {                       // :25
        const float ds = (r < 8.0f ? 0.004f : 0.01f) * r;                              // :27
        const V3<float> x0 = x;
        const float r0 = r;
        const float vlen = __builtin_sqrtf(dot3(v, v));
        rk4(x, v, st.c3, ds / vlen, rsqrt_rn_core(r2));                                // :30
        r2 = dot3(x, x);
        r = __builtin_sqrtf(r2);                                                       // :31
        const float cz = __builtin_fmaf(x0.x, x.y, -(x0.y * x.x));
        const float cd = __builtin_fmaf(x0.y, x.y, x0.x * x.x);
        const float dphi = atan2_ieee(cz, cd);                                         // :32
        const float phi1 = phi + dphi;
        for (; j < a.n_phis && (float)j * dphi_node <= phi1; ++j) {                    // :34-37
            const float den = dphi > 1e-12f ? dphi : 1e-12f;
            float t = ((float)j * dphi_node - phi) / den;
            t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
            const float lr = log_d(__builtin_fmaf(t, r - r0, r0));
            table[(int64_t)j * n + i] = lr > log_h ? lr : log_h;
        }
        phi = phi1;
        if (r < 2.0f + 0.01f || (r > a.r_max && dot3(x, v) > 0.0f)) break;            // :39
    }
    const float fill = r < 10.0f ? log_h : log_d(a.r_max);                            // :42-43
    for (; j < a.n_phis; ++j) table[(int64_t)j * n + i] = fill;
}

__global__ __launch_bounds__(kLensBlock) void k_gas_lens(const LensArgs a,
                                                         float* __restrict__ table) {
    const int32_t i = (int32_t)(blockIdx.x * kLensBlock + threadIdx.x);
    if (i < a.n_rays) lens_ray(a, i, table);
}

// Several tables of the same size in one launch (blockIdx.y: the table), so that their fans fill
// more of the chip than one table's n_rays / 64 waves.
__global__ __launch_bounds__(kLensBlock) void k_gas_lens_batch(const LensBatch b,
                                                               float* __restrict__ tables) {
    const int32_t i = (int32_t)(blockIdx.x * kLensBlock + threadIdx.x);
    const uint32_t k = blockIdx.y;
    if (i >= b.n_rays) return;
    const LensArgs a{b.cam_r[k], b.alpha_crit[k], b.r_max, b.n_rays, b.n_phis, b.max_steps};
    lens_ray(a, i, tables + (int64_t)k * b.n_rays * b.n_phis);
}

}  // namespace

hipError_t launch_gas_lens(const LensArgs& a, float* table, hipStream_t s) {
    if (a.n_rays <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gas_lens, dim3((uint32_t)((a.n_rays + kLensBlock - 1) / kLensBlock)),
                       dim3(kLensBlock), 0, s, a, table);
    return hipGetLastError();
}

hipError_t launch_gas_lens_batch(const LensBatch& b, float* tables, hipStream_t s) {
    if (b.n_rays <= 0 || b.n_tables <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gas_lens_batch,
                       dim3((uint32_t)((b.n_rays + kLensBlock - 1) / kLensBlock),
                            (uint32_t)b.n_tables),
                       dim3(kLensBlock), 0, s, b, tables);
    return hipGetLastError();
}

}  // namespace bhs
