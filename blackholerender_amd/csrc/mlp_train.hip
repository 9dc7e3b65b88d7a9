// mlp_train.hip — training-data generation and training kernels of the learned surrogate
// (forward, loss, backward on bf16 MFMA GEMMs with FP32 accumulation, Adam in FP32).
#include "mlp.hpp"

namespace bhs {
namespace mlp {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f4;

__device__ __forceinline__ float urand(uint32_t key, uint32_t i, uint32_t k) {
    return (float)(pcg(key ^ pcg(i * 16u + k)) >> 8) * 0x1p-24f;
}

// ------------------------------------------------------------------------------------------------
// Training rays: r_obs log-uniform in [8, 1000] M; beta = b/b_crit log-spaced on both sides of 1
// (40% above: 1 + 10^U(log10 0.002, log10(beta_max - 1)); 20% below: 1 - 10^U(log10 0.002,
// log10 0.7)), 35% uniform in [0.3, beta_max] and 5% uniform in [0, 0.3) (plunging rays: outside
// the network's domain, they only feed the step table); b at most 0.95 of the tangential limit
// r/sqrt(1 - 2/r); the first equatorial crossing at swept angle psi uniform in [0, pi); the orbit
// plane tilted by U(15, 89) deg about a random line of nodes. The view direction is ingoing with
// impact parameter b.
__global__ void k_sample_rays(int64_t n, uint32_t key, float* __restrict__ pos3,
                              float* __restrict__ dir3) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t ii = (uint32_t)i;
    const float r = expf(kLogRLo + (kLogRHi - kLogRLo) * urand(key, ii, 0));
    const float f = sqrtf(1.0f - 2.0f / r);
    const float bmax = 0.95f * r / f;
    const float beta_max = bmax / kBcritF;
    const float u = urand(key, ii, 1);
    float beta;
    if (u < 0.4f) {
        const float tmax = log10f(beta_max - 1.0f);
        beta = 1.0f + exp10f(-2.69897f + (tmax + 2.69897f) * urand(key, ii, 2));
    } else if (u < 0.6f) {
        beta = 1.0f - exp10f(-2.69897f + 2.5440680f * urand(key, ii, 2));   // 0.002 .. 0.7
    } else if (u < 0.95f) {
        beta = 0.3f + (beta_max - 0.3f) * urand(key, ii, 2);
    } else {
        beta = 0.3f * urand(key, ii, 2);
    }
    float b = beta * kBcritF;
    if (b > bmax) b = bmax;
    const float psi = kPi * urand(key, ii, 3);
    const float inc = (15.0f + 74.0f * urand(key, ii, 4)) * (kPi / 180.0f);
    const float az = 2.0f * kPi * urand(key, ii, 5);
    float saz, caz, sinc, cinc, sps, cps;
    sincosf(az, &saz, &caz);
    sincosf(inc, &sinc, &cinc);
    sincosf(psi, &sps, &cps);
    const V3<float> N{caz, 0.0f, saz};
    const V3<float> M{-saz * cinc, sinc, caz * cinc};
    const V3<float> e1{cps * N.x - sps * M.x, cps * N.y - sps * M.y, cps * N.z - sps * M.z};
    const V3<float> e2{sps * N.x + cps * M.x, sps * N.y + cps * M.y, sps * N.z + cps * M.z};
    const float sa = fminf(1.0f, b * f / r);
    const float ca = sqrtf(fmaxf(0.0f, 1.0f - sa * sa));
    pos3[3 * i] = r * e1.x;
    pos3[3 * i + 1] = r * e1.y;
    pos3[3 * i + 2] = r * e1.z;
    dir3[3 * i] = -ca * e1.x + sa * e2.x;
    dir3[3 * i + 1] = -ca * e1.y + sa * e2.y;
    dir3[3 * i + 2] = -ca * e1.z + sa * e2.z;
}

// Features, targets and masks from the integrator's outputs (FP64 march, root-refined crossings,
// crossings recorded at any radius). Rays outside the surrogate's domain get all-zero masks.
// steps_max[step_bin] collects the largest step count per (log r_obs, x2) bin of the domain rays.
__global__ void k_targets(int64_t n, const float* __restrict__ pos3, const float* __restrict__ dir3,
                          const SoA out, float* __restrict__ X, float* __restrict__ Y,
                          float* __restrict__ M, uint32_t* __restrict__ steps_max) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const V3<float> P{pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2]};
    const V3<float> d{dir3[3 * i], dir3[3 * i + 1], dir3[3 * i + 2]};
    const Setup<float> s = lawS_setup(P, d);
    const RayGeom g = ray_geom(P, d, s);
    float x[3];
    features(g, x);
    const bool dom = in_domain(g);
    const uint32_t fl = out.flags[i], nc = out.n_cross[i];
    const bool maxed = (fl & 1u) != 0;
    const bool escaped = (fl & 2u) != 0 && g.bm > 0.0f;
    float y0 = 0.0f, yu[2] = {0.0f, 0.0f}, mu[2] = {0.0f, 0.0f};
    if (dom) {
        const Orbit ob = orbit(g);
        if (escaped) {
            const V3<float> e{out.esc[i], out.esc[i + out.n], out.esc[i + 2 * out.n]};
            const float phim = atan2f(dot3(e, g.e2), dot3(e, g.e1));
            const float k = rintf((ob.phi_end - phim) * (0.5f / kPi));
            y0 = (phim + 2.0f * kPi * k - ob.phi_end) * (1.0f / kResScale);
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const float phi = g.psi + (float)q * kPi;
            if ((uint32_t)q >= nc || !(phi < ob.phi_end)) continue;
            const float ue = orbit_u(ob, phi);
            const float ut = 1.0f / out.disk[i + q * out.n];
            const float t = logf(ut / ue) * (1.0f / kURes);
            if (ue > 0.0f && fabsf(t) < 1000.0f) {
                yu[q] = t;
                mu[q] = 1.0f;
            }
        }
    }
    const float m = dom ? 1.0f : 0.0f;
    for (int k = 0; k < 3; ++k) X[3 * i + k] = x[k];
    Y[5 * i + 0] = y0;
    Y[5 * i + 1] = yu[0];
    Y[5 * i + 2] = yu[1];
    Y[5 * i + 3] = nc >= 1 ? 1.0f : 0.0f;
    Y[5 * i + 4] = nc >= 2 ? 1.0f : 0.0f;
    M[5 * i + 0] = (escaped && !maxed) ? m : 0.0f;
    M[5 * i + 1] = mu[0] * m;
    M[5 * i + 2] = mu[1] * m;
    M[5 * i + 3] = !maxed ? m : 0.0f;    // out-of-steps rays: reach unknown
    M[5 * i + 4] = !maxed ? m : 0.0f;
    if ((dom || plunging(g)) && steps_max)
        atomicMax(&steps_max[step_bin(g.r, x2_of(g.bm))], (uint32_t)out.steps[i]);
}

// ------------------------------------------------------------------------------------------------
// Minibatch: index b -> dataset row (hash of step key and b), first layer (FP32):
// z1 = W1 x + b1, a1 = silu(z1).
__global__ void k_l1_fwd(int32_t B, uint32_t key, int32_t ident, int64_t n,
                         const float* __restrict__ X, const float* __restrict__ P,
                         int32_t* __restrict__ idx, float* __restrict__ z1, float* __restrict__ a1) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)B * kH) return;
    const int32_t b = (int32_t)(t / kH), j = (int32_t)(t % kH);
    int64_t row = b;
    if (!ident) row = n > 0 ? (int64_t)(pcg(key ^ pcg((uint32_t)b * 2654435761u)) % (uint64_t)n) : 0;
    if (j == 0) idx[b] = (int32_t)row;
    float x[3] = {X[3 * row], X[3 * row + 1], X[3 * row + 2]};
    const float z = l1_pre(P + kW1 + 3 * j, P[kB1 + j], x);
    z1[t] = z;
    a1[t] = silu(z);
}

// Generic bf16 MFMA GEMM: C[m][n] = sum_k A(m,k) B(n,k) with A(m,k) = A[m*sam + k*sak] and
// B(n,k) = B[n*sbn + k*sbk] read from FP32 and rounded to bf16 (RNE), FP32 accumulation. 64x64
// tile per 256-thread block (each wave 32x32 = 2x2 tiles of v_mfma_f32_16x16x32_bf16), K in
// chunks of 32 through LDS; blockIdx.z splits K into ranges of kspl. Epilogues:
//   EPI_STORE      C[m*scm + n*scn] = acc                       (split-K partials: C + z*M*N)
//   EPI_BIAS_SILU  z = acc + bias[n]; C = z; C2 = silu(z)        (forward layer)
//   EPI_DSILU      C = acc * silu'(Z[m*scm + n*scn])              (backward through activation)
enum { EPI_STORE = 0, EPI_BIAS_SILU = 1, EPI_DSILU = 2 };

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, int kspl,
                                              const float* __restrict__ A, int64_t sam, int64_t sak,
                                              const float* __restrict__ B, int64_t sbn, int64_t sbk,
                                              float* __restrict__ C, int64_t scm, int64_t scn,
                                              const float* __restrict__ bias,
                                              const float* __restrict__ Z, float* __restrict__ C2) {
    __shared__ uint16_t As[64][40];
    __shared__ uint16_t Bs[64][40];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
    const int k_lo = blockIdx.z * kspl;
    const int k_hi = k_lo + kspl < K ? k_lo + kspl : K;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int g = lane >> 4, c16 = lane & 15;
    f4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = k_lo; k0 < k_hi; k0 += 32) {
        for (int e = tid; e < 64 * 32; e += 256) {
            const int rr = e >> 5, kk = e & 31;
            const int gm = m0 + rr, gn = n0 + rr, gk = k0 + kk;
            const bool kin = gk < k_hi;
            As[rr][kk] = (gm < M && kin) ? bf16_bits(A[gm * sam + gk * sak]) : (uint16_t)0;
            Bs[rr][kk] = (gn < N && kin) ? bf16_bits(B[gn * sbn + gk * sbk]) : (uint16_t)0;
        }
        __syncthreads();
        bf16x8 fa[2], fb[2];
        for (int a = 0; a < 2; ++a)
            for (int j = 0; j < 8; ++j) {
                fa[a][j] = __builtin_bit_cast(__bf16, As[wm + 16 * a + c16][8 * g + j]);
                fb[a][j] = __builtin_bit_cast(__bf16, Bs[wn + 16 * a + c16][8 * g + j]);
            }
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
    float* Cz = C + (EPI == EPI_STORE ? (int64_t)blockIdx.z * M * N : 0);
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int jj = 0; jj < 4; ++jj) {
                const int m = m0 + wm + 16 * a + 4 * g + jj, nn = n0 + wn + 16 * b + c16;
                if (m >= M || nn >= N) continue;
                const int64_t o = m * scm + nn * scn;
                const float v = acc[a][b][jj];
                if (EPI == EPI_STORE) {
                    Cz[o] = v;
                } else if (EPI == EPI_BIAS_SILU) {
                    const float z = v + bias[nn];
                    C[o] = z;
                    C2[o] = silu(z);
                } else {
                    C[o] = v * dsilu(Z[o]);
                }
            }
}

// out[i] = sum over S split-K partial planes of size cnt.
__global__ void k_sum_splits(int S, int64_t cnt, const float* __restrict__ ws, float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    float s = 0.0f;
    for (int z = 0; z < S; ++z) s += ws[z * cnt + i];
    out[i] = s;
}

// Output layer (bf16 operands, FP32 accumulation, as the MFMA inference), loss and its gradient for
// one sample per thread. Loss (mean over the batch): sum_o w_o m_o (y_o - t_o)^2 over the
// regressions (w = 1, 8, 8) + 0.1 * mean of the two logistic losses.
__global__ void k_head(int32_t B, const float* __restrict__ a3, const int32_t* __restrict__ idx,
                       const float* __restrict__ Y, const float* __restrict__ Mk,
                       const float* __restrict__ P, float reg_norm, float* __restrict__ dy,
                       float* __restrict__ loss) {
    const int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (b < B) {
        const float* a = a3 + (int64_t)b * kH;
        float y[kOut];
        for (int o = 0; o < kOut; ++o) {   // bf16 operands, as the MFMA output layer
            float s = 0.0f;
            const float* w = P + kW4 + o * kH;
            for (int k = 0; k < kH; ++k) s = __builtin_fmaf(bf16_round(w[k]), bf16_round(a[k]), s);
            y[o] = s + P[kB4 + o];
        }
        const int64_t row = idx[b];
        const float* t = Y + 5 * row;
        const float* m = Mk + 5 * row;
        const float wr[3] = {1.0f, 8.0f, 8.0f};
        for (int o = 0; o < 3; ++o) {
            const float e = y[o] - t[o];
            l += wr[o] * m[o] * e * e * reg_norm;
            dy[b * kOut + o] = 2.0f * wr[o] * m[o] * e * reg_norm;
        }
        for (int o = 3; o < 5; ++o) {
            const float z = y[o];
            const float sp = z > 0.0f ? z + log1pf(__expf(-z)) : log1pf(__expf(z));
            const float sg = 1.0f / (1.0f + __expf(-z));
            const float wb = 0.1f / (2.0f * (float)B) * m[o];
            l += wb * (sp - t[o] * z);
            dy[b * kOut + o] = wb * (sg - t[o]);
        }
    }
    // block sum of the loss into *loss (reporting only)
    __shared__ float red[256];
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// dz3 = (W4^T dy) * silu'(z3), one (sample, neuron) per thread.
__global__ void k_head_bwd(int32_t B, const float* __restrict__ dy, const float* __restrict__ P,
                           const float* __restrict__ z3, float* __restrict__ dz3) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)B * kH) return;
    const int32_t b = (int32_t)(t / kH), k = (int32_t)(t % kH);
    float s = 0.0f;
    for (int o = 0; o < kOut; ++o) s = __builtin_fmaf(dy[b * kOut + o], bf16_round(P[kW4 + o * kH + k]), s);
    dz3[t] = s * dsilu(z3[t]);
}

// Gradients of the small parameter blocks (W1, b1, b2, b3, W4, b4), one parameter per thread,
// summed over the batch in a fixed order.
__global__ void k_small_grads(int32_t B, const float* __restrict__ X, const int32_t* __restrict__ idx,
                              const float* __restrict__ dz1, const float* __restrict__ dz2,
                              const float* __restrict__ dz3, const float* __restrict__ dy,
                              const float* __restrict__ a3, float* __restrict__ grad) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    float s = 0.0f;
    if (t < kH * kIn) {                              // W1[j][k]
        const int j = t / kIn, k = t % kIn;
        for (int b = 0; b < B; ++b) s = __builtin_fmaf(dz1[(int64_t)b * kH + j], X[3 * (int64_t)idx[b] + k], s);
        grad[kW1 + t] = s;
    } else if (t < kH * kIn + kH) {                  // b1
        const int j = t - kH * kIn;
        for (int b = 0; b < B; ++b) s += dz1[(int64_t)b * kH + j];
        grad[kB1 + j] = s;
    } else if (t < kH * kIn + 2 * kH) {              // b2
        const int j = t - kH * kIn - kH;
        for (int b = 0; b < B; ++b) s += dz2[(int64_t)b * kH + j];
        grad[kB2 + j] = s;
    } else if (t < kH * kIn + 3 * kH) {              // b3
        const int j = t - kH * kIn - 2 * kH;
        for (int b = 0; b < B; ++b) s += dz3[(int64_t)b * kH + j];
        grad[kB3 + j] = s;
    } else if (t < kH * kIn + 3 * kH + kOut * kH) {  // W4[o][k]
        const int q = t - kH * kIn - 3 * kH, o = q / kH, k = q % kH;
        for (int b = 0; b < B; ++b) s = __builtin_fmaf(dy[b * kOut + o], bf16_round(a3[(int64_t)b * kH + k]), s);
        grad[kW4 + q] = s;
    } else if (t < kH * kIn + 3 * kH + kOut * kH + kOut) {   // b4
        const int o = t - kH * kIn - 3 * kH - kOut * kH;
        for (int b = 0; b < B; ++b) s += dy[b * kOut + o];
        grad[kB4 + o] = s;
    }
}

// Adam (Kingma & Ba 2015) with bias correction, FP32.
__global__ void k_adam(int32_t n, float* __restrict__ P, const float* __restrict__ G,
                       float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = G[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    P[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

__global__ void k_fill(float* p, int n, float v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_copy1(const float* src, float* dst) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = *src;
}

// He-style initialisation from a hashed uniform: W ~ U(-1, 1) * sqrt(3 / fan_in) * gain, b = 0.
__global__ void k_init(uint32_t key, float* __restrict__ P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    float fan, scale = 1.0f;
    if (i < kB1) fan = kIn;
    else if (i < kW2) { P[i] = 0.0f; return; }
    else if (i < kB2) fan = kH;
    else if (i < kW3) { P[i] = 0.0f; return; }
    else if (i < kB3) fan = kH;
    else if (i < kW4) { P[i] = 0.0f; return; }
    else if (i < kB4) { fan = kH; scale = 0.1f; }
    else { P[i] = 0.0f; return; }
    const float u = 2.0f * urand(key, (uint32_t)i, 7) - 1.0f;
    P[i] = u * sqrtf(3.0f / fan) * 1.7f * scale;
}

}  // namespace

// ------------------------------------------------------------------------------------------------
hipError_t launch_sample_rays(int64_t n, uint32_t key, float* pos3, float* dir3, hipStream_t s) {
    hipLaunchKernelGGL(k_sample_rays, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, key, pos3, dir3);
    return hipGetLastError();
}

hipError_t launch_targets(int64_t n, const float* pos3, const float* dir3, const SoA& out, float* X,
                          float* Y, float* M, uint32_t* steps_max, hipStream_t s) {
    hipLaunchKernelGGL(k_targets, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, pos3, dir3,
                       out, X, Y, M, steps_max);
    return hipGetLastError();
}

hipError_t launch_init(uint32_t key, float* P, hipStream_t s) {
    hipLaunchKernelGGL(k_init, dim3((kParams + 255) / 256), dim3(256), 0, s, key, P);
    return hipGetLastError();
}

// Split-K factor of the weight-gradient GEMMs (K = batch): about 64 rows per block and split.
static int splits_for(int B) {
    int s = B / 512;
    if (s < 1) s = 1;
    if (s > 64) s = 64;
    return s;
}

size_t train_workspace_floats(int32_t B) {
    const size_t bh = (size_t)B * kH;
    // z1 a1 z2 a2 z3 a3 dz3 dz2 dz1 (9 B*H) + dy (5B) + idx (B) + split-K planes + loss
    return 9 * bh + 5 * (size_t)B + (size_t)B + (size_t)splits_for(B) * kH * kH + 64;
}

hipError_t launch_train_step(const TrainArgs& a, hipStream_t s) {
    const int32_t B = a.batch;
    const int64_t bh = (int64_t)B * kH;
    float* w = a.ws;
    float* z1 = w;       float* a1 = w + bh;  float* z2 = w + 2 * bh; float* a2 = w + 3 * bh;
    float* z3 = w + 4 * bh; float* a3 = w + 5 * bh; float* dz3 = w + 6 * bh; float* dz2 = w + 7 * bh;
    float* dz1 = w + 8 * bh;
    float* dy = w + 9 * bh;
    int32_t* idx = (int32_t*)(w + 9 * bh + 5 * B);
    float* split = w + 9 * bh + 6 * B;
    const int S = splits_for(B);
    float* loss = split + (size_t)S * kH * kH;
    hipError_t e;
    const dim3 ew((unsigned)((bh + 255) / 256));
    hipLaunchKernelGGL(k_l1_fwd, ew, dim3(256), 0, s, B, a.key, a.ident, a.n, a.X, a.P, idx, z1, a1);
    const dim3 gf((B + 63) / 64, kH / 64, 1);
    // z2 = a1 W2^T + b2, a2 = silu(z2)
    hipLaunchKernelGGL(k_gemm<EPI_BIAS_SILU>, gf, dim3(256), 0, s, B, kH, kH, kH, a1, (int64_t)kH,
                       (int64_t)1, a.P + kW2, (int64_t)kH, (int64_t)1, z2, (int64_t)kH, (int64_t)1,
                       a.P + kB2, (const float*)nullptr, a2);
    hipLaunchKernelGGL(k_gemm<EPI_BIAS_SILU>, gf, dim3(256), 0, s, B, kH, kH, kH, a2, (int64_t)kH,
                       (int64_t)1, a.P + kW3, (int64_t)kH, (int64_t)1, z3, (int64_t)kH, (int64_t)1,
                       a.P + kB3, (const float*)nullptr, a3);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, s, loss, 1, 0.0f);
    int32_t nreg = a.reg_count > 0 ? a.reg_count : B;
    hipLaunchKernelGGL(k_head, dim3((B + 255) / 256), dim3(256), 0, s, B, a3, idx, a.Y, a.M, a.P,
                       1.0f / (float)nreg, dy, loss);
    hipLaunchKernelGGL(k_head_bwd, ew, dim3(256), 0, s, B, dy, a.P, z3, dz3);
    // dW3 = dz3^T a2 (K = batch, split), da2 = dz3 W3 -> dz2
    const int kspl = (B + S - 1) / S;
    const dim3 gw(kH / 64, kH / 64, S);
    hipLaunchKernelGGL(k_gemm<EPI_STORE>, gw, dim3(256), 0, s, kH, kH, B, ((kspl + 31) / 32) * 32,
                       dz3, (int64_t)1, (int64_t)kH, a2, (int64_t)1, (int64_t)kH, split, (int64_t)kH,
                       (int64_t)1, (const float*)nullptr, (const float*)nullptr, (float*)nullptr);
    hipLaunchKernelGGL(k_sum_splits, dim3((kH * kH + 255) / 256), dim3(256), 0, s, S,
                       (int64_t)kH * kH, split, a.G + kW3);
    hipLaunchKernelGGL(k_gemm<EPI_DSILU>, gf, dim3(256), 0, s, B, kH, kH, kH, dz3, (int64_t)kH,
                       (int64_t)1, a.P + kW3, (int64_t)1, (int64_t)kH, dz2, (int64_t)kH, (int64_t)1,
                       (const float*)nullptr, z2, (float*)nullptr);
    // dW2 = dz2^T a1, da1 = dz2 W2 -> dz1
    hipLaunchKernelGGL(k_gemm<EPI_STORE>, gw, dim3(256), 0, s, kH, kH, B, ((kspl + 31) / 32) * 32,
                       dz2, (int64_t)1, (int64_t)kH, a1, (int64_t)1, (int64_t)kH, split, (int64_t)kH,
                       (int64_t)1, (const float*)nullptr, (const float*)nullptr, (float*)nullptr);
    hipLaunchKernelGGL(k_sum_splits, dim3((kH * kH + 255) / 256), dim3(256), 0, s, S,
                       (int64_t)kH * kH, split, a.G + kW2);
    hipLaunchKernelGGL(k_gemm<EPI_DSILU>, gf, dim3(256), 0, s, B, kH, kH, kH, dz2, (int64_t)kH,
                       (int64_t)1, a.P + kW2, (int64_t)1, (int64_t)kH, dz1, (int64_t)kH, (int64_t)1,
                       (const float*)nullptr, z1, (float*)nullptr);
    const int nsmall = kH * kIn + 3 * kH + kOut * kH + kOut;
    hipLaunchKernelGGL(k_small_grads, dim3((nsmall + 255) / 256), dim3(256), 0, s, B, a.X, idx, dz1,
                       dz2, dz3, dy, a3, a.G);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (a.adam) {
        hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, s, (int32_t)kParams,
                           a.P, a.G, a.m, a.v, a.lr, a.b1, a.b2, a.eps, a.c1, a.c2);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (a.loss_dev) {
        hipLaunchKernelGGL(k_copy1, dim3(1), dim3(64), 0, s, loss, a.loss_dev);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace mlp
}  // namespace bhs
