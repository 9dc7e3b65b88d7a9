// capi_mlp.hip — the C-ABI of the learned surrogate (bhs_mlp_*, bhs_render_surrogate): model
// storage, training-set generation, the training loop, validation / error map and the surrogate
// render with integrator fallback.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/bhs_trace.h"
#include "capi_internal.hpp"
#include "mlp.hpp"

static_assert(BHS_MLP_PARAMS == bhs::mlp::kParams, "parameter count");

struct bhs_mlp {
    int device = 0;
    bhs_march_params trained_for{};
    float* P = nullptr;        // parameters, gradient, Adam moments: 4 x kParams floats
    float* G = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    void* packed = nullptr;    // inference weights
    bool packed_ok = false;
    uint32_t* steps_dev = nullptr;
    // training set: X (3n floats), Y and M (5n each)
    void* data = nullptr;
    int64_t n_data = 0;
    // training workspace and per-step losses
    void* ws = nullptr;
    size_t ws_bytes = 0;
    void* loss = nullptr;
    size_t loss_bytes = 0;
    // fallback lists of the surrogate render: [counters (256 B) | late ids | early ids]
    void* fb = nullptr;
    size_t fb_bytes = 0;
    // table mode: [x2 ranges (512 B) | slots (256 B) | pad to 1 KiB | ncam x kPhiNodes swept
    // angles | ncam n2 n3 nodes x 4 floats]; tab_n2 = 0: per-ray inference
    void* tab = nullptr;
    size_t tab_bytes = 0;
    int32_t tab_n2 = 0, tab_n3 = 0;
    bhs::mlp::SurConst k{};
    float tol_phi = 1e-4f;
    int64_t adam_t = 0;
    uint32_t seed = 0;
};

using namespace bhs::capi;
namespace M = bhs::mlp;

namespace {

constexpr float kAdamB1 = 0.9f, kAdamB2 = 0.999f, kAdamEps = 1e-8f;

void free_model(bhs_mlp* m) {
    if (!m) return;
    DeviceGuard g(m->device);
    (void)hipDeviceSynchronize();
    for (void* p : {(void*)m->P, m->packed, (void*)m->steps_dev, m->data, m->ws, m->loss, m->fb,
                    m->tab})
        if (p) (void)hipFree(p);
    delete m;
}

// Grows a model-owned buffer (ensure() with the context's error reporting).
int grow(bhs_ctx* ctx, void** buf, size_t* cap, size_t need, const char* what) {
    return ensure(ctx, buf, cap, need, what);
}

int pack(bhs_ctx* ctx, bhs_mlp* m, hipStream_t s) {
    if (m->packed_ok) return BHS_OK;
    hipError_t e = M::launch_pack(m->P, m->packed, s);
    if (e != hipSuccess) return hip_fail(ctx, e, "bhs_mlp pack");
    m->packed_ok = true;
    return BHS_OK;
}

bool same_march(const bhs_march_params& a, const bhs_march_params& b) {
    return a.disk_mode == b.disk_mode && a.step_scale == b.step_scale &&
           a.disk_step == b.disk_step && a.disk_inner == b.disk_inner &&
           a.disk_outer == b.disk_outer && a.disk_thickness == b.disk_thickness;
}

// Device ray set: pos3, dir3 (3n floats each) and the FP64 law-S outputs of the rays with the
// crossings recorded at any radius (root-refined when root64, else the march's chord crossings).
struct RaySet {
    void* mem = nullptr;
    float* pos3 = nullptr;
    float* dir3 = nullptr;
    bhs::SoA soa{};
    ~RaySet() {
        if (mem) (void)hipFree(mem);
    }
};

int march_set(bhs_ctx* ctx, bhs_mlp* m, int64_t n, uint32_t key, bool root64, RaySet* rs,
              hipStream_t s) {
    const size_t bytes = (size_t)n * (24 + 1 + 1 + 1 + 2 + 12 + 24 + 4) + 4096;
    hipError_t e = hipMalloc(&rs->mem, bytes);
    if (e != hipSuccess) {
        rs->mem = nullptr;
        (void)hipGetLastError();
        return fail(ctx, BHS_E_ALLOC, "device allocation of %s failed", "surrogate ray set");
    }
    char* p = (char*)rs->mem;
    rs->pos3 = (float*)p;
    rs->dir3 = rs->pos3 + 3 * n;
    float* esc = rs->dir3 + 3 * n;
    float* disk = esc + 3 * n;
    float* b = disk + 6 * n;
    uint16_t* steps = (uint16_t*)(b + n);
    uint8_t* u8 = (uint8_t*)(steps + n);
    rs->soa.esc = esc;
    rs->soa.disk = disk;
    rs->soa.b = b;
    rs->soa.steps = steps;
    rs->soa.cls = u8;
    rs->soa.flags = u8 + n;
    rs->soa.n_cross = u8 + 2 * n;
    rs->soa.n = n;
    if ((e = M::launch_sample_rays(n, key, rs->pos3, rs->dir3, s)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_mlp sample rays");
    bhs_march_params mp = m->trained_for;
    mp.precision = BHS_PREC_F64;
    mp.root64 = root64 ? 1 : 0;
    mp.cross_all = 1;
    if (mp.max_steps < 4096) mp.max_steps = 4096;
    bhs::MarchConst c = make_const(&mp);
    bhs::Scratch sc;
    int st;
    if ((st = prepare_scratch(ctx, 0, 0, false, &sc)) != BHS_OK) return st;
    if ((e = bhs::launch_zero(ctx->scratch, kStatsBytes, s)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_mlp clear");
    if ((e = bhs::launch_march_rays(rs->pos3, rs->dir3, n, c, rs->soa, sc, s, nullptr, nullptr)) !=
        hipSuccess)
        return hip_fail(ctx, e, "bhs_mlp march");
    return BHS_OK;
}

uint32_t mix(uint32_t a, uint32_t b) {
    uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u + (a << 6) + (a >> 2));
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    return h;
}

float quantile(std::vector<float>& v, double q) {
    if (v.empty()) return 0.0f;
    const size_t k = std::min(v.size() - 1, (size_t)(q * (double)(v.size() - 1) + 0.5));
    std::nth_element(v.begin(), v.begin() + (long)k, v.end());
    return v[k];
}

// Table-mode buffers for ncam cameras (tab->n2 = 0 when the model runs per ray).
int table_for(bhs_ctx* ctx, bhs_mlp* m, int32_t ncam, M::SurTable* tab) {
    *tab = M::SurTable{nullptr, nullptr, nullptr, nullptr, 0, 0, nullptr};
    if (m->tab_n2 <= 0) return BHS_OK;
    const size_t nodes = (size_t)m->tab_n2 * (size_t)m->tab_n3 * (size_t)ncam;
    const size_t phi_bytes = (size_t)ncam * M::kPhiNodes * 4;
    int st;
    // [range (64 x 2 floats) | slot (64) | per-camera setup terms (64 x 16 words) | phi | nodes]
    constexpr size_t kHead = 512 + 512 + 64 * 64;
    if ((st = grow(ctx, &m->tab, &m->tab_bytes, kHead + phi_bytes + nodes * 16,
                   "surrogate table")) != BHS_OK)
        return st;
    char* base = (char*)m->tab;
    *tab = M::SurTable{(float*)(base + kHead + phi_bytes), (float*)(base + kHead), (float*)base,
                       (uint32_t*)(base + 512), m->tab_n2, m->tab_n3, (float*)(base + 1024)};
    return BHS_OK;
}

// Fallback lists for n_items work items: [counters (256 B) | late ids | early ids].
int lists_for(bhs_ctx* ctx, bhs_mlp* m, uint32_t n_items, M::SurLists* fl) {
    int st;
    // statistics partials: 4 words per table-pass block (at most one block per 256 items)
    const size_t chunks = ((size_t)n_items + 255) / 256 + 1;
    const size_t part_words = 2 * 4 * chunks;
    if ((st = grow(ctx, &m->fb, &m->fb_bytes,
                   256 + (size_t)n_items * 12 + (part_words + chunks) * 4, "fallback lists")) !=
        BHS_OK)
        return st;
    // [counters: late, early, early work, late band, late band work | late | early | late band |
    //  statistics partials]
    uint32_t* cnt = (uint32_t*)m->fb;
    uint32_t* late = (uint32_t*)((char*)m->fb + 256);
    *fl = M::SurLists{late,        cnt,     late + n_items, cnt + 1, cnt + 2,
                      late + 2 * (size_t)n_items, cnt + 3, cnt + 4, n_items,
                      late + 3 * (size_t)n_items, late + 3 * (size_t)n_items + part_words};
    return BHS_OK;
}

// File header of a saved model.
struct FileHead {
    char magic[8];
    uint32_t version, params;
    bhs_march_params trained_for;
    int64_t dataset_rays, adam_steps;
    float tol_phi, pad_;
    bhs::mlp::SurConst k;
};

}  // namespace

extern "C" {

bhs_mlp* bhs_mlp_create(bhs_ctx* ctx, const bhs_march_params* trained_for, uint32_t seed,
                        int* status) {
    int dummy;
    int* st = status ? status : &dummy;
    if (!ctx) {
        *st = BHS_E_INVALID;
        return nullptr;
    }
    if ((*st = check_params(ctx, trained_for)) != BHS_OK) return nullptr;
    if (trained_for->cross_all) {
        *st = fail(ctx, BHS_E_INVALID, "bhs_mlp_create: cross_all must be 0%s");
        return nullptr;
    }
    DeviceGuard g(ctx->device);
    bhs_mlp* m = new (std::nothrow) bhs_mlp;
    if (!m) {
        *st = fail(ctx, BHS_E_ALLOC, "bhs_mlp_create: out of host memory%s");
        return nullptr;
    }
    m->device = ctx->device;
    m->trained_for = *trained_for;
    m->seed = seed;
    hipError_t e;
    if ((e = hipMalloc((void**)&m->P, 4 * sizeof(float) * M::kParams)) != hipSuccess ||
        (e = hipMalloc(&m->packed, M::kPackedBytes)) != hipSuccess ||
        (e = hipMalloc((void**)&m->steps_dev, sizeof m->k.steps_max)) != hipSuccess) {
        (void)hipGetLastError();
        free_model(m);
        *st = fail(ctx, BHS_E_ALLOC, "bhs_mlp_create: device allocation failed%s");
        return nullptr;
    }
    m->G = m->P + M::kParams;
    m->m = m->G + M::kParams;
    m->v = m->m + M::kParams;
    if ((e = bhs::launch_zero(m->G, 3 * sizeof(float) * M::kParams, nullptr)) != hipSuccess ||
        (e = M::launch_init(seed, m->P, nullptr)) != hipSuccess ||
        (e = hipDeviceSynchronize()) != hipSuccess) {
        *st = hip_fail(ctx, e, "bhs_mlp_create init");
        free_model(m);
        return nullptr;
    }
    // Nothing is served until bhs_mlp_validate has built the error map.
    std::memset(m->k.serve, 0, sizeof m->k.serve);
    for (uint32_t& v : m->k.steps_max) v = 0xFFFFu;
    m->k.beta_eps = M::kBetaEps;
    for (int i = 0; i < M::kStepBins; ++i) m->k.radius_margin[i] = 0.1f;
    m->k.logit_margin = 2.0f;
    m->k.steps_slack = 1.15f;
    *st = BHS_OK;
    return m;
}

void bhs_mlp_destroy(bhs_mlp* m) { free_model(m); }

int bhs_mlp_make_dataset(bhs_ctx* ctx, bhs_mlp* m, int64_t n_rays, uint32_t seed) {
    if (!ctx || !m) return BHS_E_INVALID;
    if (n_rays < 1 || n_rays > ((int64_t)1 << 28))
        return fail(ctx, BHS_E_INVALID, "bhs_mlp_make_dataset: n_rays must be in [1, 2^28]%s");
    DeviceGuard g(ctx->device);
    hipError_t e;
    if (m->data) {
        (void)hipDeviceSynchronize();
        (void)hipFree(m->data);
        m->data = nullptr;
        m->n_data = 0;
    }
    if ((e = hipMalloc(&m->data, (size_t)n_rays * 13 * sizeof(float))) != hipSuccess) {
        m->data = nullptr;
        (void)hipGetLastError();
        return fail(ctx, BHS_E_ALLOC, "device allocation of %s failed", "training set");
    }
    RaySet rs;
    int st;
    if ((st = march_set(ctx, m, n_rays, mix(seed, 0x51u), true, &rs, nullptr)) != BHS_OK) return st;
    float* X = (float*)m->data;
    float* Y = X + 3 * n_rays;
    float* Mk = Y + 5 * n_rays;
    // pos3/dir3 of the set are not kept: the features, targets and masks are all training needs
    if ((e = bhs::launch_zero(m->steps_dev, sizeof m->k.steps_max, nullptr)) != hipSuccess ||
        (e = M::launch_targets(n_rays, rs.pos3, rs.dir3, rs.soa, X, Y, Mk, m->steps_dev, nullptr)) !=
            hipSuccess ||
        (e = hipDeviceSynchronize()) != hipSuccess)
        return hip_fail(ctx, e, "bhs_mlp_make_dataset");
    if ((e = hipMemcpy(m->k.steps_max, m->steps_dev, sizeof m->k.steps_max, hipMemcpyDeviceToHost)) !=
        hipSuccess)
        return hip_fail(ctx, e, "bhs_mlp_make_dataset copy");
    m->n_data = n_rays;
    ctx->last_stream = nullptr;
    ctx->has_last = false;
    return BHS_OK;
}

int bhs_mlp_train(bhs_ctx* ctx, bhs_mlp* m, int32_t steps, int32_t batch, float lr_start,
                  float lr_end, float* loss_host) {
    if (!ctx || !m) return BHS_E_INVALID;
    if (!m->data || m->n_data < 1)
        return fail(ctx, BHS_E_STATE, "bhs_mlp_train: no training set (bhs_mlp_make_dataset)%s");
    if (steps < 0 || batch < 1 || batch > (1 << 20) || !(lr_start >= 0.0f) || !(lr_end >= 0.0f))
        return fail(ctx, BHS_E_INVALID, "bhs_mlp_train: bad steps / batch / learning rate%s");
    if (steps == 0) return BHS_OK;
    DeviceGuard g(ctx->device);
    int st;
    if ((st = grow(ctx, &m->ws, &m->ws_bytes, M::train_workspace_floats(batch) * sizeof(float),
                   "training workspace")) != BHS_OK)
        return st;
    if ((st = grow(ctx, &m->loss, &m->loss_bytes, (size_t)steps * sizeof(float), "loss history")) !=
        BHS_OK)
        return st;
    const float* X = (const float*)m->data;
    M::TrainArgs a{};
    a.X = X;
    a.Y = X + 3 * m->n_data;
    a.M = X + 8 * m->n_data;
    a.n = m->n_data;
    a.P = m->P;
    a.G = m->G;
    a.m = m->m;
    a.v = m->v;
    a.ws = (float*)m->ws;
    a.batch = batch;
    a.ident = 0;
    a.reg_count = batch;
    a.adam = 1;
    a.b1 = kAdamB1;
    a.b2 = kAdamB2;
    a.eps = kAdamEps;
    hipError_t e;
    for (int32_t t = 0; t < steps; ++t) {
        const double ph = steps > 1 ? (double)t / (double)(steps - 1) : 1.0;
        a.lr = (float)(lr_end + 0.5 * (lr_start - lr_end) * (1.0 + std::cos(M_PI * ph)));
        ++m->adam_t;
        a.c1 = (float)(1.0 / (1.0 - std::pow((double)kAdamB1, (double)m->adam_t)));
        a.c2 = (float)(1.0 / (1.0 - std::pow((double)kAdamB2, (double)m->adam_t)));
        a.key = mix(m->seed, (uint32_t)m->adam_t);
        a.loss_dev = (float*)m->loss + t;
        if ((e = M::launch_train_step(a, nullptr)) != hipSuccess)
            return hip_fail(ctx, e, "bhs_mlp_train step");
    }
    m->packed_ok = false;
    if ((e = hipDeviceSynchronize()) != hipSuccess) return hip_fail(ctx, e, "bhs_mlp_train");
    if (loss_host &&
        (e = hipMemcpy(loss_host, m->loss, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost)) !=
            hipSuccess)
        return hip_fail(ctx, e, "bhs_mlp_train loss copy");
    return BHS_OK;
}

int bhs_mlp_grads(bhs_ctx* ctx, bhs_mlp* m, const float* x_host, const float* y_host,
                  const float* mask_host, int32_t n, float* grad_host, float* loss_host) {
    if (!ctx || !m) return BHS_E_INVALID;
    if (n < 1 || n > (1 << 20) || !x_host || !y_host || !mask_host || !grad_host)
        return fail(ctx, BHS_E_INVALID, "bhs_mlp_grads: NULL buffer or n out of [1, 2^20]%s");
    DeviceGuard g(ctx->device);
    void* buf = nullptr;
    const size_t rows = (size_t)n * 13 * sizeof(float);
    const size_t ws = M::train_workspace_floats(n) * sizeof(float);
    hipError_t e = hipMalloc(&buf, rows + ws + 256);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, BHS_E_ALLOC, "device allocation of %s failed", "gradient batch");
    }
    float* X = (float*)buf;
    float* Y = X + 3 * (size_t)n;
    float* Mk = Y + 5 * (size_t)n;
    float* W = Mk + 5 * (size_t)n;
    float* L = (float*)((char*)W + ws);
    M::TrainArgs a{};
    a.X = X;
    a.Y = Y;
    a.M = Mk;
    a.n = n;
    a.P = m->P;
    a.G = m->G;
    a.m = m->m;
    a.v = m->v;
    a.ws = W;
    a.batch = n;
    a.ident = 1;
    a.reg_count = n;
    a.adam = 0;
    a.loss_dev = L;
    float loss = 0.0f;
    if ((e = hipMemcpy(X, x_host, 3 * sizeof(float) * n, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(Y, y_host, 5 * sizeof(float) * n, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(Mk, mask_host, 5 * sizeof(float) * n, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = M::launch_train_step(a, nullptr)) != hipSuccess ||
        (e = hipDeviceSynchronize()) != hipSuccess ||
        (e = hipMemcpy(grad_host, m->G, sizeof(float) * M::kParams, hipMemcpyDeviceToHost)) !=
            hipSuccess ||
        (e = hipMemcpy(&loss, L, sizeof(float), hipMemcpyDeviceToHost)) != hipSuccess) {
        (void)hipFree(buf);
        return hip_fail(ctx, e, "bhs_mlp_grads");
    }
    (void)hipFree(buf);
    if (loss_host) *loss_host = loss;
    return BHS_OK;
}

int bhs_mlp_forward(bhs_ctx* ctx, bhs_mlp* m, const float* x_dev, int64_t n, float* y_dev,
                    void* stream) {
    if (!ctx || !m) return BHS_E_INVALID;
    if (n < 0 || (n > 0 && (!x_dev || !y_dev)))
        return fail(ctx, BHS_E_INVALID, "bhs_mlp_forward: NULL buffer or n < 0%s");
    DeviceGuard g(ctx->device);
    hipStream_t s = (hipStream_t)stream;
    if (int os = order_stream(ctx, s)) return os;
    int st;
    if ((st = pack(ctx, m, s)) != BHS_OK) return st;
    if (n == 0) return BHS_OK;
    hipError_t e = M::launch_forward(x_dev, n, m->packed, y_dev, s);
    if (e != hipSuccess) return hip_fail(ctx, e, "bhs_mlp_forward");
    note_stream(ctx, s);
    return BHS_OK;
}

int bhs_mlp_validate(bhs_ctx* ctx, bhs_mlp* m, int64_t n_rays, uint32_t seed, float tol_phi,
                     bhs_mlp_validation* out_host) {
    if (!ctx || !m) return BHS_E_INVALID;
    if (n_rays < 1 || n_rays > ((int64_t)1 << 26) || !(tol_phi > 0.0f))
        return fail(ctx, BHS_E_INVALID, "bhs_mlp_validate: n_rays in [1, 2^26], tol_phi > 0%s");
    DeviceGuard g(ctx->device);
    RaySet rs;
    int st;
    // The reference semantics: chord crossings, so the edge margins cover the march's own chord
    // error as well as the surrogate's.
    if ((st = march_set(ctx, m, n_rays, mix(seed, 0xA5u), false, &rs, nullptr)) != BHS_OK) return st;
    if ((st = pack(ctx, m, nullptr)) != BHS_OK) return st;
    void* buf = nullptr;
    hipError_t e = hipMalloc(&buf, (size_t)n_rays * 12);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, BHS_E_ALLOC, "device allocation of %s failed", "validation errors");
    }
    float* ep_d = (float*)buf;
    float* er_d = ep_d + n_rays;
    int32_t* ci_d = (int32_t*)(er_d + n_rays);
    std::vector<float> ep(n_rays), er(n_rays);
    std::vector<int32_t> ci(n_rays);
    bhs_march_params rp = m->trained_for;
    rp.cross_all = 0;
    const bhs::MarchConst c = make_const(&rp);
    M::SurConst k = m->k;
    for (int i = 0; i < M::kMapWords; ++i) k.serve[i] = 0xFFFFFFFFu;
    auto run = [&]() -> int {
        hipError_t e2;
        if ((e2 = M::launch_validate(n_rays, rs.pos3, rs.dir3, c, k, rs.soa, m->packed, ep_d, er_d,
                                     ci_d, ctx->num_cus, nullptr)) != hipSuccess ||
            (e2 = hipDeviceSynchronize()) != hipSuccess ||
            (e2 = hipMemcpy(ep.data(), ep_d, n_rays * 4, hipMemcpyDeviceToHost)) != hipSuccess ||
            (e2 = hipMemcpy(er.data(), er_d, n_rays * 4, hipMemcpyDeviceToHost)) != hipSuccess ||
            (e2 = hipMemcpy(ci.data(), ci_d, n_rays * 4, hipMemcpyDeviceToHost)) != hipSuccess)
            return hip_fail(ctx, e2, "bhs_mlp_validate");
        return BHS_OK;
    };
    // Pass 1: radius errors near the annulus decide the edge margin of each log r_obs bin (the
    // largest error of the bin and its neighbours, x1.25, within [1e-3, 0.1]).
    if ((st = run()) != BHS_OK) {
        (void)hipFree(buf);
        return st;
    }
    std::vector<float> pos(3 * n_rays);
    if ((e = hipMemcpy(pos.data(), rs.pos3, 12 * n_rays, hipMemcpyDeviceToHost)) != hipSuccess) {
        (void)hipFree(buf);
        return hip_fail(ctx, e, "bhs_mlp_validate copy");
    }
    float bin_max[M::kStepBins] = {};
    for (int64_t i = 0; i < n_rays; ++i) {
        if (!(er[i] >= 0.0f && er[i] < 1e29f)) continue;
        const double r = std::sqrt((double)pos[3 * i] * pos[3 * i] + (double)pos[3 * i + 1] * pos[3 * i + 1] +
                                   (double)pos[3 * i + 2] * pos[3 * i + 2]);
        int bin = (int)((std::log(r) - M::kLogRLo) * (M::kStepBins / (M::kLogRHi - M::kLogRLo)));
        bin = std::max(0, std::min(M::kStepBins - 1, bin));
        bin_max[bin] = std::max(bin_max[bin], er[i]);
    }
    float margin = 0.0f;
    for (int i = 0; i < M::kStepBins; ++i) {
        float mx = bin_max[i];
        if (i > 0) mx = std::max(mx, bin_max[i - 1]);
        if (i + 1 < M::kStepBins) mx = std::max(mx, bin_max[i + 1]);
        k.radius_margin[i] = std::min(0.1f, std::max(1e-3f, 1.25f * (mx > 0.0f ? mx : 0.08f)));
        margin = std::max(margin, k.radius_margin[i]);
    }
    // Pass 2: per-cell maxima with the final margin.
    if ((st = run()) != BHS_OK) {
        (void)hipFree(buf);
        return st;
    }
    (void)hipFree(buf);
    const int cells = M::kMapR * M::kMapB;
    std::vector<float> cmax(cells, 0.0f);
    std::vector<int64_t> ccnt(cells, 0);
    int64_t in_dom = 0;
    for (int64_t i = 0; i < n_rays; ++i) {
        if (ci[i] < 0) continue;
        ccnt[ci[i]]++;
        cmax[ci[i]] = std::max(cmax[ci[i]], ep[i]);
    }
    uint32_t serve[M::kMapWords] = {};
    int served_cells = 0;
    for (int c2 = 0; c2 < cells; ++c2)
        if (ccnt[c2] >= 8 && cmax[c2] <= tol_phi) {
            serve[c2 >> 5] |= 1u << (c2 & 31);
            ++served_cells;
        }
    bhs_mlp_validation v{};
    std::vector<float> pe, re;
    for (int64_t i = 0; i < n_rays; ++i) {
        if (ep[i] >= 0.0f) ++in_dom;
        if (ci[i] < 0 || !((serve[ci[i] >> 5] >> (ci[i] & 31)) & 1u)) continue;
        ++v.served;
        if (ep[i] >= 1e29f) {
            ++v.mismatches;
            continue;
        }
        if (er[i] >= 0.0f) re.push_back(er[i]);
        if (ep[i] > 0.0f) {
            ++v.escaping;
            pe.push_back(ep[i]);
        }
    }
    v.rays = n_rays;
    v.in_domain = in_dom;
    v.phi_err_max = pe.empty() ? 0.0f : quantile(pe, 1.0);
    v.phi_err_p99 = quantile(pe, 0.99);
    v.phi_err_p50 = quantile(pe, 0.5);
    v.r_err_max = re.empty() ? 0.0f : quantile(re, 1.0);
    v.r_err_p99 = quantile(re, 0.99);
    v.r_err_p50 = quantile(re, 0.5);
    v.radius_margin = margin;
    v.served_cells = served_cells;
    std::memcpy(m->k.serve, serve, sizeof serve);
    std::memcpy(m->k.radius_margin, k.radius_margin, sizeof k.radius_margin);
    m->tol_phi = tol_phi;
    if (out_host) *out_host = v;
    ctx->last_stream = nullptr;
    ctx->has_last = false;
    return BHS_OK;
}

int bhs_mlp_get_params(bhs_ctx* ctx, bhs_mlp* m, float* host, int64_t n) {
    if (!ctx || !m) return BHS_E_INVALID;
    if (!host || n != M::kParams)
        return fail(ctx, BHS_E_INVALID, "bhs_mlp_get_params: need BHS_MLP_PARAMS floats%s");
    DeviceGuard g(ctx->device);
    hipError_t e;
    if ((e = hipDeviceSynchronize()) != hipSuccess ||
        (e = hipMemcpy(host, m->P, sizeof(float) * M::kParams, hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_mlp_get_params");
    return BHS_OK;
}

int bhs_mlp_set_params(bhs_ctx* ctx, bhs_mlp* m, const float* host, int64_t n) {
    if (!ctx || !m) return BHS_E_INVALID;
    if (!host || n != M::kParams)
        return fail(ctx, BHS_E_INVALID, "bhs_mlp_set_params: need BHS_MLP_PARAMS floats%s");
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(host[i]))
            return fail(ctx, BHS_E_INVALID, "bhs_mlp_set_params: non-finite parameter%s");
    DeviceGuard g(ctx->device);
    hipError_t e;
    if ((e = hipDeviceSynchronize()) != hipSuccess ||
        (e = hipMemcpy(m->P, host, sizeof(float) * M::kParams, hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_mlp_set_params");
    m->packed_ok = false;
    return BHS_OK;
}

int bhs_mlp_get_info(bhs_ctx* ctx, bhs_mlp* m, bhs_mlp_info* out_host) {
    if (!ctx || !m || !out_host) return BHS_E_INVALID;
    bhs_mlp_info i{};
    i.trained_for = m->trained_for;
    i.dataset_rays = m->n_data;
    i.adam_steps = m->adam_t;
    i.tol_phi = m->tol_phi;
    i.radius_margin = *std::max_element(m->k.radius_margin, m->k.radius_margin + M::kStepBins);
    i.logit_margin = m->k.logit_margin;
    i.steps_slack = m->k.steps_slack;
    for (int r = 0; r < M::kStepBins; ++r) {   // per log r_obs bin: the largest over x2
        uint32_t mx = 0;
        for (int b = 0; b < M::kStepBinsB; ++b)
            mx = std::max(mx, m->k.steps_max[r * M::kStepBinsB + b]);
        i.steps_max[r] = mx;
    }
    int cells = 0;
    for (int w = 0; w < M::kMapWords; ++w) cells += __builtin_popcount(m->k.serve[w]);
    i.served_cells = cells;
    *out_host = i;
    return BHS_OK;
}

int bhs_mlp_save(bhs_ctx* ctx, bhs_mlp* m, const char* path) {
    if (!ctx || !m) return BHS_E_INVALID;
    if (!path) return fail(ctx, BHS_E_INVALID, "bhs_mlp_save: path is NULL%s");
    std::vector<float> p(M::kParams);
    int st;
    if ((st = bhs_mlp_get_params(ctx, m, p.data(), M::kParams)) != BHS_OK) return st;
    FileHead h{};
    std::memcpy(h.magic, "BHSMLP04", 8);
    h.version = 4;
    h.params = M::kParams;
    h.trained_for = m->trained_for;
    h.dataset_rays = m->n_data;
    h.adam_steps = m->adam_t;
    h.tol_phi = m->tol_phi;
    h.k = m->k;
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(ctx, BHS_E_IO, "bhs_mlp_save: cannot open %s", path);
    const bool ok = std::fwrite(&h, sizeof h, 1, f) == 1 &&
                    std::fwrite(p.data(), sizeof(float), p.size(), f) == p.size();
    if (std::fclose(f) != 0 || !ok) return fail(ctx, BHS_E_IO, "bhs_mlp_save: write failed: %s", path);
    return BHS_OK;
}

bhs_mlp* bhs_mlp_load(bhs_ctx* ctx, const char* path, int* status) {
    int dummy;
    int* st = status ? status : &dummy;
    if (!ctx || !path) {
        *st = ctx ? fail(ctx, BHS_E_INVALID, "bhs_mlp_load: path is NULL%s") : BHS_E_INVALID;
        return nullptr;
    }
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        *st = fail(ctx, BHS_E_IO, "bhs_mlp_load: cannot open %s", path);
        return nullptr;
    }
    FileHead h{};
    std::vector<float> p(M::kParams);
    const bool ok = std::fread(&h, sizeof h, 1, f) == 1 &&
                    std::fread(p.data(), sizeof(float), p.size(), f) == p.size();
    std::fclose(f);
    if (!ok || std::memcmp(h.magic, "BHSMLP04", 8) != 0 || h.version != 4 ||
        h.params != (uint32_t)M::kParams) {
        *st = fail(ctx, BHS_E_IO, "bhs_mlp_load: not a model file of this version: %s", path);
        return nullptr;
    }
    bhs_mlp* m = bhs_mlp_create(ctx, &h.trained_for, 0u, st);
    if (!m) return nullptr;
    if ((*st = bhs_mlp_set_params(ctx, m, p.data(), M::kParams)) != BHS_OK) {
        free_model(m);
        return nullptr;
    }
    m->adam_t = h.adam_steps;
    m->tol_phi = h.tol_phi;
    m->k = h.k;
    return m;
}

int bhs_mlp_set_table(bhs_ctx* ctx, bhs_mlp* m, int32_t n2, int32_t n3) {
    if (!ctx || !m) return BHS_E_INVALID;
    if (n2 == 0) {
        m->tab_n2 = m->tab_n3 = 0;
        return BHS_OK;
    }
    if (n2 < 2 || n3 < 2 || n2 > 16384 || n3 > 16384 || (int64_t)n2 * n3 > ((int64_t)1 << 24))
        return fail(ctx, BHS_E_INVALID,
                    "bhs_mlp_set_table: n2, n3 in [2, 16384] with n2 * n3 <= 2^24 (or n2 = 0)%s");
    m->tab_n2 = n2;
    m->tab_n3 = n3;
    return BHS_OK;
}

int bhs_render_surrogate(bhs_ctx* ctx, bhs_mlp* m, const bhs_camera* cam,
                         const bhs_march_params* params, const bhs_ray_soa* out, void* stream) {
    if (!ctx || !m) return BHS_E_INVALID;
    int st;
    if ((st = check_camera(ctx, cam)) != BHS_OK) return st;
    if ((st = check_params(ctx, params)) != BHS_OK) return st;
    if (!out) return fail(ctx, BHS_E_INVALID, "bhs_render_surrogate: out is NULL%s");
    if (params->cross_all)
        return fail(ctx, BHS_E_INVALID, "bhs_render_surrogate: cross_all must be 0%s");
    if (!same_march(*params, m->trained_for))
        return fail(ctx, BHS_E_INVALID,
                    "bhs_render_surrogate: the model was trained for other step-control / disk "
                    "parameters%s");
    if (params->root64 && (!out->n_cross_dev || !out->flags_dev || !out->disk_dev))
        return fail(ctx, BHS_E_INVALID,
                    "bhs_render_surrogate: root64 needs n_cross, flags and disk outputs%s");
    DeviceGuard g(ctx->device);
    const int64_t n = (int64_t)cam->width * cam->height;
    const uint32_t n_items = bhs::camera_items(cam->width, cam->height);
    bhs::MarchConst c = make_const(params);
    fill_camera(c, cam);
    c.ray_offset = 0;
    const bool roots = params->root64 && params->precision != BHS_PREC_F64;
    bhs::Scratch sc;
    if ((st = prepare_scratch(ctx, n, n_items, roots, &sc)) != BHS_OK) return st;
    M::SurLists fl;
    if ((st = lists_for(ctx, m, n_items, &fl)) != BHS_OK) return st;
    const M::SurStreams ss{ctx->aux,      ctx->aux2,    ctx->ev_fork, ctx->ev_join,
                           ctx->ev_fork2, ctx->ev_join2, ctx->ev_tab0, ctx->ev_tab1};
    M::SurTable tab{};
    if ((st = table_for(ctx, m, 1, &tab)) != BHS_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    if (int os = order_stream(ctx, s)) return os;
    if ((st = pack(ctx, m, s)) != BHS_OK) return st;
    hipError_t e;
    bhs_ctx::TimingRec* tr = next_timer(ctx);
    if (tr) (void)hipEventRecord(tr->call0, s);
    if ((e = bhs::launch_zero(ctx->scratch, kStatsBytes, s)) != hipSuccess ||
        (e = bhs::launch_zero(m->fb, 256, s)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_render_surrogate clear");
    if ((e = M::launch_surrogate(c, m->k, m->packed, tab.n2 > 0 ? &tab : nullptr,
                                 make_soa(out, n), sc, fl, ctx->tuning, ctx->num_cus, s, ss,
                                 tr ? tr->k0 : nullptr, tr ? tr->k1 : nullptr, nullptr)) !=
        hipSuccess)
        return hip_fail(ctx, e, "bhs_render_surrogate launch");
    if (tr) (void)hipEventRecord(tr->call1, s);
    note_stream(ctx, s);
    return BHS_OK;
}

int bhs_render_surrogate_batch(bhs_ctx* ctx, bhs_mlp* m, const bhs_camera* cams,
                               const bhs_march_params* params, int32_t ncam,
                               const bhs_ray_soa* out, void* stream) {
    if (!ctx || !m) return BHS_E_INVALID;
    if (!out) return fail(ctx, BHS_E_INVALID, "bhs_render_surrogate_batch: out is NULL%s");
    BatchTable t;
    int st;
    if ((st = build_batch(ctx, cams, params, ncam, &t)) != BHS_OK) return st;
    for (int32_t k = 0; k < ncam; ++k) {
        if (params[k].cross_all)
            return fail(ctx, BHS_E_INVALID, "bhs_render_surrogate_batch: cross_all must be 0%s");
        if (!same_march(params[k], m->trained_for))
            return fail(ctx, BHS_E_INVALID,
                        "bhs_render_surrogate_batch: the model was trained for other step-control "
                        "/ disk parameters%s");
    }
    if (t.any_root && (!out->n_cross_dev || !out->flags_dev || !out->disk_dev))
        return fail(ctx, BHS_E_INVALID,
                    "bhs_render_surrogate_batch: root64 needs n_cross, flags and disk outputs%s");
    DeviceGuard g(ctx->device);
    hipStream_t s = (hipStream_t)stream;
    if (int os = order_stream(ctx, s)) return os;
    void* table_dev = nullptr;
    if ((st = upload_batch(ctx, t, s, &table_dev)) != BHS_OK) return st;
    bhs::Scratch sc;
    if ((st = prepare_scratch(ctx, t.n_rays, t.n_items, t.any_root, &sc)) != BHS_OK) return st;
    M::SurLists fl;
    if ((st = lists_for(ctx, m, t.n_items, &fl)) != BHS_OK) return st;
    M::SurTable tab;
    if ((st = table_for(ctx, m, ncam, &tab)) != BHS_OK) return st;
    const M::SurStreams ss{ctx->aux,      ctx->aux2,    ctx->ev_fork, ctx->ev_join,
                           ctx->ev_fork2, ctx->ev_join2, ctx->ev_tab0, ctx->ev_tab1};
    if ((st = pack(ctx, m, s)) != BHS_OK) return st;
    hipError_t e;
    bhs_ctx::TimingRec* tr = next_timer(ctx);
    if (tr) (void)hipEventRecord(tr->call0, s);
    if ((e = bhs::launch_zero(ctx->scratch, kStatsBytes, s)) != hipSuccess ||
        (e = bhs::launch_zero(m->fb, 256, s)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_render_surrogate_batch clear");
    const bhs::MarchConst* table = (const bhs::MarchConst*)table_dev;
    const uint32_t* base = (const uint32_t*)((char*)table_dev + table_cams_bytes(ncam));
    if ((e = M::launch_surrogate_batch(table, base, (uint32_t)ncam, t.n_items, t.any_band, m->k,
                                       m->packed, tab.n2 > 0 ? &tab : nullptr,
                                       make_soa(out, t.n_rays), sc, fl, ctx->tuning,
                                       ctx->num_cus, s, ss, tr ? tr->k0 : nullptr,
                                       tr ? tr->k1 : nullptr, nullptr)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_render_surrogate_batch launch");
    if (tr) (void)hipEventRecord(tr->call1, s);
    note_stream(ctx, s);
    return BHS_OK;
}

}  // extern "C"
