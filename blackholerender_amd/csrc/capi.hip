// capi.hip — the C-ABI of include/bhs_trace.h: context, argument checking, scratch management and
// stream ordering around the kernels of kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/bhs_trace.h"
#include "capi_internal.hpp"
#include "footprint.hpp"
#include "gas_lens.hpp"
#include "kernels.hpp"

// A captured batch launch (bhs_graph_*): owns its camera table and scratch so later calls on the
// context cannot move them, and its own capture/fork streams and events.
struct bhs_graph {
    int device = 0;
    void* mem = nullptr;          // [table | scratch]
    size_t bytes = 0;
    int32_t ncam = 0;
    int64_t n_rays = 0;
    uint32_t n_items = 0;
    std::vector<int32_t> dims;    // width, height per camera (updates must keep them)
    bool any_band = false, any_root = false;
    hipStream_t cap = nullptr, aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace bhs {
namespace capi {

int fail(bhs_ctx* ctx, int code, const char* fmt, const char* detail) {
    if (ctx) {
        char buf[512];
        std::snprintf(buf, sizeof buf, fmt, detail ? detail : "");
        ctx->err = buf;
    }
    return code;
}

int hip_fail(bhs_ctx* ctx, hipError_t e, const char* where) {
    if (ctx) ctx->err = std::string(where) + ": " + hipGetErrorString(e);
    return (e == hipErrorOutOfMemory) ? BHS_E_ALLOC : BHS_E_DEVICE;
}

// Grows a device buffer, after draining work that may still use the old one.
int ensure(bhs_ctx* ctx, void** buf, size_t* cap, size_t need, const char* what) {
    if (*cap >= need) return BHS_OK;
    if (*buf) {
        (void)hipDeviceSynchronize();
        (void)hipFree(*buf);
        *buf = nullptr;
        *cap = 0;
    }
    size_t want = need + need / 4 + 4096;
    hipError_t e = hipMalloc(buf, want);
    if (e != hipSuccess) {
        *buf = nullptr;
        (void)hipGetLastError();
        return fail(ctx, BHS_E_ALLOC, "device allocation of %s failed", what);
    }
    *cap = want;
    return BHS_OK;
}

bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

int check_params(bhs_ctx* ctx, const bhs_march_params* p) {
    if (!p) return fail(ctx, BHS_E_INVALID, "params is NULL%s");
    if (p->max_steps < 0) return fail(ctx, BHS_E_INVALID, "max_steps < 0%s");
    if (p->disk_mode != BHS_DISK_NONE && p->disk_mode != BHS_DISK_THIN)
        return fail(ctx, BHS_E_INVALID, "disk_mode must be BHS_DISK_NONE or BHS_DISK_THIN%s");
    if (p->precision < BHS_PREC_F32 || p->precision > BHS_PREC_F64)
        return fail(ctx, BHS_E_INVALID, "unknown precision%s");
    if (!(p->step_scale > 0.0f) || !std::isfinite(p->step_scale))
        return fail(ctx, BHS_E_INVALID, "step_scale must be finite and > 0%s");
    if (!std::isfinite(p->disk_step) || !std::isfinite(p->disk_inner) ||
        !std::isfinite(p->disk_outer) || !std::isfinite(p->disk_thickness) ||
        !std::isfinite(p->band_rel) || p->band_rel < 0.0f)
        return fail(ctx, BHS_E_INVALID, "disk/band parameters must be finite (band_rel >= 0)%s");
    return BHS_OK;
}

bhs::MarchConst make_const(const bhs_march_params* p) {
    bhs::MarchConst c;
    std::memset(&c, 0, sizeof c);
    c.max_steps = p->max_steps;
    c.thin = p->disk_mode == BHS_DISK_THIN;
    c.step_scale = p->step_scale;
    c.disk_step = p->disk_step;
    c.disk_inner = p->disk_inner;
    c.disk_outer = p->disk_outer;
    c.disk_thickness = p->disk_thickness;
    c.band_abs = p->band_rel * bhs::kBcritF;
    c.precision = p->precision;
    c.root64 = p->root64 ? 1 : 0;
    c.cross_all = p->cross_all ? 1 : 0;
    return c;
}

uint32_t pcg_host(uint32_t v) {
    const uint32_t state = v * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

// Seed of the hashed per-pixel jitter: pcg chain over (sample, camera, seed high, seed low).
uint32_t jitter_base(uint64_t seed, uint32_t camera_id, uint32_t sample) {
    uint32_t h = pcg_host(sample);
    h = pcg_host(h ^ camera_id);
    h = pcg_host(h ^ (uint32_t)(seed >> 32));
    h = pcg_host(h ^ (uint32_t)seed);
    return h;
}

int check_camera(bhs_ctx* ctx, const bhs_camera* cam) {
    if (!cam) return fail(ctx, BHS_E_INVALID, "camera is NULL%s");
    if (cam->width <= 0 || cam->height <= 0)
        return fail(ctx, BHS_E_INVALID, "camera width and height must be > 0%s");
    if ((int64_t)cam->width * cam->height > (int64_t)0x7fffffff)
        return fail(ctx, BHS_E_INVALID, "camera has more than 2^31-1 pixels%s");
    if (!finite3(cam->pos) || !std::isfinite(cam->tan_half_fov) || !(cam->tan_half_fov > 0.0f))
        return fail(ctx, BHS_E_INVALID, "camera position / tan_half_fov must be finite (tan > 0)%s");
    for (int k = 0; k < 9; ++k)
        if (!std::isfinite(cam->basis[k])) return fail(ctx, BHS_E_INVALID, "camera basis not finite%s");
    const float r2 = cam->pos[0] * cam->pos[0] + cam->pos[1] * cam->pos[1] + cam->pos[2] * cam->pos[2];
    if (!(r2 > 0.0f)) return fail(ctx, BHS_E_INVALID, "camera at the origin%s");
    if (!std::isfinite(cam->jitter[0]) || !std::isfinite(cam->jitter[1]))
        return fail(ctx, BHS_E_INVALID, "camera jitter not finite%s");
    return BHS_OK;
}

void fill_camera(bhs::MarchConst& c, const bhs_camera* cam) {
    for (int k = 0; k < 3; ++k) c.pos[k] = cam->pos[k];
    for (int k = 0; k < 9; ++k) c.basis[k] = cam->basis[k];
    c.tan_half_fov = cam->tan_half_fov;
    c.width = cam->width;
    c.height = cam->height;
    c.jitter[0] = cam->jitter[0];
    c.jitter[1] = cam->jitter[1];
    c.jitter_hashed = cam->jitter_hashed ? 1 : 0;
    c.jitter_base = jitter_base(cam->seed, cam->camera_id, cam->sample);
}

bhs::SoA make_soa(const bhs_ray_soa* o, int64_t n) {
    bhs::SoA s;
    s.cls = o->cls_dev;
    s.flags = o->flags_dev;
    s.n_cross = o->n_cross_dev;
    s.steps = o->steps_dev;
    s.esc = o->esc_dev;
    s.disk = o->disk_dev;
    s.b = o->b_dev;
    s.n = n;
    return s;
}

// Next timer record of the context (null when timing is off or events cannot be created).
bhs_ctx::TimingRec* next_timer(bhs_ctx* ctx) {
    if (!ctx->timing) return nullptr;
    if (ctx->timers_used == ctx->timers.size()) {
        bhs_ctx::TimingRec r{};
        if (hipEventCreate(&r.call0) != hipSuccess || hipEventCreate(&r.k0) != hipSuccess ||
            hipEventCreate(&r.k1) != hipSuccess || hipEventCreate(&r.call1) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        ctx->timers.push_back(r);
    }
    return &ctx->timers[ctx->timers_used++];
}

int order_stream(bhs_ctx* ctx, hipStream_t s) {
    if (!ctx->has_last || ctx->last_stream == s) return BHS_OK;
    hipError_t e;
    if ((e = hipEventRecord(ctx->ev_order, ctx->last_stream)) != hipSuccess ||
        (e = hipStreamWaitEvent(s, ctx->ev_order, 0)) != hipSuccess)
        return hip_fail(ctx, e, "stream ordering");
    return BHS_OK;
}

// Render scratch for n_rays rays and n_items work items: [Stats | band count | work counter]
// (kStatsBytes, cleared per launch) [band ids] [root states, 2 per ray].
size_t scratch_bytes(int64_t n_rays, int64_t n_items, bool roots) {
    const size_t band_bytes = ((size_t)n_items * 4 + 255) & ~(size_t)255;
    const size_t root_bytes = roots ? (size_t)n_rays * 2 * sizeof(bhs::RootState<float>) : 0;
    return kStatsBytes + band_bytes + root_bytes;
}

void layout_scratch(void* mem, int64_t n_items, bool roots, bhs::Scratch* sc) {
    const size_t band_bytes = ((size_t)n_items * 4 + 255) & ~(size_t)255;
    char* base = (char*)mem;
    sc->stats = (bhs::Stats*)base;
    sc->band_count = (uint32_t*)(base + sizeof(bhs::Stats));
    sc->work = (uint32_t*)(base + sizeof(bhs::Stats) + 4);
    sc->band_work = (uint32_t*)(base + sizeof(bhs::Stats) + 8);
    sc->seg_done = (unsigned long long*)(base + sizeof(bhs::Stats) + 16);
    sc->seg = (uint32_t*)(base + 256);
    sc->band_ids = (uint32_t*)(base + kStatsBytes);
    sc->band_cap = (uint32_t)n_items;
    sc->roots = roots ? (bhs::RootState<float>*)(base + kStatsBytes + band_bytes) : nullptr;
}

int prepare_scratch(bhs_ctx* ctx, int64_t n_rays, int64_t n_items, bool roots, bhs::Scratch* sc) {
    int st = ensure(ctx, &ctx->scratch, &ctx->scratch_bytes, scratch_bytes(n_rays, n_items, roots),
                    "render scratch");
    if (st != BHS_OK) return st;
    layout_scratch(ctx->scratch, n_items, roots, sc);
    ctx->stats_src = nullptr;
    return BHS_OK;
}

// Host-side camera table of a batch: constants, output offsets (prefix sums of width*height) and
// item prefix sums. Checks every camera and its params.
int build_batch(bhs_ctx* ctx, const bhs_camera* cams, const bhs_march_params* params, int32_t ncam,
                BatchTable* t) {
    if (!cams || !params) return fail(ctx, BHS_E_INVALID, "batch: cameras or params is NULL%s");
    if (ncam < 1 || (uint32_t)ncam > bhs::kMaxBatchCams)
        return fail(ctx, BHS_E_INVALID, "batch: ncam must be in 1..64%s");
    t->cams.resize((size_t)ncam);
    t->item_base.assign((size_t)ncam + 1, 0u);
    uint64_t items = 0;
    int64_t rays = 0;
    for (int32_t k = 0; k < ncam; ++k) {
        int st;
        if ((st = check_camera(ctx, &cams[k])) != BHS_OK) return st;
        if ((st = check_params(ctx, &params[k])) != BHS_OK) return st;
        if (params[k].precision == BHS_PREC_F64)
            return fail(ctx, BHS_E_INVALID, "batch: BHS_PREC_F64 is not supported in batches%s");
        bhs::MarchConst c = make_const(&params[k]);
        fill_camera(c, &cams[k]);
        c.ray_offset = rays;
        t->cams[(size_t)k] = c;
        t->item_base[(size_t)k] = (uint32_t)items;
        items += bhs::camera_items(cams[k].width, cams[k].height);
        rays += (int64_t)cams[k].width * cams[k].height;
        if (items > 0xFFFF0000ull) return fail(ctx, BHS_E_INVALID, "batch: more than 2^32 rays%s");
        t->any_band = t->any_band || params[k].precision == BHS_PREC_F32_BAND64;
        t->any_root = t->any_root || params[k].root64 != 0;
    }
    t->item_base[(size_t)ncam] = (uint32_t)items;
    t->n_items = (uint32_t)items;
    t->n_rays = rays;
    return BHS_OK;
}

size_t table_cams_bytes(int32_t ncam) {
    return ((size_t)ncam * sizeof(bhs::MarchConst) + 255) & ~(size_t)255;
}
size_t table_bytes(int32_t ncam) {
    return table_cams_bytes(ncam) + ((((size_t)ncam + 1) * sizeof(uint32_t) + 255) & ~(size_t)255);
}

// Copies a batch table into one host buffer laid out as on the device.
void pack_table(const BatchTable& t, char* dst) {
    const int32_t ncam = (int32_t)t.cams.size();
    std::memcpy(dst, t.cams.data(), t.cams.size() * sizeof(bhs::MarchConst));
    std::memcpy(dst + table_cams_bytes(ncam), t.item_base.data(), t.item_base.size() * sizeof(uint32_t));
}

int upload_batch(bhs_ctx* ctx, const BatchTable& t, hipStream_t s, void** table_dev) {
    const int32_t ncam = (int32_t)t.cams.size();
    const size_t tb = table_bytes(ncam);
    int st;
    if ((st = ensure(ctx, &ctx->table_dev, &ctx->table_dev_bytes, tb, "camera table")) != BHS_OK)
        return st;
    if (ctx->table_host_bytes < tb) {
        if (ctx->upload_pending) (void)hipEventSynchronize(ctx->ev_upload);
        if (ctx->table_host) (void)hipHostFree(ctx->table_host);
        ctx->table_host = nullptr;
        ctx->table_host_bytes = 0;
        if (hipHostMalloc(&ctx->table_host, tb, 0) != hipSuccess) {
            (void)hipGetLastError();
            ctx->table_host = nullptr;
            return fail(ctx, BHS_E_ALLOC, "pinned allocation of the camera table failed%s");
        }
        ctx->table_host_bytes = tb;
    }
    if (ctx->upload_pending) (void)hipEventSynchronize(ctx->ev_upload);   // staging buffer is free
    pack_table(t, (char*)ctx->table_host);
    hipError_t e;
    if ((e = hipMemcpyAsync(ctx->table_dev, ctx->table_host, tb, hipMemcpyHostToDevice, s)) !=
            hipSuccess ||
        (e = hipEventRecord(ctx->ev_upload, s)) != hipSuccess)
        return hip_fail(ctx, e, "camera table upload");
    ctx->upload_pending = true;
    *table_dev = ctx->table_dev;
    return BHS_OK;
}

}  // namespace capi
}  // namespace bhs

// ------------------------------------------------------------------ static scratch (no-alloc render)
// Scratch of bhs_render_noalloc: stats and counters, the FP64 band list and the root states of one
// render, in a device array of fixed size instead of memory the context allocates. Each GPU holds
// its own copy at its own address; every no-alloc call on that GPU shares it and the array keeps
// the last call's contents (the counters are cleared by a kernel at the start of each call).
constexpr size_t kNoAllocScratchBytes = BHS_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};   // device address of `scratch`, per device

// Device address of `scratch` on `device` (which must be current), resolved once per device. The
// array's name is a host-side stand-in: only this address may reach kernels and copies.
int resolve_static_scratch(int device, void** ptr) {
    if (device < 0 || device >= kMaxDevices) return BHS_E_DEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return BHS_E_DEVICE;
        }
        g_scratch_dev[device] = p;
    }
    *ptr = g_scratch_dev[device];
    return BHS_OK;
}
}  // namespace

using namespace bhs::capi;

extern "C" {

int bhs_version(void) { return BHS_ABI_VERSION; }

bhs_ctx* bhs_create(int device, int* status) {
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0 || device < 0 || device >= ndev) {
        (void)hipGetLastError();
        if (status) *status = BHS_E_DEVICE;
        return nullptr;
    }
    bhs_ctx* ctx = new (std::nothrow) bhs_ctx();
    if (!ctx) {
        if (status) *status = BHS_E_ALLOC;
        return nullptr;
    }
    ctx->device = device;
    DeviceGuard g(device);
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        cus > 0)
        ctx->num_cus = cus;
    if (hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->aux2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_fork2, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_join2, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_tab0, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_tab1, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_upload, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_order, hipEventDisableTiming) != hipSuccess ||
        resolve_static_scratch(device, &ctx->static_scratch) != BHS_OK) {
        (void)hipGetLastError();
        bhs_destroy(ctx);
        if (status) *status = BHS_E_DEVICE;
        return nullptr;
    }
    if (status) *status = BHS_OK;
    return ctx;
}

void bhs_destroy(bhs_ctx* ctx) {
    if (!ctx) return;
    {
        DeviceGuard g(ctx->device);
        (void)hipDeviceSynchronize();
        if (ctx->scratch) (void)hipFree(ctx->scratch);
        if (ctx->tstage) (void)hipFree(ctx->tstage);
        for (auto& r : ctx->timers) {
            (void)hipEventDestroy(r.call0);
            (void)hipEventDestroy(r.k0);
            (void)hipEventDestroy(r.k1);
            (void)hipEventDestroy(r.call1);
        }
        if (ctx->table_dev) (void)hipFree(ctx->table_dev);
        if (ctx->noise_vol) (void)hipFree(ctx->noise_vol);
        if (ctx->bb_lut) (void)hipFree(ctx->bb_lut);
        if (ctx->shade_list) (void)hipFree(ctx->shade_list);
        if (ctx->table_host) (void)hipHostFree(ctx->table_host);
        if (ctx->ev_upload) (void)hipEventDestroy(ctx->ev_upload);
        if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
        if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
        if (ctx->ev_order) (void)hipEventDestroy(ctx->ev_order);
        if (ctx->aux) (void)hipStreamDestroy(ctx->aux);
        if (ctx->ev_fork2) (void)hipEventDestroy(ctx->ev_fork2);
        if (ctx->ev_join2) (void)hipEventDestroy(ctx->ev_join2);
        if (ctx->ev_tab0) (void)hipEventDestroy(ctx->ev_tab0);
        if (ctx->ev_tab1) (void)hipEventDestroy(ctx->ev_tab1);
        if (ctx->aux2) (void)hipStreamDestroy(ctx->aux2);
    }
    delete ctx;
}

const char* bhs_last_error(const bhs_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int bhs_sync(bhs_ctx* ctx) {
    if (!ctx) return BHS_E_INVALID;
    DeviceGuard g(ctx->device);
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_fail(ctx, e, "bhs_sync");
    return BHS_OK;
}

int bhs_trace_rays(bhs_ctx* ctx, const double* pos3_host, const double* dir3_host, int64_t n,
                   int32_t max_steps, double step_scale, double esc_tol, double r_esc_min,
                   bhs_trace_result* out_host) {
    if (!ctx) return BHS_E_INVALID;
    if (n < 0 || (n > 0 && (!pos3_host || !dir3_host || !out_host)))
        return fail(ctx, BHS_E_INVALID, "bhs_trace_rays: NULL buffer or n < 0%s");
    if (n == 0) return BHS_OK;
    if (max_steps < 0) return fail(ctx, BHS_E_INVALID, "bhs_trace_rays: max_steps < 0%s");
    DeviceGuard g(ctx->device);
    const size_t in_bytes = (size_t)n * 3 * sizeof(double);
    const size_t out_bytes = (size_t)n * sizeof(bhs_trace_result);
    int st = ensure(ctx, &ctx->tstage, &ctx->tstage_bytes, 2 * in_bytes + out_bytes, "trace staging");
    if (st != BHS_OK) return st;
    char* base = (char*)ctx->tstage;
    double* dpos = (double*)base;
    double* ddir = (double*)(base + in_bytes);
    void* dout = base + 2 * in_bytes;
    hipStream_t s = ctx->aux;
    hipError_t e;
    if ((e = hipMemcpyAsync(dpos, pos3_host, in_bytes, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(ddir, dir3_host, in_bytes, hipMemcpyHostToDevice, s)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_trace_rays upload");
    if ((e = bhs::launch_trace_T(dpos, ddir, n, max_steps, step_scale, esc_tol, r_esc_min, dout, s)) !=
        hipSuccess)
        return hip_fail(ctx, e, "bhs_trace_rays launch");
    if ((e = hipMemcpyAsync(out_host, dout, out_bytes, hipMemcpyDeviceToHost, s)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_trace_rays download");
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(ctx, e, "bhs_trace_rays");
    return BHS_OK;
}

int bhs_render(bhs_ctx* ctx, const bhs_camera* cam, const bhs_march_params* params,
               const bhs_ray_soa* out, void* stream) {
    if (!ctx) return BHS_E_INVALID;
    int st;
    if ((st = check_camera(ctx, cam)) != BHS_OK) return st;
    if ((st = check_params(ctx, params)) != BHS_OK) return st;
    if (!out) return fail(ctx, BHS_E_INVALID, "bhs_render: out is NULL%s");
    if (params->root64 && (!out->n_cross_dev || !out->flags_dev || !out->disk_dev))
        return fail(ctx, BHS_E_INVALID, "bhs_render: root64 needs n_cross, flags and disk outputs%s");
    DeviceGuard g(ctx->device);
    const int64_t n = (int64_t)cam->width * cam->height;
    bhs::MarchConst c = make_const(params);
    fill_camera(c, cam);
    c.ray_offset = 0;
    const bool roots = params->root64 && params->precision != BHS_PREC_F64;
    bhs::Scratch sc;
    if ((st = prepare_scratch(ctx, n, bhs::camera_items(cam->width, cam->height), roots, &sc)) !=
        BHS_OK)
        return st;
    hipStream_t s = (hipStream_t)stream;
    if (int os = order_stream(ctx, s)) return os;
    hipError_t e;
    bhs_ctx::TimingRec* tr = next_timer(ctx);
    if (tr) (void)hipEventRecord(tr->call0, s);
    if ((e = bhs::launch_zero(ctx->scratch, kStatsBytes, s)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_render clear");
    if ((e = bhs::launch_render(c, make_soa(out, n), sc, ctx->tuning, ctx->num_cus, s, ctx->aux,
                                ctx->ev_fork, ctx->ev_join, tr ? tr->k0 : nullptr,
                                tr ? tr->k1 : nullptr)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_render launch");
    if (tr) (void)hipEventRecord(tr->call1, s);
    note_stream(ctx, s);
    return BHS_OK;
}

int bhs_render_noalloc(bhs_ctx* ctx, const bhs_camera* cam, const bhs_march_params* params,
                       const bhs_ray_soa* out, void* stream) {
    if (!ctx) return BHS_E_INVALID;
    int st;
    if ((st = check_camera(ctx, cam)) != BHS_OK) return st;
    if ((st = check_params(ctx, params)) != BHS_OK) return st;
    if (!out) return fail(ctx, BHS_E_INVALID, "bhs_render_noalloc: out is NULL%s");
    if (params->root64 && (!out->n_cross_dev || !out->flags_dev || !out->disk_dev))
        return fail(ctx, BHS_E_INVALID,
                    "bhs_render_noalloc: root64 needs n_cross, flags and disk outputs%s");
    const int64_t n = (int64_t)cam->width * cam->height;
    const uint32_t n_items = bhs::camera_items(cam->width, cam->height);
    const bool roots = params->root64 && params->precision != BHS_PREC_F64;
    const size_t need = scratch_bytes(n, n_items, roots);
    if (need > kNoAllocScratchBytes) {
        char what[96];
        std::snprintf(what, sizeof what, "%zu bytes", need);
        return fail(ctx, BHS_E_INVALID,
                    "bhs_render_noalloc: this render needs %s of scratch, more than the 48 MiB "
                    "static maximum (4 bytes per work item, 56 more per ray with root64)", what);
    }
    if (!ctx->static_scratch)
        return fail(ctx, BHS_E_STATE, "bhs_render_noalloc: static scratch not resolved%s");
    DeviceGuard g(ctx->device);
    bhs::MarchConst c = make_const(params);
    fill_camera(c, cam);
    c.ray_offset = 0;
    bhs::Scratch sc;
    layout_scratch(ctx->static_scratch, n_items, roots, &sc);
    hipStream_t s = (hipStream_t)stream;
    if (int os = order_stream(ctx, s)) return os;
    hipError_t e;
    if ((e = bhs::launch_zero(ctx->static_scratch, kStatsBytes, s)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_render_noalloc clear");
    if ((e = bhs::launch_render(c, make_soa(out, n), sc, ctx->tuning, ctx->num_cus, s, ctx->aux,
                                ctx->ev_fork, ctx->ev_join, nullptr, nullptr)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_render_noalloc launch");
    note_stream(ctx, s);
    ctx->stats_src = ctx->static_scratch;
    return BHS_OK;
}

int bhs_march_rays(bhs_ctx* ctx, const float* pos3_dev, const float* dir3_dev, int64_t n,
                   const bhs_march_params* params, const bhs_ray_soa* out, void* stream) {
    if (!ctx) return BHS_E_INVALID;
    int st;
    if ((st = check_params(ctx, params)) != BHS_OK) return st;
    if (n < 0 || (n > 0 && (!pos3_dev || !dir3_dev || !out)))
        return fail(ctx, BHS_E_INVALID, "bhs_march_rays: NULL buffer or n < 0%s");
    DeviceGuard g(ctx->device);
    bhs::Scratch sc;
    if ((st = prepare_scratch(ctx, 0, 0, false, &sc)) != BHS_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    if (int os = order_stream(ctx, s)) return os;
    hipError_t e;
    if ((e = bhs::launch_zero(ctx->scratch, kStatsBytes, s)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_march_rays clear");
    note_stream(ctx, s);
    if (n == 0) return BHS_OK;
    bhs::MarchConst c = make_const(params);
    bhs_ctx::TimingRec* tr = next_timer(ctx);
    if (tr) (void)hipEventRecord(tr->call0, s);
    if ((e = bhs::launch_march_rays(pos3_dev, dir3_dev, n, c, make_soa(out, n), sc, s,
                                    tr ? tr->k0 : nullptr, tr ? tr->k1 : nullptr)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_march_rays launch");
    if (tr) (void)hipEventRecord(tr->call1, s);
    return BHS_OK;
}

int bhs_render_batch(bhs_ctx* ctx, const bhs_camera* cams, const bhs_march_params* params,
                     int32_t ncam, const bhs_ray_soa* out, void* stream) {
    if (!ctx) return BHS_E_INVALID;
    if (!out) return fail(ctx, BHS_E_INVALID, "bhs_render_batch: out is NULL%s");
    BatchTable t;
    int st;
    if ((st = build_batch(ctx, cams, params, ncam, &t)) != BHS_OK) return st;
    if (t.any_root && (!out->n_cross_dev || !out->flags_dev || !out->disk_dev))
        return fail(ctx, BHS_E_INVALID, "bhs_render_batch: root64 needs n_cross, flags and disk outputs%s");
    DeviceGuard g(ctx->device);
    hipStream_t s = (hipStream_t)stream;
    if (int os = order_stream(ctx, s)) return os;
    void* table_dev = nullptr;
    if ((st = upload_batch(ctx, t, s, &table_dev)) != BHS_OK) return st;
    bhs::Scratch sc;
    if ((st = prepare_scratch(ctx, t.n_rays, t.n_items, t.any_root, &sc)) != BHS_OK) return st;
    hipError_t e;
    bhs_ctx::TimingRec* tr = next_timer(ctx);
    if (tr) (void)hipEventRecord(tr->call0, s);
    if ((e = bhs::launch_zero(ctx->scratch, kStatsBytes, s)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_render_batch clear");
    const bhs::MarchConst* table = (const bhs::MarchConst*)ctx->table_dev;
    const uint32_t* base = (const uint32_t*)((char*)ctx->table_dev + table_cams_bytes(ncam));
    if ((e = bhs::launch_batch(table, base, (uint32_t)ncam, t.n_items, t.any_band, t.any_root,
                               make_soa(out, t.n_rays), sc, ctx->tuning, ctx->num_cus, s, ctx->aux,
                               ctx->ev_fork, ctx->ev_join, tr ? tr->k0 : nullptr,
                               tr ? tr->k1 : nullptr)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_render_batch launch");
    if (tr) (void)hipEventRecord(tr->call1, s);
    note_stream(ctx, s);
    return BHS_OK;
}

int bhs_sky_footprint(bhs_ctx* ctx, const bhs_camera* cam, const bhs_ray_soa* rays,
                      int32_t coarse, const bhs_footprint_soa* out, void* stream) {
    if (!ctx) return BHS_E_INVALID;
    int st;
    if ((st = check_camera(ctx, cam)) != BHS_OK) return st;
    if (!rays || !rays->flags_dev || !rays->esc_dev || !out)
        return fail(ctx, BHS_E_INVALID, "bhs_sky_footprint: rays (flags, esc) or out is NULL%s");
    DeviceGuard g(ctx->device);
    bhs::MarchConst c;
    std::memset(&c, 0, sizeof c);
    fill_camera(c, cam);
    c.ray_offset = 0;
    const int64_t n = (int64_t)cam->width * cam->height;
    const bhs::FootprintPlanes o{out->ddx_dev, out->ddy_dev,  out->e1_dev,       out->e2_dev,
                                 out->rinv_dev, out->area_dev, out->mu_dev, out->fallback_dev, n};
    hipStream_t s = (hipStream_t)stream;
    if (int os = order_stream(ctx, s)) return os;
    hipError_t e;
    if ((e = bhs::launch_footprint(c, bhs::FootprintIn{rays->flags_dev, rays->esc_dev}, o,
                                   coarse != 0, s)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_sky_footprint launch");
    note_stream(ctx, s);
    return BHS_OK;
}

int bhs_sky_footprint_batch(bhs_ctx* ctx, const bhs_camera* cams, int32_t ncam,
                            const bhs_ray_soa* rays, int32_t coarse,
                            const bhs_footprint_soa* out, void* stream) {
    if (!ctx) return BHS_E_INVALID;
    if (!cams) return fail(ctx, BHS_E_INVALID, "bhs_sky_footprint_batch: cameras is NULL%s");
    if (ncam < 1 || (uint32_t)ncam > bhs::kMaxBatchCams)
        return fail(ctx, BHS_E_INVALID, "bhs_sky_footprint_batch: ncam must be in 1..64%s");
    if (!rays || !rays->flags_dev || !rays->esc_dev || !out)
        return fail(ctx, BHS_E_INVALID, "bhs_sky_footprint_batch: rays (flags, esc) or out is NULL%s");
    BatchTable t;
    t.cams.resize((size_t)ncam);
    t.item_base.assign((size_t)ncam + 1, 0u);
    int64_t max_quads = 0;
    bool widths_even = true;
    for (int32_t k = 0; k < ncam; ++k) {
        int st;
        if ((st = check_camera(ctx, &cams[k])) != BHS_OK) return st;
        bhs::MarchConst c;
        std::memset(&c, 0, sizeof c);
        fill_camera(c, &cams[k]);
        c.ray_offset = t.n_rays;
        t.cams[(size_t)k] = c;
        t.n_rays += (int64_t)cams[k].width * cams[k].height;
        max_quads = std::max(max_quads, (int64_t)((cams[k].width + 1) / 2) *
                                            ((cams[k].height + 1) / 2));
        widths_even = widths_even && (cams[k].width & 1) == 0;
    }
    DeviceGuard g(ctx->device);
    hipStream_t s = (hipStream_t)stream;
    if (int os = order_stream(ctx, s)) return os;
    void* table_dev = nullptr;
    int st;
    if ((st = upload_batch(ctx, t, s, &table_dev)) != BHS_OK) return st;
    const bhs::FootprintPlanes o{out->ddx_dev, out->ddy_dev,  out->e1_dev,       out->e2_dev,
                                 out->rinv_dev, out->area_dev, out->mu_dev, out->fallback_dev,
                                 t.n_rays};
    hipError_t e;
    if ((e = bhs::launch_footprint_batch((const bhs::MarchConst*)table_dev, (uint32_t)ncam,
                                         max_quads, widths_even,
                                         bhs::FootprintIn{rays->flags_dev, rays->esc_dev}, o,
                                         coarse != 0, s)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_sky_footprint_batch launch");
    note_stream(ctx, s);
    return BHS_OK;
}

int bhs_gas_lens_table(bhs_ctx* ctx, float cam_r, float alpha_crit, int32_t n_rays,
                       int32_t n_phis, float* table_dev, void* stream) {
    if (!ctx) return BHS_E_INVALID;
    if (!table_dev) return fail(ctx, BHS_E_INVALID, "bhs_gas_lens_table: table is NULL%s");
    if (!std::isfinite(cam_r) || !(cam_r > 2.01f))
        return fail(ctx, BHS_E_INVALID, "bhs_gas_lens_table: cam_r must be finite and > 2.01%s");
    if (!std::isfinite(alpha_crit) || !(alpha_crit > 0.0f) || !(alpha_crit < 3.14159265f))
        return fail(ctx, BHS_E_INVALID, "bhs_gas_lens_table: alpha_crit must lie in (0, pi)%s");
    if (n_rays < 1 || n_phis < 2 || (int64_t)n_rays * n_phis > (int64_t)1 << 31)
        return fail(ctx, BHS_E_INVALID, "bhs_gas_lens_table: need n_rays >= 1, n_phis >= 2%s");
    DeviceGuard g(ctx->device);
    const bhs::LensArgs a{cam_r, alpha_crit, BHS_LENS_RMAX, n_rays, n_phis, BHS_LENS_MAX_STEPS};
    hipStream_t s = (hipStream_t)stream;
    if (int os = order_stream(ctx, s)) return os;
    hipError_t e;
    if ((e = bhs::launch_gas_lens(a, table_dev, s)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_gas_lens_table launch");
    note_stream(ctx, s);
    return BHS_OK;
}

int bhs_gas_lens_tables(bhs_ctx* ctx, int32_t n_tables, const float* cam_r,
                        const float* alpha_crit, int32_t n_rays, int32_t n_phis,
                        float* tables_dev, void* stream) {
    if (!ctx) return BHS_E_INVALID;
    if (!tables_dev || !cam_r || !alpha_crit)
        return fail(ctx, BHS_E_INVALID, "bhs_gas_lens_tables: NULL argument%s");
    if (n_tables < 1 || n_tables > bhs::kLensBatchMax)
        return fail(ctx, BHS_E_INVALID, "bhs_gas_lens_tables: need 1 <= n_tables <= 64%s");
    if (n_rays < 1 || n_phis < 2 || (int64_t)n_rays * n_phis * n_tables > (int64_t)1 << 31)
        return fail(ctx, BHS_E_INVALID,
                    "bhs_gas_lens_tables: need n_rays >= 1, n_phis >= 2, at most 2^31 entries%s");
    bhs::LensBatch b{};
    for (int32_t k = 0; k < n_tables; ++k) {
        if (!std::isfinite(cam_r[k]) || !(cam_r[k] > 2.01f))
            return fail(ctx, BHS_E_INVALID, "bhs_gas_lens_tables: cam_r must be finite and > 2.01%s");
        if (!std::isfinite(alpha_crit[k]) || !(alpha_crit[k] > 0.0f) ||
            !(alpha_crit[k] < 3.14159265f))
            return fail(ctx, BHS_E_INVALID, "bhs_gas_lens_tables: alpha_crit must lie in (0, pi)%s");
        b.cam_r[k] = cam_r[k];
        b.alpha_crit[k] = alpha_crit[k];
    }
    b.r_max = BHS_LENS_RMAX;
    b.n_rays = n_rays;
    b.n_phis = n_phis;
    b.max_steps = BHS_LENS_MAX_STEPS;
    b.n_tables = n_tables;
    DeviceGuard g(ctx->device);
    hipStream_t s = (hipStream_t)stream;
    if (int os = order_stream(ctx, s)) return os;
    hipError_t e;
    if ((e = bhs::launch_gas_lens_batch(b, tables_dev, s)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_gas_lens_tables launch");
    note_stream(ctx, s);
    return BHS_OK;
}

void bhs_graph_destroy(bhs_graph* g) {
    if (!g) return;
    {
        DeviceGuard dg(g->device);
        if (g->cap) (void)hipStreamSynchronize(g->cap);
        (void)hipDeviceSynchronize();
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        if (g->ev_fork) (void)hipEventDestroy(g->ev_fork);
        if (g->ev_join) (void)hipEventDestroy(g->ev_join);
        if (g->aux) (void)hipStreamDestroy(g->aux);
        if (g->cap) (void)hipStreamDestroy(g->cap);
        if (g->mem) (void)hipFree(g->mem);
    }
    delete g;
}

int bhs_graph_create(bhs_ctx* ctx, const bhs_camera* cams, const bhs_march_params* params,
                     int32_t ncam, const bhs_ray_soa* out, bhs_graph** graph) {
    if (!ctx) return BHS_E_INVALID;
    if (!out || !graph) return fail(ctx, BHS_E_INVALID, "bhs_graph_create: out or graph is NULL%s");
    *graph = nullptr;
    BatchTable t;
    int st;
    if ((st = build_batch(ctx, cams, params, ncam, &t)) != BHS_OK) return st;
    if (t.any_root && (!out->n_cross_dev || !out->flags_dev || !out->disk_dev))
        return fail(ctx, BHS_E_INVALID, "bhs_graph_create: root64 needs n_cross, flags and disk outputs%s");
    DeviceGuard dg(ctx->device);
    bhs_graph* g = new (std::nothrow) bhs_graph();
    if (!g) return fail(ctx, BHS_E_ALLOC, "bhs_graph_create: out of host memory%s");
    g->device = ctx->device;
    g->ncam = ncam;
    g->n_rays = t.n_rays;
    g->n_items = t.n_items;
    g->any_band = t.any_band;
    g->any_root = t.any_root;
    for (int32_t k = 0; k < ncam; ++k) {
        g->dims.push_back(cams[k].width);
        g->dims.push_back(cams[k].height);
    }
    const size_t tb = table_bytes(ncam);
    g->bytes = tb + scratch_bytes(t.n_rays, t.n_items, t.any_root);
    hipError_t e;
    if ((e = hipMalloc(&g->mem, g->bytes)) != hipSuccess) {
        g->mem = nullptr;
        bhs_graph_destroy(g);
        return hip_fail(ctx, e, "bhs_graph_create alloc");
    }
    std::vector<char> host(tb);
    pack_table(t, host.data());
    if ((e = hipMemcpy(g->mem, host.data(), tb, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&g->aux, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&g->ev_fork, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&g->ev_join, hipEventDisableTiming)) != hipSuccess) {
        bhs_graph_destroy(g);
        return hip_fail(ctx, e, "bhs_graph_create setup");
    }
    bhs::Scratch sc;
    void* smem = (char*)g->mem + tb;
    layout_scratch(smem, t.n_items, t.any_root, &sc);
    const bhs::MarchConst* table = (const bhs::MarchConst*)g->mem;
    const uint32_t* base = (const uint32_t*)((char*)g->mem + table_cams_bytes(ncam));
    if ((e = hipStreamBeginCapture(g->cap, hipStreamCaptureModeThreadLocal)) != hipSuccess) {
        bhs_graph_destroy(g);
        return hip_fail(ctx, e, "bhs_graph_create begin capture");
    }
    hipError_t le = bhs::launch_zero(smem, kStatsBytes, g->cap);
    if (le == hipSuccess)
        le = bhs::launch_batch(table, base, (uint32_t)ncam, t.n_items, t.any_band, t.any_root,
                               make_soa(out, t.n_rays), sc, ctx->tuning, ctx->num_cus, g->cap,
                               g->aux, g->ev_fork, g->ev_join, nullptr, nullptr);
    e = hipStreamEndCapture(g->cap, &g->graph);
    if (le != hipSuccess || e != hipSuccess) {
        bhs_graph_destroy(g);
        return hip_fail(ctx, le != hipSuccess ? le : e, "bhs_graph_create capture");
    }
    if ((e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0)) != hipSuccess) {
        bhs_graph_destroy(g);
        return hip_fail(ctx, e, "bhs_graph_create instantiate");
    }
    *graph = g;
    return BHS_OK;
}

int bhs_graph_update(bhs_ctx* ctx, bhs_graph* g, const bhs_camera* cams,
                     const bhs_march_params* params) {
    if (!ctx) return BHS_E_INVALID;
    if (!g) return fail(ctx, BHS_E_INVALID, "bhs_graph_update: graph is NULL%s");
    BatchTable t;
    int st;
    if ((st = build_batch(ctx, cams, params, g->ncam, &t)) != BHS_OK) return st;
    for (int32_t k = 0; k < g->ncam; ++k)
        if (cams[k].width != g->dims[2 * k] || cams[k].height != g->dims[2 * k + 1])
            return fail(ctx, BHS_E_INVALID, "bhs_graph_update: camera sizes must not change%s");
    if (t.any_band != g->any_band || t.any_root != g->any_root)
        return fail(ctx, BHS_E_INVALID, "bhs_graph_update: band/root settings must not change%s");
    DeviceGuard dg(g->device);
    hipError_t e;
    if ((e = hipStreamSynchronize(g->cap)) != hipSuccess) return hip_fail(ctx, e, "bhs_graph_update");
    (void)hipDeviceSynchronize();
    std::vector<char> host(table_bytes(g->ncam));
    pack_table(t, host.data());
    if ((e = hipMemcpy(g->mem, host.data(), host.size(), hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_graph_update copy");
    return BHS_OK;
}

int bhs_graph_launch(bhs_ctx* ctx, bhs_graph* g, void* stream) {
    if (!ctx) return BHS_E_INVALID;
    if (!g) return fail(ctx, BHS_E_INVALID, "bhs_graph_launch: graph is NULL%s");
    DeviceGuard dg(g->device);
    hipStream_t s = (hipStream_t)stream;
    bhs_ctx::TimingRec* tr = next_timer(ctx);
    if (tr) {
        (void)hipEventRecord(tr->call0, s);
        (void)hipEventRecord(tr->k0, s);
    }
    hipError_t e = hipGraphLaunch(g->exec, s);
    if (e != hipSuccess) return hip_fail(ctx, e, "bhs_graph_launch");
    if (tr) {
        (void)hipEventRecord(tr->k1, s);
        (void)hipEventRecord(tr->call1, s);
    }
    return BHS_OK;
}

int bhs_graph_stats(bhs_ctx* ctx, bhs_graph* g, bhs_stats* out_host) {
    if (!ctx) return BHS_E_INVALID;
    if (!g || !out_host) return fail(ctx, BHS_E_INVALID, "bhs_graph_stats: NULL argument%s");
    DeviceGuard dg(g->device);
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_fail(ctx, e, "bhs_graph_stats sync");
    bhs::Stats st;
    if ((e = hipMemcpy(&st, (char*)g->mem + table_bytes(g->ncam), sizeof st, hipMemcpyDeviceToHost)) !=
        hipSuccess)
        return hip_fail(ctx, e, "bhs_graph_stats copy");
    std::memset(out_host, 0, sizeof *out_host);
    out_host->rays = st.rays;
    out_host->steps = st.steps;
    out_host->crossings = st.crossings;
    out_host->escapes = st.escapes;
    out_host->band64 = st.band64;
    out_host->fallback = st.fallback;
    return BHS_OK;
}

int bhs_set_tuning(bhs_ctx* ctx, const bhs_tuning* t) {
    if (!ctx) return BHS_E_INVALID;
    if (!t) return fail(ctx, BHS_E_INVALID, "tuning is NULL%s");
    if ((t->mode != 0 && t->mode != 1) || (t->block != 64 && t->block != 256) ||
        t->waves_per_simd < 1 || t->waves_per_simd > 8 || t->refill_min < 1 || t->refill_min > 64)
        return fail(ctx, BHS_E_INVALID,
                    "tuning out of range (mode 0|1, block 64|256, waves_per_simd 1..8, refill_min 1..64)%s");
    ctx->tuning = bhs::Tuning{t->mode, t->block, t->waves_per_simd, t->refill_min};
    return BHS_OK;
}

int bhs_get_tuning(bhs_ctx* ctx, bhs_tuning* t) {
    if (!ctx || !t) return BHS_E_INVALID;
    t->mode = ctx->tuning.mode;
    t->block = ctx->tuning.block;
    t->waves_per_simd = ctx->tuning.waves_per_simd;
    t->refill_min = ctx->tuning.refill_min;
    return BHS_OK;
}

int bhs_autotune(bhs_ctx* ctx, const bhs_camera* cam, const bhs_march_params* params,
                 const bhs_ray_soa* out, int32_t reps, bhs_tuning* best, double* best_ms) {
    if (!ctx) return BHS_E_INVALID;
    if (reps < 1 || reps > 1000) return fail(ctx, BHS_E_INVALID, "reps must be in 1..1000%s");
    static const bhs::Tuning cands[] = {
        {0, 256, 8, 64}, {0, 64, 8, 64},
        {1, 256, 8, 64}, {1, 256, 8, 32},
        {1, 256, 7, 64}, {1, 256, 7, 48}, {1, 256, 7, 32},
        {1, 256, 6, 64}, {1, 256, 6, 32}, {1, 256, 5, 64},
        {1, 256, 4, 64}, {1, 256, 4, 48}, {1, 256, 4, 32},
        {1, 64, 7, 64}, {1, 64, 4, 64}, {1, 64, 4, 32},
    };
    DeviceGuard g(ctx->device);
    const bhs::Tuning saved = ctx->tuning;
    const bool saved_timing = ctx->timing;
    double best_t = 1e30;
    bhs::Tuning best_c = saved;
    std::vector<double> ts;
    for (const bhs::Tuning& cand : cands) {
        ctx->tuning = cand;
        ctx->timing = true;
        ctx->timers_used = 0;
        ts.clear();
        for (int r = 0; r <= reps; ++r) {
            int st = bhs_render(ctx, cam, params, out, nullptr);
            if (st != BHS_OK) {
                ctx->tuning = saved;
                ctx->timing = saved_timing;
                ctx->timers_used = 0;
                return st;
            }
        }
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) {
            ctx->tuning = saved;
            ctx->timing = saved_timing;
            return hip_fail(ctx, e, "bhs_autotune");
        }
        for (size_t i = 1; i < ctx->timers_used; ++i) {   // first launch is the warm-up
            float a = 0.0f;
            if (hipEventElapsedTime(&a, ctx->timers[i].k0, ctx->timers[i].k1) == hipSuccess) ts.push_back(a);
        }
        ctx->timers_used = 0;
        if (ts.empty()) continue;
        std::sort(ts.begin(), ts.end());
        const double med = ts[ts.size() / 2];
        if (med < best_t) {
            best_t = med;
            best_c = cand;
        }
    }
    ctx->tuning = best_c;
    ctx->timing = saved_timing;
    if (best) {
        best->mode = best_c.mode;
        best->block = best_c.block;
        best->waves_per_simd = best_c.waves_per_simd;
        best->refill_min = best_c.refill_min;
    }
    if (best_ms) *best_ms = best_t;
    return BHS_OK;
}

int bhs_set_timing(bhs_ctx* ctx, int on) {
    if (!ctx) return BHS_E_INVALID;
    ctx->timing = on != 0;
    ctx->timers_used = 0;
    return BHS_OK;
}

int bhs_timing(bhs_ctx* ctx, double* march_ms_avg, double* total_ms_avg, int32_t* launches) {
    if (!ctx) return BHS_E_INVALID;
    DeviceGuard g(ctx->device);
    double km = 0.0, tm = 0.0;
    const size_t n = ctx->timers_used;
    for (size_t i = 0; i < n; ++i) {
        const bhs_ctx::TimingRec& r = ctx->timers[i];
        float a = 0.0f, b = 0.0f;
        hipError_t e;
        if ((e = hipEventSynchronize(r.call1)) != hipSuccess) return hip_fail(ctx, e, "bhs_timing");
        if ((e = hipEventElapsedTime(&a, r.k0, r.k1)) != hipSuccess ||
            (e = hipEventElapsedTime(&b, r.call0, r.call1)) != hipSuccess)
            return hip_fail(ctx, e, "bhs_timing elapsed");
        km += a;
        tm += b;
    }
    if (march_ms_avg) *march_ms_avg = n ? km / (double)n : 0.0;
    if (total_ms_avg) *total_ms_avg = n ? tm / (double)n : 0.0;
    if (launches) *launches = (int32_t)n;
    ctx->timers_used = 0;
    return BHS_OK;
}

int bhs_rsqrt_rn(bhs_ctx* ctx, const float* x_dev, float* y_dev, int64_t n, void* stream) {
    if (!ctx) return BHS_E_INVALID;
    if (n < 0 || (n > 0 && (!x_dev || !y_dev)))
        return fail(ctx, BHS_E_INVALID, "bhs_rsqrt_rn: NULL buffer or n < 0%s");
    if (n == 0) return BHS_OK;
    DeviceGuard g(ctx->device);
    hipStream_t s = (hipStream_t)stream;
    if (int os = order_stream(ctx, s)) return os;
    hipError_t e;
    if ((e = bhs::launch_rsqrt_rn(x_dev, y_dev, n, s)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_rsqrt_rn launch");
    note_stream(ctx, s);
    return BHS_OK;
}

int bhs_valu_probe(bhs_ctx* ctx, int32_t kind, int32_t waves_per_simd, int32_t iters, double* ms,
                   double* wave_instr) {
    if (!ctx || !ms || !wave_instr) return BHS_E_INVALID;
    if (kind < 0 || kind > 4 || waves_per_simd < 1 || waves_per_simd > 8 || iters < 1)
        return fail(ctx, BHS_E_INVALID, "bhs_valu_probe: kind 0..4, waves_per_simd 1..8, iters >= 1%s");
    DeviceGuard g(ctx->device);
    const uint32_t blocks = (uint32_t)ctx->num_cus * (uint32_t)waves_per_simd;   // 4 waves a block
    float* out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e;
    if ((e = hipMalloc((void**)&out, (size_t)blocks * 256 * sizeof(float))) != hipSuccess)
        return hip_fail(ctx, e, "bhs_valu_probe alloc");
    float t = 0.0f;
    e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = bhs::launch_valu_probe(kind, blocks, iters, out, nullptr);   // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    if (e == hipSuccess) e = bhs::launch_valu_probe(kind, blocks, iters, out, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (e != hipSuccess) return hip_fail(ctx, e, "bhs_valu_probe");
    *ms = t;
    *wave_instr = (double)blocks * 4.0 * (double)iters * 16.0 * (kind == 3 ? 2.0 : 1.0);
    return BHS_OK;
}

int bhs_last_stats(bhs_ctx* ctx, bhs_stats* out_host) {
    if (!ctx || !out_host) return BHS_E_INVALID;
    std::memset(out_host, 0, sizeof *out_host);
    const void* src = ctx->stats_src ? ctx->stats_src : ctx->scratch;
    if (!src) return BHS_OK;
    DeviceGuard g(ctx->device);
    hipError_t e = hipStreamSynchronize(ctx->last_stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "bhs_last_stats sync");
    bhs::Stats s;
    if ((e = hipMemcpy(&s, src, sizeof s, hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_last_stats copy");
    out_host->rays = s.rays;
    out_host->steps = s.steps;
    out_host->crossings = s.crossings;
    out_host->escapes = s.escapes;
    out_host->band64 = s.band64;
    out_host->fallback = s.fallback;
    out_host->surrogate = s.surrogate;
    return BHS_OK;
}

}  // extern "C"
