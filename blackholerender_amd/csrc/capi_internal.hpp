// capi_internal.hpp — the context object of the C-ABI and the helpers shared by its source files.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/bhs_trace.h"
#include "kernels.hpp"

struct bhs_ctx {
    int device = 0;
    std::string err;
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // second side stream (surrogate render: early FP32 fallback march beside the FP64 band)
    hipStream_t aux2 = nullptr;
    hipEvent_t ev_fork2 = nullptr, ev_join2 = nullptr;
    // surrogate render of one camera: the table built on aux2 beside the routing pre-pass
    hipEvent_t ev_tab0 = nullptr, ev_tab1 = nullptr;
    // Stream of the context's most recent stream-ordered call (has_last: one was made). A call on
    // another stream first waits for it (ev_order): scratch, camera table and lists are shared.
    hipStream_t last_stream = nullptr;
    bool has_last = false;
    hipEvent_t ev_order = nullptr;
    // Render scratch: [Stats | band_count (pad 256) | band_ids | roots]
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    // Device address of the library's static scratch array on this device (bhs_render_noalloc),
    // and where bhs_last_stats reads the latest call's counters (null: `scratch`).
    void* static_scratch = nullptr;
    const void* stats_src = nullptr;
    // Law-T staging (pos3, dir3, results).
    void* tstage = nullptr;
    size_t tstage_bytes = 0;
    // GPU timers: per launch, events before the call, around the main kernel, after the call.
    struct TimingRec { hipEvent_t call0, k0, k1, call1; };
    bool timing = false;
    std::vector<TimingRec> timers;
    size_t timers_used = 0;
    int num_cus = 256;
    bhs::Tuning tuning{1, 256, 6, 64};
    // Batch camera tables: device copy, pinned staging buffer and the event of its last upload.
    void* table_dev = nullptr;
    size_t table_dev_bytes = 0;
    void* table_host = nullptr;
    size_t table_host_bytes = 0;
    hipEvent_t ev_upload = nullptr;
    bool upload_pending = false;
    // Thin-disk shading: the baked noise volume (128^3 half) and blackbody table (256 x RGBA),
    // made at the first shading call, and the completion list ([count | pad | ray ids]).
    uint16_t* noise_vol = nullptr;
    float* bb_lut = nullptr;
    void* shade_list = nullptr;
    size_t shade_list_bytes = 0;
};


namespace bhs {
namespace capi {

// Launch counters, cleared per launch: Stats, band counter, work counters and the segmented
// queue's mask in the first 256 bytes, then the segmented queue's counters.
constexpr size_t kStatsBytes = 256 + (size_t)bhs::kSegs * bhs::kSegStride * 4;

// Records `fmt` (one %s for `detail`) as the context's last error and returns `code`.
int fail(bhs_ctx* ctx, int code, const char* fmt, const char* detail = nullptr);
int hip_fail(bhs_ctx* ctx, hipError_t e, const char* where);

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

// Grows a device buffer, after draining work that may still use the old one.
int ensure(bhs_ctx* ctx, void** buf, size_t* cap, size_t need, const char* what);
int check_params(bhs_ctx* ctx, const bhs_march_params* p);
int check_camera(bhs_ctx* ctx, const bhs_camera* cam);
bhs::MarchConst make_const(const bhs_march_params* p);
void fill_camera(bhs::MarchConst& c, const bhs_camera* cam);
bhs::SoA make_soa(const bhs_ray_soa* o, int64_t n);
bhs_ctx::TimingRec* next_timer(bhs_ctx* ctx);
int prepare_scratch(bhs_ctx* ctx, int64_t n_rays, int64_t n_items, bool roots, bhs::Scratch* sc);
// Orders a call on stream s after the context's previous stream-ordered call when that one was
// issued on another stream (event on the old stream, wait on s).
int order_stream(bhs_ctx* ctx, hipStream_t s);
// Records s as the stream of the context's latest call.
inline void note_stream(bhs_ctx* ctx, hipStream_t s) {
    ctx->last_stream = s;
    ctx->has_last = true;
}

// Host-side camera table of a batch: constants, output offsets (prefix sums of width*height) and
// work-item prefix sums (camera_items per camera).
struct BatchTable {
    std::vector<bhs::MarchConst> cams;
    std::vector<uint32_t> item_base;
    int64_t n_rays = 0;
    uint32_t n_items = 0;
    bool any_band = false, any_root = false;
};
int build_batch(bhs_ctx* ctx, const bhs_camera* cams, const bhs_march_params* params, int32_t ncam,
                BatchTable* t);
size_t table_cams_bytes(int32_t ncam);
// Stream-ordered upload of the table into the context's device table (pinned staging).
int upload_batch(bhs_ctx* ctx, const BatchTable& t, hipStream_t s, void** table_dev);

}  // namespace capi
}  // namespace bhs
