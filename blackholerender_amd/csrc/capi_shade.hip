// capi_shade.hip — the C-ABI of the thin-disk shading pass, the frame accumulation and the disk
// tables (include/bhs_trace.h), around the kernels of shade.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/bhs_trace.h"
#include "capi_internal.hpp"
#include "shade.hpp"
#include "sky.hpp"
#include "volume.hpp"

using namespace bhs::capi;

namespace {

constexpr size_t kNoiseTexels = (size_t)bhs::shade::kNoiseSize * bhs::shade::kNoiseSize *
                                bhs::shade::kNoiseSize;

// Bakes the context's noise volume and blackbody table once (synchronous, first call only).
int ensure_disk_tables(bhs_ctx* ctx) {
    if (ctx->noise_vol && ctx->bb_lut) return BHS_OK;
    hipError_t e;
    if (!ctx->noise_vol) {
        if ((e = hipMalloc((void**)&ctx->noise_vol, kNoiseTexels * sizeof(uint16_t))) != hipSuccess) {
            ctx->noise_vol = nullptr;
            (void)hipGetLastError();
            return fail(ctx, BHS_E_ALLOC, "device allocation of the noise volume failed%s");
        }
        if ((e = bhs::launch_noise_bake(ctx->noise_vol, bhs::shade::kNoiseSeed, nullptr)) !=
                hipSuccess ||
            (e = hipDeviceSynchronize()) != hipSuccess) {
            (void)hipFree(ctx->noise_vol);
            ctx->noise_vol = nullptr;
            return hip_fail(ctx, e, "noise volume bake");
        }
    }
    std::vector<float> lut((size_t)bhs::shade::kLutSize * 4);
    bhs::bake_blackbody_lut(lut.data());
    if ((e = hipMalloc((void**)&ctx->bb_lut, lut.size() * sizeof(float))) != hipSuccess) {
        ctx->bb_lut = nullptr;
        (void)hipGetLastError();
        return fail(ctx, BHS_E_ALLOC, "device allocation of the blackbody table failed%s");
    }
    if ((e = hipMemcpy(ctx->bb_lut, lut.data(), lut.size() * sizeof(float),
                       hipMemcpyHostToDevice)) != hipSuccess) {
        (void)hipFree(ctx->bb_lut);
        ctx->bb_lut = nullptr;
        return hip_fail(ctx, e, "blackbody table upload");
    }
    return BHS_OK;
}

int check_shading(bhs_ctx* ctx, const bhs_disk_shading* sh) {
    if (!sh) return fail(ctx, BHS_E_INVALID, "shading parameters are NULL%s");
    const float v[] = {sh->density, sh->opacity, sh->peak_temp, sh->brightness, sh->noise_scale,
                       sh->noise_strength, sh->warmth, sh->doppler, sh->time, sh->inflow};
    for (float x : v)
        if (!std::isfinite(x)) return fail(ctx, BHS_E_INVALID, "shading parameters must be finite%s");
    if (sh->noise_scale < 0.0f || sh->noise_scale > 1000.0f || std::fabs(sh->time) > 1e6f ||
        std::fabs(sh->inflow) > 1e3f || sh->density < 0.0f || sh->opacity < 0.0f ||
        sh->peak_temp < 0.0f)
        return fail(ctx, BHS_E_INVALID,
                    "shading: need density, opacity, peak_temp >= 0, noise_scale in [0, 1000], "
                    "|time| <= 1e6, |inflow| <= 1e3%s");
    return BHS_OK;
}

int shade_common(bhs_ctx* ctx, const bhs_camera* cams, const bhs_march_params* params,
                 int32_t ncam, const bhs_disk_shading* sh, const bhs_ray_soa* rays,
                 const bhs_shade_soa* out, void* stream, const char* what) {
    if (!ctx) return BHS_E_INVALID;
    if (!cams || !params) return fail(ctx, BHS_E_INVALID, "%s: cameras or params is NULL", what);
    if (ncam < 1 || (uint32_t)ncam > bhs::kMaxBatchCams)
        return fail(ctx, BHS_E_INVALID, "%s: ncam must be in 1..64", what);
    if (!rays || !rays->flags_dev || !rays->n_cross_dev || !rays->disk_dev)
        return fail(ctx, BHS_E_INVALID, "%s: rays need flags, n_cross and disk planes", what);
    if (!out || !out->rgb_dev || !out->sky_weight_dev)
        return fail(ctx, BHS_E_INVALID, "%s: out needs rgb and sky_weight planes", what);
    int st;
    if ((st = check_shading(ctx, sh)) != BHS_OK) return st;
    BatchTable t;
    t.cams.resize((size_t)ncam);
    t.item_base.assign((size_t)ncam + 1, 0u);
    for (int32_t k = 0; k < ncam; ++k) {
        if ((st = check_camera(ctx, &cams[k])) != BHS_OK) return st;
        if ((st = check_params(ctx, &params[k])) != BHS_OK) return st;
        if (params[k].disk_mode != BHS_DISK_THIN || params[k].cross_all)
            return fail(ctx, BHS_E_INVALID,
                        "%s: the rays must come from a thin-disk render without cross_all", what);
        if (params[k].disk_inner != params[0].disk_inner ||
            params[k].disk_outer != params[0].disk_outer)
            return fail(ctx, BHS_E_INVALID, "%s: every camera must share the disk annulus", what);
        bhs::MarchConst c = make_const(&params[k]);
        fill_camera(c, &cams[k]);
        c.ray_offset = t.n_rays;
        t.cams[(size_t)k] = c;
        t.n_rays += (int64_t)cams[k].width * cams[k].height;
    }
    if (t.n_rays > (int64_t)0xFFFFFFFFll)
        return fail(ctx, BHS_E_INVALID, "%s: more than 2^32 rays", what);
    DeviceGuard g(ctx->device);
    if ((st = ensure_disk_tables(ctx)) != BHS_OK) return st;
    const size_t list_bytes = 256 + (size_t)t.n_rays * sizeof(uint32_t);
    if ((st = ensure(ctx, &ctx->shade_list, &ctx->shade_list_bytes, list_bytes,
                     "shading list")) != BHS_OK)
        return st;
    hipStream_t s = (hipStream_t)stream;
    if (int os = order_stream(ctx, s)) return os;
    void* table_dev = nullptr;
    if ((st = upload_batch(ctx, t, s, &table_dev)) != BHS_OK) return st;
    const bhs::shade::ShadeConst sc{sh->density,        sh->opacity, sh->peak_temp, sh->brightness,
                                    sh->noise_scale,    sh->noise_strength, sh->warmth,
                                    sh->doppler,        sh->time,    sh->inflow,
                                    params[0].disk_inner, params[0].disk_outer};
    const bhs::ShadeList list{(uint32_t*)ctx->shade_list,
                              (uint32_t*)((char*)ctx->shade_list + 256), (uint32_t)t.n_rays};
    const bhs::ShadeOut o{out->rgb_dev, out->sky_weight_dev, t.n_rays};
    hipError_t e;
    bhs_ctx::TimingRec* tr = next_timer(ctx);
    if (tr) (void)hipEventRecord(tr->call0, s);
    if ((e = bhs::launch_zero(ctx->shade_list, 256, s)) != hipSuccess)
        return hip_fail(ctx, e, "shading clear");
    if ((e = bhs::launch_shade((const bhs::MarchConst*)table_dev, (uint32_t)ncam, t.n_rays,
                               make_soa(rays, t.n_rays), o, sc, ctx->noise_vol, ctx->bb_lut, list,
                               ctx->num_cus, s, tr ? tr->k0 : nullptr, tr ? tr->k1 : nullptr)) !=
        hipSuccess)
        return hip_fail(ctx, e, "shading launch");
    if (tr) (void)hipEventRecord(tr->call1, s);
    note_stream(ctx, s);
    return BHS_OK;
}

int volumetric_common(bhs_ctx* ctx, const bhs_camera* cams, const bhs_march_params* params,
                      int32_t ncam, const bhs_disk_shading* sh, const bhs_ray_soa* rays,
                      const bhs_shade_soa* out, void* stream, const char* what) {
    if (!ctx) return BHS_E_INVALID;
    if (!cams || !params) return fail(ctx, BHS_E_INVALID, "%s: cameras or params is NULL", what);
    if (ncam < 1 || (uint32_t)ncam > bhs::kMaxBatchCams)
        return fail(ctx, BHS_E_INVALID, "%s: ncam must be in 1..64", what);
    if (!out || !out->rgb_dev || !out->sky_weight_dev)
        return fail(ctx, BHS_E_INVALID, "%s: out needs rgb and sky_weight planes", what);
    int st;
    if ((st = check_shading(ctx, sh)) != BHS_OK) return st;
    // The batch table takes the thin-disk form of the parameters (same step control fields); a
    // single FP64 camera marches every ray in FP64.
    std::vector<bhs_march_params> pp(params, params + ncam);
    bool all_f64 = false;
    for (int32_t k = 0; k < ncam; ++k) {
        if (pp[k].disk_mode != BHS_DISK_VOLUMETRIC)
            return fail(ctx, BHS_E_INVALID, "%s: disk_mode must be BHS_DISK_VOLUMETRIC", what);
        if (pp[k].root64 || pp[k].cross_all)
            return fail(ctx, BHS_E_INVALID, "%s: root64 and cross_all do not apply", what);
        pp[k].disk_mode = BHS_DISK_THIN;
        if (pp[k].precision == BHS_PREC_F64) {
            if (ncam > 1)
                return fail(ctx, BHS_E_INVALID, "%s: BHS_PREC_F64 is for single cameras", what);
            all_f64 = true;
            pp[k].precision = BHS_PREC_F32;
        }
    }
    BatchTable t;
    if ((st = build_batch(ctx, cams, pp.data(), ncam, &t)) != BHS_OK) return st;
    DeviceGuard g(ctx->device);
    if ((st = ensure_disk_tables(ctx)) != BHS_OK) return st;
    bhs::Scratch sc;
    if ((st = prepare_scratch(ctx, t.n_rays, t.n_items, false, &sc)) != BHS_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    if (int os = order_stream(ctx, s)) return os;
    void* table_dev = nullptr;
    if ((st = upload_batch(ctx, t, s, &table_dev)) != BHS_OK) return st;
    const bhs::shade::ShadeConst sc_sh{sh->density,     sh->opacity, sh->peak_temp, sh->brightness,
                                       sh->noise_scale, sh->noise_strength, sh->warmth,
                                       sh->doppler,     sh->time,    sh->inflow, 0.0f, 0.0f};
    bhs_ray_soa none{};
    const bhs::SoA ro = make_soa(rays ? rays : &none, t.n_rays);
    const bhs::ShadeOut o{out->rgb_dev, out->sky_weight_dev, t.n_rays};
    hipError_t e;
    bhs_ctx::TimingRec* tr = next_timer(ctx);
    if (tr) (void)hipEventRecord(tr->call0, s);
    if ((e = bhs::launch_zero(ctx->scratch, kStatsBytes, s)) != hipSuccess)
        return hip_fail(ctx, e, "volumetric clear");
    const bhs::MarchConst* table = (const bhs::MarchConst*)table_dev;
    const uint32_t* base = (const uint32_t*)((char*)table_dev + table_cams_bytes(ncam));
    if ((e = bhs::launch_volumetric(table, base, (uint32_t)ncam, t.n_items, all_f64, t.any_band,
                                    sc_sh, ro, o, ctx->noise_vol, ctx->bb_lut, sc, ctx->num_cus, s,
                                    ctx->aux, ctx->ev_fork, ctx->ev_join, tr ? tr->k0 : nullptr,
                                    tr ? tr->k1 : nullptr)) != hipSuccess)
        return hip_fail(ctx, e, "volumetric launch");
    if (tr) (void)hipEventRecord(tr->call1, s);
    note_stream(ctx, s);
    return BHS_OK;
}

}  // namespace

namespace {

// The sky pass of one camera (batch false) or a batch of ncam cameras.
int sky_common(bhs_ctx* ctx, const bhs_camera* cams, int32_t ncam, bool batch,
               const bhs_ray_soa* rays, int32_t coarse, const bhs_sky_params* params,
               const bhs_sky_cube* cube, const float* sky_weight_dev, float* rgb_dev,
               void* stream) {
    const char* what = batch ? "bhs_sample_sky_batch" : "bhs_sample_sky";
    if (!ctx) return BHS_E_INVALID;
    if (!cams) return fail(ctx, BHS_E_INVALID, "sky: cameras is NULL%s");
    if (ncam < 1 || (uint32_t)ncam > bhs::kMaxBatchCams)
        return fail(ctx, BHS_E_INVALID, "sky: ncam must be in 1..64%s");
    if (!rays || !rays->flags_dev || !rays->esc_dev || !rgb_dev || !params)
        return fail(ctx, BHS_E_INVALID, "sky: rays (flags, esc), rgb or params is NULL%s");
    const float pv[4] = {params->sky_brightness, params->star_density, params->star_brightness,
                         params->doppler};
    for (float v : pv)
        if (!std::isfinite(v) || v < 0.0f)
            return fail(ctx, BHS_E_INVALID, "sky: parameters must be finite and >= 0%s");
    bhs::sky::Cube cb;
    std::memset(&cb, 0, sizeof cb);
    if (cube) {
        if (!cube->texels_dev || cube->res < 1 || cube->levels < 1 ||
            cube->levels > bhs::sky::kMaxLevels)
            return fail(ctx, BHS_E_INVALID,
                        "sky: the cube map needs texels, res >= 1 and 1..16 levels%s");
        cb.texels = (const float4*)cube->texels_dev;
        cb.res = cube->res;
        cb.levels = cube->levels;
        int64_t off = 0;
        for (int l = 0; l < cube->levels; ++l) {
            cb.off[l] = off;
            const int64_t S = std::max<int64_t>(1, (int64_t)cube->res >> l);
            off += 6 * S * S;
        }
    }
    BatchTable t;
    t.cams.resize((size_t)ncam);
    t.item_base.assign((size_t)ncam + 1, 0u);
    int64_t max_pix = 0;
    for (int32_t k = 0; k < ncam; ++k) {
        int st;
        if ((st = check_camera(ctx, &cams[k])) != BHS_OK) return st;
        bhs::MarchConst c;
        std::memset(&c, 0, sizeof c);
        fill_camera(c, &cams[k]);
        c.ray_offset = t.n_rays;
        t.cams[(size_t)k] = c;
        const int64_t np = (int64_t)cams[k].width * cams[k].height;
        t.n_rays += np;
        // one wave per 8x8-pixel tile
        const int64_t tiles = (int64_t)((cams[k].width + 7) / 8) * ((cams[k].height + 7) / 8);
        max_pix = std::max(max_pix, 64 * tiles);
    }
    DeviceGuard g(ctx->device);
    int st;
    if ((st = ensure_disk_tables(ctx)) != BHS_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    if (int os = order_stream(ctx, s)) return os;
    const void* table_dev = nullptr;
    if (batch) {
        void* td = nullptr;
        if ((st = upload_batch(ctx, t, s, &td)) != BHS_OK) return st;
        table_dev = td;
    }
    const bhs::sky::SkyConst k{params->sky_brightness, params->star_density,
                               params->star_brightness, params->doppler};
    hipError_t e;
    if ((e = bhs::launch_sky(t.cams[0], (const bhs::MarchConst*)table_dev, (uint32_t)ncam, max_pix,
                             bhs::FootprintIn{rays->flags_dev, rays->esc_dev}, t.n_rays, k, cb,
                             ctx->bb_lut, sky_weight_dev, rgb_dev, coarse != 0, s)) != hipSuccess)
        return hip_fail(ctx, e, what);
    note_stream(ctx, s);
    return BHS_OK;
}

}  // namespace

extern "C" {

int bhs_render_volumetric(bhs_ctx* ctx, const bhs_camera* cam, const bhs_march_params* params,
                          const bhs_disk_shading* shading, const bhs_ray_soa* rays,
                          const bhs_shade_soa* out, void* stream) {
    return volumetric_common(ctx, cam, params, 1, shading, rays, out, stream,
                             "bhs_render_volumetric");
}

int bhs_render_volumetric_batch(bhs_ctx* ctx, const bhs_camera* cams,
                                const bhs_march_params* params, int32_t ncam,
                                const bhs_disk_shading* shading, const bhs_ray_soa* rays,
                                const bhs_shade_soa* out, void* stream) {
    return volumetric_common(ctx, cams, params, ncam, shading, rays, out, stream,
                             "bhs_render_volumetric_batch");
}

int bhs_shade_thin_disk(bhs_ctx* ctx, const bhs_camera* cam, const bhs_march_params* params,
                        const bhs_disk_shading* shading, const bhs_ray_soa* rays,
                        const bhs_shade_soa* out, void* stream) {
    return shade_common(ctx, cam, params, 1, shading, rays, out, stream, "bhs_shade_thin_disk");
}

int bhs_shade_thin_disk_batch(bhs_ctx* ctx, const bhs_camera* cams,
                              const bhs_march_params* params, int32_t ncam,
                              const bhs_disk_shading* shading, const bhs_ray_soa* rays,
                              const bhs_shade_soa* out, void* stream) {
    return shade_common(ctx, cams, params, ncam, shading, rays, out, stream,
                        "bhs_shade_thin_disk_batch");
}

float bhs_accumulate_weight(int32_t sample_index, int32_t cap) {
    const int32_t c1 = cap - 1 > 0 ? cap - 1 : 0;
    const int32_t n = sample_index < c1 ? sample_index : c1;
    return sample_index <= 0 ? 1.0f : 1.0f / (float)(n + 1);
}

int bhs_accumulate(bhs_ctx* ctx, const float* cur_rgb_dev, float* acc_rgb_dev, int64_t n,
                   int32_t sample_index, int32_t cap, void* stream) {
    if (!ctx) return BHS_E_INVALID;
    if (n < 0 || (n > 0 && (!cur_rgb_dev || !acc_rgb_dev)))
        return fail(ctx, BHS_E_INVALID, "bhs_accumulate: NULL buffer or n < 0%s");
    if (cap < 1 || sample_index < 0)
        return fail(ctx, BHS_E_INVALID, "bhs_accumulate: need cap >= 1 and sample_index >= 0%s");
    if (n == 0) return BHS_OK;
    DeviceGuard g(ctx->device);
    hipStream_t s = (hipStream_t)stream;
    if (int os = order_stream(ctx, s)) return os;
    hipError_t e;
    if ((e = bhs::launch_accumulate(cur_rgb_dev, acc_rgb_dev, 3 * n,
                                    bhs_accumulate_weight(sample_index, cap), s)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_accumulate launch");
    note_stream(ctx, s);
    return BHS_OK;
}

void bhs_halton_jitter(int32_t sample, float* out2) {
    if (!out2) return;
    out2[0] = out2[1] = 0.0f;
    if (sample <= 0) return;
    for (int k = 0; k < 2; ++k) {
        const int base = k == 0 ? 2 : 3;
        float f = 1.0f, r = 0.0f;
        for (int i = sample; i > 0; i /= base) {
            f /= (float)base;
            r += f * (float)(i % base);
        }
        out2[k] = r - 0.5f;
    }
}

int bhs_sample_sky(bhs_ctx* ctx, const bhs_camera* cam, const bhs_ray_soa* rays, int32_t coarse,
                   const bhs_sky_params* params, const bhs_sky_cube* cube,
                   const float* sky_weight_dev, float* rgb_dev, void* stream) {
    return sky_common(ctx, cam, 1, false, rays, coarse, params, cube, sky_weight_dev, rgb_dev,
                      stream);
}

int bhs_sample_sky_batch(bhs_ctx* ctx, const bhs_camera* cams, int32_t ncam,
                         const bhs_ray_soa* rays, int32_t coarse, const bhs_sky_params* params,
                         const bhs_sky_cube* cube, const float* sky_weight_dev, float* rgb_dev,
                         void* stream) {
    return sky_common(ctx, cams, ncam, true, rays, coarse, params, cube, sky_weight_dev, rgb_dev,
                      stream);
}

int bhs_disk_textures(bhs_ctx* ctx, uint16_t* noise_host, float* lut_host) {
    if (!ctx) return BHS_E_INVALID;
    DeviceGuard g(ctx->device);
    int st;
    if ((st = ensure_disk_tables(ctx)) != BHS_OK) return st;
    hipError_t e;
    if (noise_host &&
        (e = hipMemcpy(noise_host, ctx->noise_vol, kNoiseTexels * sizeof(uint16_t),
                       hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_disk_textures noise");
    if (lut_host &&
        (e = hipMemcpy(lut_host, ctx->bb_lut, (size_t)bhs::shade::kLutSize * 4 * sizeof(float),
                       hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(ctx, e, "bhs_disk_textures table");
    return BHS_OK;
}

}  // extern "C"
