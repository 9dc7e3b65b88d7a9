/*
 * bhs_trace.h — C-ABI of the MI355X Schwarzschild null-geodesic tracer.
 *
 * Plain C, plain pointers and sizes: callable from C, C++ (include/bhs/physics/Geodesic.h
 * builds the reference's bhs::physics API on it), ctypes, or any other FFI.
 *
 * Units: G = c = M = 1 (horizon r = 2, photon sphere r = 3, b_crit = 3*sqrt(3)).
 *
 * Conventions
 *  - Every entry point returns an int status: BHS_OK (0) or a negative bhs_status. No exception
 *    or abort crosses the ABI; bhs_last_error(ctx) holds the text of the last failure.
 *  - Device buffers are caller-owned and named *_dev. Host buffers are named *_host or are plain
 *    struct arguments. The context owns its streams, scratch buffers, graphs and MLP weights.
 *  - bhs_trace_rays is synchronous (like the reference's trace()). Everything taking a `stream`
 *    argument is stream-ordered and asynchronous; `stream` is a hipStream_t (NULL = the null stream).
 *  - The calls of one context stay in issue order across streams: a call on another stream than
 *    the context's previous stream-ordered call first makes its stream wait for that call, since
 *    the context's scratch, camera table and fallback lists are shared by all of them.
 *  - A context may be used by one host thread at a time.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *  - bhs_trace_rays           : bhs::physics::trace, src/physics/Geodesic.h:107-140 (law T),
 *                               batched; TraceResult is src/physics/Geodesic.h:97-103.
 *  - bhs_render / bhs_render_batch :
 *                               the trace pass Renderer::tracePass -> setTraceUniforms ->
 *                               drawFullscreen (src/Renderer.cpp:369-387, 242-356, 154-157) running
 *                               shaders/blackhole.frag:67-253 with DISK_THIN, lensing on, far mode off,
 *                               no bodies, no gas, disk opacity 0 (law S). The GLSL colour output is
 *                               replaced by per-ray geometry in structure-of-arrays form.
 *  - bhs_march_rays           : the same law-S march (blackhole.frag:77-253) on explicit rays.
 *  - bhs_gas_lens_table       : the gas lens table of shaders/gas_lens.comp:11-44 (GasParticles'
 *                               lens program, src/GasParticles.cpp:164-166); bhs_gas_lens_tables:
 *                               several camera radii in one launch.
 *  - bhs_sky_footprint / bhs_sky_footprint_batch :
 *                               the sky footprint of shaders/blackhole.frag:297-306 (GLSL
 *                               dFdx/dFdy of escDir over 2x2 quads, the unlensed footprint at
 *                               status edges) and makeFootprint, shaders/sky.glsl:17-51, as a
 *                               post-pass over a render's ray planes.
 *  - bhs_render_volumetric[_batch] : the trace pass with the volumetric disk
 *                               (blackhole.frag:142-253 DISK_VOLUMETRIC, disk.glsl:81-110).
 *  - bhs_sample_sky[_batch]     : the sky of escaped rays (sky.glsl:57-119, blackhole.frag:297-310)
 *                               added to the colour planes.
 *  - bhs_shade_thin_disk[_batch], bhs_accumulate, bhs_halton_jitter, bhs_disk_textures :
 *                               the thin-disk colour and transparency of blackhole.frag:209-216,234
 *                               (disk.glsl:113-130, blackbody.glsl, noise_bake.comp,
 *                               DiskTextures.cpp:61-93) as a pass over a render's ray planes; the
 *                               accumulation pass (Renderer.cpp:389-410, accumulate.frag) and its
 *                               Halton jitter (Renderer.cpp:16-29).
 *  - bhs_mlp_*, bhs_render_surrogate : no counterpart in the reference (learned surrogate of the
 *                               march, falling back to it near b_crit).
 */
#ifndef BHS_TRACE_H
#define BHS_TRACE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BHS_ABI_VERSION 1

typedef enum bhs_status {
    BHS_OK = 0,
    BHS_E_INVALID = -1,   /* bad argument (NULL, size, range) */
    BHS_E_DEVICE = -2,    /* no such device / HIP runtime error */
    BHS_E_ALLOC = -3,     /* device or host allocation failed */
    BHS_E_LAUNCH = -4,    /* kernel launch or execution failed */
    BHS_E_IO = -5,        /* file read/write failed */
    BHS_E_STATE = -6      /* call not valid in the current state (e.g. no MLP weights) */
} bhs_status;

/* TraceResult::Status (src/physics/Geodesic.h:98), same numeric order. */
enum { BHS_TRACE_ESCAPED = 0, BHS_TRACE_CAPTURED = 1, BHS_TRACE_MAXSTEPS = 2 };

/* Per-ray class written by the law-S entry points. A ray whose march runs out of steps is
 * BHS_CLASS_HORIZON with BHS_FLAG_MAXSTEPS set (blackhole.frag:252). A ray is BHS_CLASS_DISK when at
 * least one equatorial crossing lands in [disk_inner, disk_outer] (opaque-disk geometry). */
enum { BHS_CLASS_HORIZON = 0, BHS_CLASS_DISK = 1, BHS_CLASS_ESCAPE = 2 };

enum {
    BHS_FLAG_MAXSTEPS = 1,    /* march hit max_steps */
    BHS_FLAG_ESCAPED = 2,     /* terminal status was escape (also set for disk rays that escaped) */
    BHS_FLAG_BAND64 = 4,      /* |b - b_crit| < band: the ray was marched in FP64 */
    BHS_FLAG_ROOT64 = 8,      /* disk crossings refined by FP64 root bracketing of the RK4 step */
    BHS_FLAG_SURROGATE = 16   /* outputs come from the MLP surrogate (steps then reads 0); rays of
                                 a surrogate render without it were marched by the integrator and
                                 carry exactly the plain render's record */
};

enum { BHS_PREC_F32 = 0, BHS_PREC_F32_BAND64 = 1, BHS_PREC_F64 = 2 };
enum { BHS_DISK_NONE = 0, BHS_DISK_THIN = 1, BHS_DISK_VOLUMETRIC = 2 };   /* VOLUMETRIC: only
                                 bhs_render_volumetric[_batch] */

/* TraceResult (src/physics/Geodesic.h:97-103). */
typedef struct bhs_trace_result {
    int32_t status;           /* BHS_TRACE_* */
    int32_t steps;
    int32_t equator_crossings;
    int32_t pad_;
    double dir[3];            /* asymptotic direction if escaped, else 0 */
    double min_r;
} bhs_trace_result;

/* One pinhole camera (uniforms uResolution, uJitter, uCamPos, uCamBasis, uTanHalfFov of
 * blackhole.frag:27-31, set by Renderer::setTraceUniforms, src/Renderer.cpp:246-252). */
typedef struct bhs_camera {
    float pos[3];
    float basis[9];           /* column-major 3x3, columns right, up, back (camera looks down -back) */
    float tan_half_fov;       /* tan of half the vertical field of view */
    int32_t width, height;
    float jitter[2];          /* offset added to the pixel centre (uJitter); unused when hashed */
    int32_t jitter_hashed;    /* 1: per-pixel u in [0,1)^2 from (seed, camera_id, pixel, sample)
                                 replaces the pixel centre */
    uint32_t camera_id;
    uint32_t sample;
    uint32_t pad_;
    uint64_t seed;
} bhs_camera;

/* March settings (Settings fields that setTraceUniforms forwards: src/Settings.h:19-37). */
typedef struct bhs_march_params {
    int32_t max_steps;        /* uMaxSteps */
    int32_t disk_mode;        /* BHS_DISK_NONE / BHS_DISK_THIN (uDiskMode) */
    int32_t precision;        /* BHS_PREC_* */
    int32_t root64;           /* 1: refine each in-annulus crossing by FP64 root bracketing */
    float step_scale;         /* uStepScale */
    float disk_step;          /* uDiskStep */
    float disk_inner, disk_outer, disk_thickness;
    float band_rel;           /* FP64 band half-width as a fraction of b_crit (0.02 keeps the FP32
                                 escape direction within 1e-4 rad of FP64 at 1080p) */
    int32_t cross_all;        /* 1: record the first two equatorial crossings at any radius (class
                                 disk then means "crossed the plane"); step control still uses
                                 the disk annulus. Used to generate surrogate training data. */
} bhs_march_params;

/* Structure-of-arrays ray output in device memory; plane k of a multi-plane array starts at
 * base + k * n. Any pointer may be NULL to skip that output. */
typedef struct bhs_ray_soa {
    uint8_t* cls_dev;         /* n   : BHS_CLASS_* */
    uint8_t* flags_dev;       /* n   : BHS_FLAG_* */
    uint8_t* n_cross_dev;     /* n   : in-annulus crossings (saturates at 255) */
    uint16_t* steps_dev;      /* n   : march iterations */
    float* esc_dev;           /* 3n  : escape direction x,y,z planes (0 unless escaped) */
    float* disk_dev;          /* 6n  : r0, r1, phi0, phi1, g0, g1 of the first two crossings */
    float* b_dev;             /* n   : impact parameter b = r sin(alpha) / sqrt(1 - 2/r) */
} bhs_ray_soa;

/* Aggregate counters of one launch (device-side, read after sync). */
typedef struct bhs_stats {
    uint64_t rays;
    uint64_t steps;           /* sum of march iterations */
    uint64_t crossings;       /* sum of in-annulus crossings */
    uint64_t escapes;         /* rays that escaped (far-field correction applied) */
    uint64_t band64;          /* rays marched in FP64 */
    uint64_t fallback;        /* surrogate rays sent back to the integrator */
    uint64_t surrogate;       /* rays served by the surrogate network */
    uint64_t pad_;
} bhs_stats;

int bhs_version(void);

/* Context on HIP device `device`. Returns NULL on failure with *status set. */
typedef struct bhs_ctx bhs_ctx;
bhs_ctx* bhs_create(int device, int* status);
void bhs_destroy(bhs_ctx* ctx);
const char* bhs_last_error(const bhs_ctx* ctx);
int bhs_sync(bhs_ctx* ctx);

/* Batched bhs::physics::trace (src/physics/Geodesic.h:107-140), FP64, law T. Host arrays, n rays,
 * pos3/dir3 interleaved xyz. Synchronous. */
int bhs_trace_rays(bhs_ctx* ctx, const double* pos3_host, const double* dir3_host, int64_t n,
                   int32_t max_steps, double step_scale, double esc_tol, double r_esc_min,
                   bhs_trace_result* out_host);

/* Law-S render of one camera: width*height rays, pixel p = row*width + col with row 0 at the
 * bottom (gl_FragCoord convention). Stream-ordered. */
int bhs_render(bhs_ctx* ctx, const bhs_camera* cam, const bhs_march_params* params,
               const bhs_ray_soa* out, void* stream);

/* bhs_render that allocates nothing: no counterpart in the reference. For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, kernels and results as bhs_render, but the render's scratch (stats
 * and counters, the FP64 band list, FP64 root states with root64) lives in a static device array
 * of the library, BHS_NOALLOC_SCRATCH_BYTES per GPU (held by each GPU the library runs on until
 * the process ends), instead of memory the context allocates. A render needs 8,448 bytes plus 4 per
 * work item (its image padded to whole 16x4-pixel tiles) plus, with root64, 56 per ray; larger renders
 * are rejected with BHS_E_INVALID (a 3840x2160 frame fits without root64). Every no-alloc call on
 * a GPU shares that GPU's array: calls that overlap in time on one GPU (other streams or other
 * contexts) race and give wrong results. bhs_last_stats reads the counters of the latest call.
 * Stream-ordered; timing (bhs_set_timing) does not record it. */
#define BHS_NOALLOC_SCRATCH_BYTES (48u << 20)
int bhs_render_noalloc(bhs_ctx* ctx, const bhs_camera* cam, const bhs_march_params* params,
                       const bhs_ray_soa* out, void* stream);

/* Law-S render of up to 64 cameras (each with its own size and march settings, FP32 or FP32 with
 * the FP64 band) in one persistent launch: camera k's rays occupy out[offset_k, offset_k + w_k h_k)
 * with offset_k = sum of w_j h_j over j < k, each camera in the pixel order of bhs_render. The
 * camera table is uploaded with the launch (stream-ordered). */
int bhs_render_batch(bhs_ctx* ctx, const bhs_camera* cams, const bhs_march_params* params,
                     int32_t ncam, const bhs_ray_soa* out, void* stream);

/* Sky footprint planes (per pixel, caller-owned device buffers; any may be NULL). Three-component
 * fields are 3 planes of n floats, like esc_dev. */
typedef struct bhs_footprint_soa {
    float* ddx_dev;           /* 3n : dFdx of escDir; of the view direction where the quad falls back */
    float* ddy_dev;           /* 3n : dFdy, likewise */
    float* e1_dev;            /* 3n : Footprint.e1, orthonormal footprint basis (sky.glsl:19) */
    float* e2_dev;            /* 3n : Footprint.e2 */
    float* rinv_dev;          /* 3n : Footprint r11, r12, r22, the inverse R factor (sky.glsl:20) */
    float* area_dev;          /* n  : la * lbp = sqrt(Footprint.det), the footprint solid angle */
    float* mu_dev;            /* n  : Footprint.mu = min(uPixelOmega0 / area, 1e4) (sky.glsl:49) */
    uint8_t* fallback_dev;    /* n  : 1 where the unlensed derivatives were used at a status edge,
                                 2 where that edge runs through a helper pixel beyond an odd image
                                 edge (the reference marches helper pixels; this pass counts them
                                 as not escaped), 0 elsewhere */
} bhs_footprint_soa;

/* Sky footprint post-pass over the ray planes of a bhs_render of `cam` (reads rays->flags_dev and
 * rays->esc_dev, which must be set; n = width * height). Per pixel: the derivatives of escDir
 * (its escape direction if BHS_FLAG_ESCAPED, else its view direction) across its 2x2 quad, or those
 * of the view direction where the quad's pair mixes escaped and non-escaped pixels
 * (blackhole.frag:297-306), and for escaped pixels makeFootprint(escDir, ddx, ddy, uPixelOmega0)
 * (sky.glsl:27-51; zeros elsewhere). coarse: 0 fine derivatives (each pixel's own row and column
 * pair), 1 coarse (the quad's first row and column). Pixels beyond an odd edge count as not
 * escaped (fallback 2 where that decides the fallback). Stream-ordered. */
int bhs_sky_footprint(bhs_ctx* ctx, const bhs_camera* cam, const bhs_ray_soa* rays,
                      int32_t coarse, const bhs_footprint_soa* out, void* stream);
/* The same over the planes of a bhs_render_batch / bhs_render_surrogate_batch of ncam <= 64
 * cameras (camera k at the prefix sum of width * height), in one launch. */
int bhs_sky_footprint_batch(bhs_ctx* ctx, const bhs_camera* cams, int32_t ncam,
                            const bhs_ray_soa* rays, int32_t coarse,
                            const bhs_footprint_soa* out, void* stream);

/* ------------------------------------------------------------------------------------------------
 * Thin-disk shading and transparency, accumulation of jittered frames.
 *
 * The colour half of the trace pass for the thin disk (shaders/blackhole.frag:209-216,234 with
 * thinDiskSample, shaders/disk.glsl:113-130; blackbody, shaders/blackbody.glsl; the noise texture of
 * shaders/noise_bake.comp and the blackbody table of src/DiskTextures.cpp:61-93), over the ray
 * planes of a bhs_render / bhs_render_batch / bhs_render_surrogate[_batch] of the same cameras and
 * march parameters. Every in-annulus crossing is composited front to back (colour += T rgb alpha,
 * T *= 1 - alpha) until T < 0.004, where the reference's march stops (no sky). The recorded first
 * two crossings are used as recorded; a ray with more crossings that is still transparent after
 * two is marched again from its pixel (same precision as rendered) for the rest.
 * ------------------------------------------------------------------------------------------------ */

/* Disk shading uniforms (Settings fields forwarded by Renderer::setTraceUniforms,
 * src/Renderer.cpp:328-343; defaults of src/Settings.h:38-45: density 1, opacity 2, peak_temp 6000,
 * brightness 5.5, noise_scale 1, noise_strength 0.85, warmth 0.4, doppler 0, time 0, inflow 0).
 * The annulus comes from the march parameters. */
typedef struct bhs_disk_shading {
    float density;            /* uDiskDensity */
    float opacity;            /* uDiskOpacity */
    float peak_temp;          /* uDiskPeakTemp (K at the hottest ring) */
    float brightness;         /* uDiskBrightness */
    float noise_scale;        /* uDiskNoiseScale, in [0, 1000] */
    float noise_strength;     /* uDiskNoiseStrength */
    float warmth;             /* uDiskWarmth */
    float doppler;            /* uDopplerStrength (0 film look, 1 physical) */
    float time;               /* uDiskTime (flow clock in M) */
    float inflow;             /* uDiskInflow */
} bhs_disk_shading;

/* Shading output planes (caller-owned device buffers). */
typedef struct bhs_shade_soa {
    float* rgb_dev;           /* 3n : linear radiance of the disk crossings, composited */
    float* sky_weight_dev;    /* n  : transmittance left for the sky (0 unless the ray escaped and
                                 the disk did not make it opaque): pixel = rgb + sky_weight * sky */
} bhs_shade_soa;

#define BHS_NOISE_SIZE 128    /* noise volume: 128^3 half-float texels, x fastest */
#define BHS_BB_LUT_SIZE 256   /* blackbody table: 256 x (r, g, b, luminance / luminance(6500 K)) */

/* Shading pass over the planes of a render of `cam` (needs flags, n_cross and disk planes; n =
 * width * height). Stream-ordered; the first call of a context bakes its disk tables. */
int bhs_shade_thin_disk(bhs_ctx* ctx, const bhs_camera* cam, const bhs_march_params* params,
                        const bhs_disk_shading* shading, const bhs_ray_soa* rays,
                        const bhs_shade_soa* out, void* stream);
/* The same over a batch of ncam <= 64 cameras laid out as bhs_render_batch (one annulus). */
int bhs_shade_thin_disk_batch(bhs_ctx* ctx, const bhs_camera* cams,
                              const bhs_march_params* params, int32_t ncam,
                              const bhs_disk_shading* shading, const bhs_ray_soa* rays,
                              const bhs_shade_soa* out, void* stream);

/* Volumetric disk render (shaders/blackhole.frag:142-253 with uDiskMode = DISK_VOLUMETRIC and
 * shaders/disk.glsl:81-110, the High and Ultra presets' disk): law S with the volumetric slab rule
 * (steps inside the slab capped at 6 disk_step), emission and absorption (volumetricDiskStep,
 * diskDensity with the vertical profile, noise and puffing) sampled at ceil(seg / disk_step)
 * points of the cubic Hermite interpolant of every RK4 step that starts inside the slab, and the
 * march stopped where the transmittance falls below 0.004. params[k].disk_mode must be
 * BHS_DISK_VOLUMETRIC (root64 and cross_all 0; FP64 band rule as in bhs_render; BHS_PREC_F64 for
 * a single camera only). out receives the colour and sky weight; rays (may be NULL, any plane
 * NULL) receives class (horizon / disk = absorbed / escape), flags, steps, escape direction and
 * b, with n_cross and the disk planes 0. Stream-ordered. */
int bhs_render_volumetric(bhs_ctx* ctx, const bhs_camera* cam, const bhs_march_params* params,
                          const bhs_disk_shading* shading, const bhs_ray_soa* rays,
                          const bhs_shade_soa* out, void* stream);
int bhs_render_volumetric_batch(bhs_ctx* ctx, const bhs_camera* cams,
                                const bhs_march_params* params, int32_t ncam,
                                const bhs_disk_shading* shading, const bhs_ray_soa* rays,
                                const bhs_shade_soa* out, void* stream);

/* Sky (shaders/sky.glsl:57-119, sampleSky and starLayer; added at shaders/blackhole.frag:297-310):
 * the uniforms of the sky and the backdrop cube map (src/SkyBaker.cpp:83-91: a mip-mapped cube map,
 * GL_LINEAR_MIPMAP_LINEAR, seamless, src/App.cpp:216). */
typedef struct bhs_sky_params {
    float sky_brightness;     /* uSkyBrightness (Settings.h: 1) */
    float star_density;       /* uStarDensity (1) */
    float star_brightness;    /* uStarBrightness (1) */
    float doppler;            /* uDopplerStrength (0 film look, 1 physical) */
} bhs_sky_params;

#define BHS_SKY_MAX_LEVELS 16
/* Backdrop cube map (device memory, read only): level l in 0..levels-1 has faces of
 * S_l = max(1, res >> l) texels; texel (i, j) of face f (order +X, -X, +Y, -Y, +Z, -Z; s = (i +
 * 0.5) / S_l, t = (j + 0.5) / S_l as OpenGL's glTexImage2D rows) at float4 index
 * off_l + (f * S_l + j) * S_l + i, off_l = 6 (S_0^2 + ... + S_(l-1)^2); each float4 holds r, g, b
 * and an unused fourth value. The reference stores R11G11B10F texels; convert them to float. */
typedef struct bhs_sky_cube {
    const float* texels_dev;
    int32_t res;              /* face size of level 0, >= 1 */
    int32_t levels;           /* 1..BHS_SKY_MAX_LEVELS */
} bhs_sky_cube;

/* Sky pass over the planes of a render of `cam` (reads rays->flags_dev and rays->esc_dev; n =
 * width * height): for every escaped pixel with weight w != 0 (w = sky_weight_dev[i], the shading
 * pass's transmittance, or 1 where sky_weight_dev is NULL),
 *   rgb_dev[i + k n] += w * sampleSky(escDir, dFdx, dFdy, uPixelOmega0, rcam)[k]   (k = r, g, b)
 * with the derivatives and footprint of bhs_sky_footprint (coarse as there) and rcam = |cam->pos|.
 * cube may be NULL (no backdrop: the star layers only). The cube lookup is the OpenGL
 * specification's exact textureGrad (level of detail log2 of the larger screen-axis footprint in
 * level-0 texels, trilinear, seamless edges and corners). Other pixels are left unchanged.
 * Stream-ordered; the first call of a context bakes its blackbody table. */
int bhs_sample_sky(bhs_ctx* ctx, const bhs_camera* cam, const bhs_ray_soa* rays, int32_t coarse,
                   const bhs_sky_params* params, const bhs_sky_cube* cube,
                   const float* sky_weight_dev, float* rgb_dev, void* stream);
/* The same over the planes of a batch render of ncam <= 64 cameras (camera k at the prefix sum of
 * width * height), in one launch. */
int bhs_sample_sky_batch(bhs_ctx* ctx, const bhs_camera* cams, int32_t ncam,
                         const bhs_ray_soa* rays, int32_t coarse, const bhs_sky_params* params,
                         const bhs_sky_cube* cube, const float* sky_weight_dev, float* rgb_dev,
                         void* stream);

/* Progressive supersampling (Renderer::accumulatePass, src/Renderer.cpp:389-410, and
 * shaders/accumulate.frag:8-11): acc = w >= 1 ? cur : acc (1 - w) + cur w over 3n floats (rgb
 * planes), w = bhs_accumulate_weight(sample_index, cap) = 1 for sample 0, else
 * 1 / (min(sample_index, cap - 1) + 1). Stream-ordered. */
float bhs_accumulate_weight(int32_t sample_index, int32_t cap);
int bhs_accumulate(bhs_ctx* ctx, const float* cur_rgb_dev, float* acc_rgb_dev, int64_t n,
                   int32_t sample_index, int32_t cap, void* stream);
/* The sub-pixel jitter of sample `sample` (jitterFor, src/Renderer.cpp:16-29): (halton(s, 2) - 0.5,
 * halton(s, 3) - 0.5) in float, (0, 0) for sample 0; for bhs_camera.jitter. */
void bhs_halton_jitter(int32_t sample, float* out2);
/* Host copies of the context's disk tables (bakes them first if needed); either may be NULL:
 * noise_host 128^3 uint16 half-float texels, lut_host 256 x 4 floats. Synchronous. */
int bhs_disk_textures(bhs_ctx* ctx, uint16_t* noise_host, float* lut_host);

/* Gas lens table (shaders/gas_lens.comp:11-44 with lensAlpha, shaders/gas_lens.glsl:13-23; built
 * at src/GasParticles.cpp:164-166): n_rays fan rays (local angle alpha_i = lensAlpha((i + 0.5) /
 * n_rays, alpha_crit) from the inward radial) leave a static camera at (cam_r, 0, 0) turning
 * towards +y; each is marched with law S's float setup and RK4 at ds = (r < 8 ? 0.004 : 0.01) r for
 * at most BHS_LENS_MAX_STEPS steps. table_dev: n_phis rows of n_rays floats; row j holds log r at
 * swept angle j * 2 pi / (n_phis - 1) (linear in r between steps, at least log 2; log 2 once the
 * ray fell in, log BHS_LENS_RMAX once it escaped). The reference's sizes are 1024 x 1024 and
 * alpha_crit = shadowAngularRadius(cam_r). Stream-ordered. */
#define BHS_LENS_RMAX 2.0e5f
#define BHS_LENS_MAX_STEPS 20000
int bhs_gas_lens_table(bhs_ctx* ctx, float cam_r, float alpha_crit, int32_t n_rays,
                       int32_t n_phis, float* table_dev, void* stream);
/* n_tables <= 64 such tables of the same size in one launch (no counterpart in the reference,
 * which builds one table per camera radius, src/GasParticles.cpp:164-166; for callers with several
 * cameras, e.g. the poses of a batch): table k, for cam_r[k] and alpha_crit[k] (host arrays), at
 * tables_dev + k * n_phis * n_rays, equal to bhs_gas_lens_table's. One table's fan is n_rays / 64
 * waves, a sixteenth of the chip at 1024 rays; the batch fills it. Stream-ordered. */
int bhs_gas_lens_tables(bhs_ctx* ctx, int32_t n_tables, const float* cam_r,
                        const float* alpha_crit, int32_t n_rays, int32_t n_phis,
                        float* tables_dev, void* stream);

/* The same batch captured once into a HIP graph and replayed: bhs_graph_create captures (and
 * instantiates) the whole launch sequence for these cameras, settings and output buffers with the
 * context's current tuning; bhs_graph_launch replays it on `stream`; bhs_graph_update replaces the
 * camera poses/settings (sizes and band/root choices must stay) before the next replay;
 * bhs_graph_stats reads the counters of the last replay (synchronises). The graph owns its camera
 * table and scratch; `out` must stay allocated while the graph is used. */
typedef struct bhs_graph bhs_graph;
int bhs_graph_create(bhs_ctx* ctx, const bhs_camera* cams, const bhs_march_params* params,
                     int32_t ncam, const bhs_ray_soa* out, bhs_graph** graph);
int bhs_graph_update(bhs_ctx* ctx, bhs_graph* graph, const bhs_camera* cams,
                     const bhs_march_params* params);
int bhs_graph_launch(bhs_ctx* ctx, bhs_graph* graph, void* stream);
int bhs_graph_stats(bhs_ctx* ctx, bhs_graph* graph, bhs_stats* out_host);
void bhs_graph_destroy(bhs_graph* graph);

/* Law-S march of explicit rays: pos3_dev/dir3_dev are device float arrays (xyz interleaved).
 * Stream-ordered. */
int bhs_march_rays(bhs_ctx* ctx, const float* pos3_dev, const float* dir3_dev, int64_t n,
                   const bhs_march_params* params, const bhs_ray_soa* out, void* stream);

/* Aggregate counters of the most recent bhs_render / bhs_march_rays / bhs_render_batch call
 * (synchronises the context's last stream). */
int bhs_last_stats(bhs_ctx* ctx, bhs_stats* out_host);

/* The float march's reciprocal square root over n device floats: y_dev[i] = 1/sqrt(x_dev[i])
 * rounded once to float (IEEE special values: +-0 -> +-inf, +inf -> +0, negative -> NaN). No
 * counterpart in the reference (the GLSL march's inversesqrt has no exact rounding): it exposes the
 * arithmetic primitive the march's results depend on, so a caller or a test can check it against
 * any correctly rounded implementation. Stream-ordered. */
int bhs_rsqrt_rn(bhs_ctx* ctx, const float* x_dev, float* y_dev, int64_t n, void* stream);

/* VALU issue-rate probe (no counterpart in the reference; the measurement behind the VALU
 * rooflines): waves_per_simd resident waves on every SIMD each run `iters` rounds of 16
 * independent instructions of one class (kind 0 v_fma_f32, 1 v_mul_f32, 2 v_fma_f64, 3 v_xor_b32 +
 * v_add_u32, 4 v_rsq_f32). Synchronous; *ms the kernel's time (HIP events), *wave_instr the
 * wave64 instructions of the class it issued. */
int bhs_valu_probe(bhs_ctx* ctx, int32_t kind, int32_t waves_per_simd, int32_t iters, double* ms,
                   double* wave_instr);

/* Launch shape of the FP32 march. mode 0: one lane per pixel on a static grid of 8x8-pixel wave
 * tiles; mode 1: resident waves pull rays from a work queue and refill finished lanes: with
 * refill_min 64, whole 16x4-pixel tiles from a queue split into 64 ranges with their own counters;
 * below 64, the idle lanes with one atomic per refill on a single counter (wave ballot + prefix
 * count). A context starts with {1, 256, 6, 64}; chosen per context by bhs_autotune or set by hand;
 * bit-identical results in every shape. */
typedef struct bhs_tuning {
    int32_t mode;             /* 0 static tiles, 1 persistent work queue */
    int32_t block;            /* threads per block: 64 or 256 */
    int32_t waves_per_simd;   /* mode 1: resident waves per SIMD, 1..8 */
    int32_t refill_min;       /* mode 1: refill when at least this many lanes are idle, 1..64 */
} bhs_tuning;

int bhs_set_tuning(bhs_ctx* ctx, const bhs_tuning* t);
int bhs_get_tuning(bhs_ctx* ctx, bhs_tuning* t);

/* Project autotuner: times bhs_render of `cam` over a fixed candidate set of launch shapes (each
 * `reps` times after one warm-up, median of the main-kernel GPU time), keeps the fastest in the
 * context and returns it in *best (and its median ms in *best_ms if non-NULL). `out` receives the
 * renders. Synchronous. */
int bhs_autotune(bhs_ctx* ctx, const bhs_camera* cam, const bhs_march_params* params,
                 const bhs_ray_soa* out, int32_t reps, bhs_tuning* best, double* best_ms);

/* GPU timers (the trace-pass GL_TIME_ELAPSED timer of Renderer::render, src/Renderer.cpp:159-178).
 * While timing is on, every launch records HIP events on its own stream around its main march
 * kernel and around the whole call. bhs_timing waits for them and returns the averages (ms) over
 * the launches recorded since the previous bhs_timing call, then starts a new window. */
int bhs_set_timing(bhs_ctx* ctx, int on);
int bhs_timing(bhs_ctx* ctx, double* march_ms_avg, double* total_ms_avg, int32_t* launches);

/* ------------------------------------------------------------------------------------------------
 * Learned surrogate of the law-S march (no counterpart in the reference).
 *
 * A 3 -> 128 -> 128 -> 128 -> 5 SiLU network maps (log r_obs, b/b_crit, psi) of a ray to the residual
 * of its escape angle over the exact Schwarzschild swept angle, the radii of its first two
 * equatorial crossings and the logits of reaching them. A model is bound to the march parameters
 * it was trained for (step control and disk; max_steps, precision, root64 and band_rel may differ
 * at render time). Training data come from the FP64 march itself; training runs on the GPU
 * (bf16 MFMA forward/backward, FP32 accumulation, Adam in FP32); bhs_mlp_validate compares the
 * decoded outputs with FP64 marches of fresh rays and builds the error map that decides, per
 * (r_obs, b) cell, where the network may serve rays. bhs_render_surrogate serves the rays of
 * served cells whose outcome is decided (reach, annulus edges, step budget) and sends all others
 * back to the integrator (FP64 band rule of `params`, FP32 elsewhere) in the same call.
 * ------------------------------------------------------------------------------------------------ */
typedef struct bhs_mlp bhs_mlp;

#define BHS_MLP_PARAMS 34181

typedef struct bhs_mlp_info {
    bhs_march_params trained_for;
    int64_t dataset_rays;
    int64_t adam_steps;
    float tol_phi;            /* escape-direction tolerance of the error map (rad) */
    float radius_margin;      /* relative margin around the annulus edges */
    float logit_margin;       /* |logit| below this: reach undecided */
    float steps_slack;        /* served only if the training step maximum * slack < max_steps */
    uint32_t steps_max[16];   /* training step maximum per log r_obs bin over [8, 1000] (the
                                 largest over b; the model keeps a (r_obs, b) table) */
    int32_t served_cells;     /* error-map cells (of 32 x 64) that serve rays */
    int32_t pad_;
} bhs_mlp_info;

typedef struct bhs_mlp_validation {
    int64_t rays;             /* validation rays */
    int64_t in_domain;        /* rays in the surrogate's domain */
    int64_t served;           /* rays the render would serve (error map applied) */
    int64_t escaping;         /* served rays that escape */
    int64_t mismatches;       /* served rays whose rendered record disagrees with FP64 */
    float phi_err_max, phi_err_p99, phi_err_p50;   /* escape-direction error (rad), served */
    float r_err_max, r_err_p99, r_err_p50;         /* relative crossing-radius error, served */
    float radius_margin;      /* margin chosen from the radius errors */
    int32_t served_cells;
} bhs_mlp_validation;

/* New model (He-initialised from `seed`) for the march parameters `trained_for`. */
bhs_mlp* bhs_mlp_create(bhs_ctx* ctx, const bhs_march_params* trained_for, uint32_t seed,
                        int* status);
void bhs_mlp_destroy(bhs_mlp* m);
/* Training set: n_rays rays sampled from `seed` (r_obs log-uniform in [8, 1000], b log-spaced on
 * both sides of b_crit), marched in FP64 with refined crossings recorded at any radius. */
int bhs_mlp_make_dataset(bhs_ctx* ctx, bhs_mlp* m, int64_t n_rays, uint32_t seed);
/* `steps` Adam steps on minibatches of `batch` dataset rays, learning rate cosine-annealed from
 * lr_start to lr_end. loss_host (steps floats) receives the minibatch losses if non-NULL.
 * Synchronous. */
int bhs_mlp_train(bhs_ctx* ctx, bhs_mlp* m, int32_t steps, int32_t batch, float lr_start,
                  float lr_end, float* loss_host);
/* Loss and gradient (BHS_MLP_PARAMS floats) on n explicit rows: x_host n x 3 features, y_host
 * n x 5 targets, mask_host n x 5 masks. Parameters unchanged. Synchronous. */
int bhs_mlp_grads(bhs_ctx* ctx, bhs_mlp* m, const float* x_host, const float* y_host,
                  const float* mask_host, int32_t n, float* grad_host, float* loss_host);
/* Raw network outputs (n x 5) of n feature rows (n x 3), device arrays, MFMA inference. */
int bhs_mlp_forward(bhs_ctx* ctx, bhs_mlp* m, const float* x_dev, int64_t n, float* y_dev,
                    void* stream);
/* Validation on n_rays fresh rays from `seed`; sets the radius margin and the error map (cells
 * whose served rays all stay within tol_phi and render the FP64 record). Synchronous. */
int bhs_mlp_validate(bhs_ctx* ctx, bhs_mlp* m, int64_t n_rays, uint32_t seed, float tol_phi,
                     bhs_mlp_validation* out_host);
int bhs_mlp_get_params(bhs_ctx* ctx, bhs_mlp* m, float* host, int64_t n);
int bhs_mlp_set_params(bhs_ctx* ctx, bhs_mlp* m, const float* host, int64_t n);
int bhs_mlp_get_info(bhs_ctx* ctx, bhs_mlp* m, bhs_mlp_info* out_host);
/* Model file: parameters, error map, margins, step table and the march parameters it serves. */
int bhs_mlp_save(bhs_ctx* ctx, bhs_mlp* m, const char* path);
bhs_mlp* bhs_mlp_load(bhs_ctx* ctx, const char* path, int* status);
/* Table mode (n2, n3 >= 2; n2 = 0 returns to per-ray inference): every bhs_render_surrogate then
 * evaluates the network once per node of an n2 x n3 grid over (b/b_crit, psi) at the camera's r_obs,
 * spanning the camera's in-domain rays, and each ray interpolates the outputs bilinearly. */
int bhs_mlp_set_table(bhs_ctx* ctx, bhs_mlp* m, int32_t n2, int32_t n3);
/* Surrogate render of one camera with integrator fallback, outputs as bhs_render (steps = 0 and
 * BHS_FLAG_SURROGATE on served rays). Stream-ordered. With timing on, the main-kernel interval is
 * the network kernel and the call interval includes the fallback march. */
int bhs_render_surrogate(bhs_ctx* ctx, bhs_mlp* m, const bhs_camera* cam,
                         const bhs_march_params* params, const bhs_ray_soa* out, void* stream);
/* Surrogate render of ncam <= 64 cameras in one call, laid out as bhs_render_batch (camera k's
 * rays start at the prefix sum of width * height; no BHS_PREC_F64). The FP64 band and the
 * fallback marches of all cameras run together, so their serial tails are paid once per call. */
int bhs_render_surrogate_batch(bhs_ctx* ctx, bhs_mlp* m, const bhs_camera* cams,
                               const bhs_march_params* params, int32_t ncam,
                               const bhs_ray_soa* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* BHS_TRACE_H */
