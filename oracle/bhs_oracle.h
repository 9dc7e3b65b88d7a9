/*
 * bhs_oracle.h — CPU oracle of the Schwarzschild null-geodesic tracer.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker of the HIP product path: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. Nothing in the product
 * (blackholerender_amd/, include/) links or calls it.
 *
 * It restates, in plain C, two control laws of the reference:
 *   law T  bhs::physics::trace and its helpers, src/physics/Geodesic.h:15-140 (FP64, GLM operation
 *          order, no fused multiply-add: the reference's CPU build has no FMA contraction);
 *   law S  the thin-disk march of shaders/blackhole.frag:67-253 (DISK_THIN, lensing on, far mode off,
 *          no bodies, no gas, disk opacity 0), with geodesic.glsl and disk.glsl:24,33-42,113-117.
 *          Law S is written once (lawS_body.inc) and compiled twice, in float and in double, with one
 *          pinned operation order (explicit fma, IEEE division and square root, no contraction) so a
 *          device implementation with the same order can be compared bit for bit.
 *
 * Pins (tests/test_oracle_*.py): the 15 checks of tests/geodesic_tests.cpp restated; closed forms
 * (impact parameter, shadow radius, redshift); exact Schwarzschild deflection by quadrature; RK4
 * step-halving order; orbit-plane invariance; float vs double law S agreement; brute-force
 * convergence of class on tiny images.
 */
#ifndef BHS_ORACLE_H
#define BHS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- law T: src/physics/Geodesic.h ------------------------------------------------------------ */
typedef struct oracle_trace_result {
    int32_t status; /* 0 escaped, 1 captured, 2 max steps (Geodesic.h:98) */
    int32_t steps;
    int32_t equator_crossings;
    int32_t pad_;
    double dir[3];
    double min_r;
} oracle_trace_result;

double oracle_critical_impact_parameter(void);                                     /* :19     */
void oracle_accel(const double x[3], double h2, double out[3]);                    /* :21-25  */
void oracle_rk4_step(double x[3], double v[3], double h2, double dt);               /* :27-34  */
void oracle_initial_velocity(const double cam[3], const double dir[3], double out[3]); /* :40-47 */
double oracle_impact_parameter(const double cam[3], const double dir[3]);          /* :50-54  */
double oracle_impact_parameter_about_axis(const double cam[3], const double dir[3],
                                          const double axis[3]);                   /* :58-61  */
void oracle_far_field_direction(const double x[3], const double v[3], double h2,
                                double out[3]);                                    /* :65-75  */
double oracle_shadow_angular_radius(double r);                                     /* :78-82  */
double oracle_kepler_omega(double r);                                              /* :85     */
double oracle_redshift_factor(double r_emit, double b_axis, double r_cam);         /* :90-95  */
void oracle_trace_T(const double cam[3], const double dir[3], int32_t max_steps, double step_scale,
                    double esc_tol, double r_esc_min, oracle_trace_result* out);   /* :107-140 */
/* n rays, pos3/dir3 interleaved xyz; `threads` worker threads (<=0: 1). */
void oracle_trace_T_batch(const double* pos3, const double* dir3, int64_t n, int32_t max_steps,
                          double step_scale, double esc_tol, double r_esc_min,
                          oracle_trace_result* out, int32_t threads);

/* ---- law S: shaders/blackhole.frag ------------------------------------------------------------ */
typedef struct oracle_lawS_params {
    int32_t max_steps, disk_mode; /* disk_mode 0 none, 1 thin */
    int32_t root64;               /* 1: FP64 root bracketing of each in-annulus crossing */
    int32_t cross_all;            /* 1: record equatorial crossings at any radius (step control
                                     still uses the disk annulus) */
    float step_scale, disk_step, disk_inner, disk_outer, disk_thickness;
    float band_rel;               /* band half-width / b_crit used by oracle_lawS_band */
} oracle_lawS_params;

/* One ray's law-S result; float fields hold the value the march produced (float march: exact;
 * double march: rounded to float once, at the end). */
typedef struct oracle_lawS_ray {
    int32_t cls;      /* 0 horizon, 1 disk, 2 escape */
    int32_t flags;    /* 1 max steps, 2 escaped, 4 marched in FP64, 8 root-refined crossings */
    int32_t steps;
    int32_t n_cross;
    float esc[3];
    float disk_r[2], disk_phi[2], disk_g[2];
    float b;          /* impact parameter r sin(alpha)/sqrt(1-2/r) of the march's own precision */
} oracle_lawS_ray;

/* Ray generation (blackhole.frag:68-72 with Renderer.cpp:246-252 and common.glsl:28-32 pcg for the
 * hashed jitter). basis is column-major (right, up, back). Writes the float view direction. */
void oracle_ray_dir(const float pos[3], const float basis[9], float tan_half_fov, int32_t width,
                    int32_t height, const float jitter[2], int32_t jitter_hashed, uint64_t seed,
                    uint32_t camera_id, uint32_t sample, int64_t pixel, float dir_out[3]);
uint32_t oracle_pcg(uint32_t v);
/* As oracle_ray_dir for an explicit column and row (which may lie outside the image; the hashed
 * jitter then uses pixel = row * width + col). */
void oracle_ray_dir_at(const float pos[3], const float basis[9], float tan_half_fov, int32_t width,
                       int32_t height, const float jitter[2], int32_t jitter_hashed, uint64_t seed,
                       uint32_t camera_id, uint32_t sample, int64_t pixel, int64_t col, int64_t row,
                       float dir_out[3]);

/* ---- sky footprint: shaders/blackhole.frag:297-306, shaders/sky.glsl:17-51 -------------------- */
typedef struct oracle_footprint {
    float e1[3], e2[3];   /* orthonormal footprint basis in sky space */
    float rinv[3];        /* r11, r12, r22 of the inverse R factor */
    float area;           /* la * lbp = sqrt(det) */
    float mu;             /* min(omega0 / area, 1e4) */
} oracle_footprint;
void oracle_make_footprint(const float d[3], const float a[3], const float b[3], float omega0,
                           oracle_footprint* f);
/* Footprint image of one camera's ray planes (flags: n bytes, bit 2 = escaped; esc: 3 planes of n
 * floats). Outputs are planes of n floats (ddx, ddy, e1, e2, rinv: 3 planes), any may be NULL;
 * non-escaped pixels get derivatives and a zero footprint. coarse: 0 fine derivatives, 1 coarse. */
void oracle_sky_footprint(const float pos[3], const float basis[9], float tan_half_fov,
                          int32_t width, int32_t height, const float jitter[2],
                          int32_t jitter_hashed, uint64_t seed, uint32_t camera_id,
                          uint32_t sample, int32_t coarse, const uint8_t* flags, const float* esc,
                          float* ddx, float* ddy, float* e1, float* e2, float* rinv, float* area,
                          float* mu, uint8_t* fallback);

/* Law S in float and in double (same operation order). Inputs are the float camera position and
 * float direction; the double march promotes them exactly. */
void oracle_lawS_f32(const float pos[3], const float dir[3], const oracle_lawS_params* p,
                     oracle_lawS_ray* out);
void oracle_lawS_f64(const float pos[3], const float dir[3], const oracle_lawS_params* p,
                     oracle_lawS_ray* out);
/* Float march with the FP64 band: rays whose float impact parameter b satisfies
 * |b - b_crit| < band_rel * b_crit are marched in double instead (flag 4). */
void oracle_lawS_band(const float pos[3], const float dir[3], const oracle_lawS_params* p,
                      oracle_lawS_ray* out);
/* The float setup's h = |x0 x v0| and impact parameter b for a ray. */
float oracle_lawS_h_f32(const float pos[3], const float dir[3]);
float oracle_lawS_b_f32(const float pos[3], const float dir[3]);

/* Image of one camera: rays [first, first+count) of width*height pixels, `mode` 0 float, 1 double,
 * 2 float with FP64 band; out[count]. */
void oracle_render(const float pos[3], const float basis[9], float tan_half_fov, int32_t width,
                   int32_t height, const float jitter[2], int32_t jitter_hashed, uint64_t seed,
                   uint32_t camera_id, uint32_t sample, const oracle_lawS_params* p, int32_t mode,
                   int64_t first, int64_t count, int64_t stride, oracle_lawS_ray* out,
                   int32_t threads);

/* Law S on n explicit rays (pos3/dir3 float xyz interleaved), `mode` 0 float, 1 double, 2 float
 * with FP64 band, `threads` workers. */
void oracle_march_batch(const float* pos3, const float* dir3, int64_t n, const oracle_lawS_params* p,
                        int32_t mode, oracle_lawS_ray* out, int32_t threads);

/* 1/sqrt(x) correctly rounded to float: the reciprocal square root of the float march. */
float oracle_rsqrtf(float x);
/* 1/sqrt(x) of the double march: oracle_rsqrtf((float)x) refined by one step in double. */
double oracle_rsqrt_d(double x);

/* Atan2 built from IEEE operations only (range reduction + degree-4 polynomial in t^2); identical
 * in any IEEE-754 float implementation that honours fmaf. */
float oracle_atan2f(float y, float x);

/* ---- gas lens table: shaders/gas_lens.comp:11-44, shaders/gas_lens.glsl:13-23 -------------------
 * table: n_phis rows of n_rays floats; row j holds log r of every fan ray at swept angle
 * j * 2 pi / (n_phis - 1) (log 2 once fallen in, log r_max once escaped). The reference's sizes:
 * 1024 x 1024, max_steps 20000, r_max 2e5; alpha_crit = shadowAngularRadius(cam_r). */
void oracle_gas_lens_ray(float cam_r, float alpha_crit, int32_t n_rays, int32_t n_phis,
                         int32_t max_steps, float r_max, int32_t i, float* table);
void oracle_gas_lens_table(float cam_r, float alpha_crit, int32_t n_rays, int32_t n_phis,
                           int32_t max_steps, float r_max, float* table);

#ifdef __cplusplus
}
#endif
void oracle_camera_dirs(const float pos[3], const float B[9], float tan_half_fov, int32_t width,
                        int32_t height, const float jitter[2], int32_t jitter_hashed, uint64_t seed,
                        uint32_t camera_id, uint32_t sample, int64_t first, int64_t count,
                        int64_t stride, float* dir3_out);

/* Thin-disk shading (shade.inc). */
typedef struct oracle_shading {
    float density, opacity, peak_temp, brightness, noise_scale, noise_strength, warmth, doppler,
        time, inflow;
    float inner, outer;
} oracle_shading;
float oracle_exp2f(float x);
float oracle_log2f(float x);
float oracle_logf(float x);
float oracle_expf(float x);
float oracle_powf(float x, float y);
uint16_t oracle_f2h(float x);
float oracle_h2f(uint16_t h);
void oracle_grad3(int cx, int cy, int cz, uint32_t salt, float g[3]);
float oracle_pgnoise3(float px, float py, float pz, int period, uint32_t salt);
float oracle_noise_texel(int x, int y, int z, uint32_t seed);
void oracle_noise_bake(uint16_t* vol, uint32_t seed, int32_t threads);
float oracle_sample_noise(const uint16_t* vol, float s, float t, float r);
void oracle_blackbody_lut(float* lut);
void oracle_blackbody(const float* lut, float T, float o[4]);
float oracle_temp_profile(const oracle_shading* c, float r);
float oracle_radial_profile(const oracle_shading* c, float r);
float oracle_disk_noise(const oracle_shading* c, const uint16_t* vol, float r, float phi,
                        float y_over_h);
void oracle_thin_disk_sample(const oracle_shading* c, const uint16_t* vol, const float* lut,
                             float lapse_cam, float r, float phi, float gphys, float out[4]);
void oracle_shade_rays(const float pos[3], const float basis[9], float tan_half_fov, int32_t width,
                       int32_t height, const float jitter[2], int32_t jitter_hashed, uint64_t seed,
                       uint32_t camera_id, uint32_t sample, const oracle_lawS_params* p,
                       const oracle_shading* c, const uint16_t* vol, const float* lut,
                       int64_t first, int64_t count, int64_t stride, const uint8_t* flags,
                       const uint8_t* n_cross, const float* disk_r, const float* disk_phi,
                       const float* disk_g, float* rgb_out, float* sky_out);
void oracle_shade_march(const float pos[3], const float basis[9], float tan_half_fov, int32_t width,
                        int32_t height, const float jitter[2], int32_t jitter_hashed,
                        uint64_t seed, uint32_t camera_id, uint32_t sample,
                        const oracle_lawS_params* p, const oracle_shading* c,
                        const uint16_t* vol, const float* lut, int64_t first, int64_t count,
                        int64_t stride, int32_t mode, float* rgb_out, float* sky_out);
float oracle_accumulate_weight(int32_t sample_index, int32_t cap);
void oracle_accumulate(const float* cur, float* acc, int64_t n, int32_t sample_index, int32_t cap);
void oracle_halton_jitter(int32_t sample, float out[2]);

/* Sky of escaped rays (sky.inc: sky.glsl:57-119 with the backdrop cube map lookup). Cube maps:
 * float4 texels, levels x 6 faces (+X -X +Y -Y +Z -Z) x S_l x S_l, S_l = max(1, res >> l). */
typedef struct oracle_sky_params {
    float sky_brightness, star_density, star_brightness, doppler;
} oracle_sky_params;
int oracle_cube_face(const float d[3], float out_st[2]);
void oracle_cube_sample(const float* texels, int32_t res, int32_t levels, const float d[3],
                        const float dx[3], const float dy[3], float o[3], float* lod);
void oracle_star_layer(const float* lut, const float d[3], const oracle_footprint* fp, float N,
                       float prob, float flux, float sig_star, float omega0, uint32_t salt,
                       float g_shift, float o[3]);
void oracle_sample_sky(const oracle_sky_params* k, const float* texels, int32_t res, int32_t levels,
                       const float* lut, float rcam, const float d[3], const float dx[3],
                       const float dy[3], const oracle_footprint* fp, float omega0, float o[3]);
/* The sky pass over a frame: for every escaped pixel (flags bit 2) of the strided sample with
 * weight w != 0 (weight NULL: 1), rgb[p + k n] += w * sampleSky(...)[k]; the derivatives and the
 * footprint as oracle_sky_footprint (coarse as there). texels may be NULL. */
void oracle_sky_pass(const float pos[3], const float B[9], float tan_half_fov, int32_t width,
                     int32_t height, const float jitter[2], int32_t jitter_hashed, uint64_t seed,
                     uint32_t camera_id, uint32_t sample, int32_t coarse, const uint8_t* flags,
                     const float* esc, const oracle_sky_params* k, const float* texels,
                     int32_t res, int32_t levels, const float* lut, const float* weight,
                     float* rgb, int64_t first, int64_t stride, int32_t threads);

void oracle_volumetric(const float pos[3], const float basis[9], float tan_half_fov, int32_t width,
                       int32_t height, const float jitter[2], int32_t jitter_hashed, uint64_t seed,
                       uint32_t camera_id, uint32_t sample, const oracle_lawS_params* p,
                       const oracle_shading* c, const uint16_t* vol, const float* lut,
                       int64_t first, int64_t count, int64_t stride, int32_t mode, float* rgb_out,
                       float* sky_out, uint8_t* cls_out, uint8_t* flags_out, int32_t* steps_out,
                       float* esc_out, int32_t threads);
int32_t oracle_vol_ray(const float pos[3], const float dir[3], const oracle_lawS_params* p,
                       const oracle_shading* c, const uint16_t* vol, const float* lut, int32_t mode,
                       float out[8]);
void oracle_vol_sample(const oracle_shading* c, float thickness, float lapse, const uint16_t* vol,
                       const float* lut, float px, float py, float pz, float r, float ds,
                       float b_axis, const float rgb_t[4], float out[4]);
void oracle_hermite_f64(const double x0[3], const double m0[3], const double x1[3],
                        const double m1[3], double t, double out[3]);

#endif
