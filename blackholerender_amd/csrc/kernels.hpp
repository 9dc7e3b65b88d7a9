// kernels.hpp — kernel entry points of the tracer library and their launch helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lawS_device.hpp"

namespace bhs {

// Output planes (device pointers, any may be null) — bhs_ray_soa of include/bhs_trace.h.
struct SoA {
    uint8_t* cls;
    uint8_t* flags;
    uint8_t* n_cross;
    uint16_t* steps;
    float* esc;    // 3 planes of n
    float* disk;   // 6 planes of n: r0 r1 phi0 phi1 g0 g1
    float* b;
    int64_t n;     // plane stride
};

// Device counters of one launch (layout of bhs_stats).
struct Stats {
    unsigned long long rays, steps, crossings, escapes, band64, fallback, surrogate, pad1;
};

// Scratch used by one render launch.
struct Scratch {
    Stats* stats;
    uint32_t* work;              // work-queue counter of the persistent kernels
    uint32_t* band_work;         // work-queue counter of the FP64 band march
    uint32_t* band_count;        // number of rays in the FP64 band list
    uint32_t* band_ids;          // work items (camera tile order) of the FP64 band rays
    uint32_t band_cap;           // capacity of band_ids (>= the launch's work items)
    RootState<float>* roots;     // 2 per ray (global ray index) when root refinement is on
    uint32_t* seg;               // segmented work queue: kSegs counters, kSegStride words apart
    unsigned long long* seg_done;   // its used-up segments (bit mask)
};

// Segmented work queue of the persistent march (claim_seg): counters per contiguous range of
// tiles, each on its own 128-byte line; cleared with the launch counters.
constexpr uint32_t kSegs = 64;
constexpr uint32_t kSegStride = 32;

// Cameras of one batched launch: the camera table and item prefix sums are staged in LDS.
constexpr uint32_t kMaxBatchCams = 64;

// Work-item order of the integrator's kernels: a camera's items run over kItemTileW x kItemTileH
// pixel tiles (row-major over the tile grid), 64 items per tile (row-major inside it), so the 64
// consecutive items a wave claims are one tile of neighbouring rays.
#ifndef BHS_ITEM_TILE_W
#define BHS_ITEM_TILE_W 16
#endif
constexpr uint32_t kItemTileW = BHS_ITEM_TILE_W;
constexpr uint32_t kItemTileH = 64u / kItemTileW;
static_assert(kItemTileW * kItemTileH == 64u && (kItemTileW & (kItemTileW - 1u)) == 0u,
              "a work-item tile is 64 pixels, power-of-two wide");

__host__ __device__ __forceinline__ uint32_t item_tiles_x(uint32_t width) {
    return (width + kItemTileW - 1u) / kItemTileW;
}

// Pixel (col, row) of lane l of tile t of a camera with tiles_x tiles per tile row (col may lie
// beyond the width and row beyond the height in the padding of the last tile column / row).
__device__ __forceinline__ void tile_col_row(uint32_t tiles_x, uint32_t t, uint32_t l,
                                             uint32_t& col, uint32_t& row) {
    col = (t % tiles_x) * kItemTileW + (l % kItemTileW);
    row = (t / tiles_x) * kItemTileH + (l / kItemTileW);
}
// Pixel of work item q (tile q / 64, lane q % 64).
__device__ __forceinline__ void item_col_row(uint32_t tiles_x, uint32_t q, uint32_t& col,
                                             uint32_t& row) {
    tile_col_row(tiles_x, q >> 6, q & 63u, col, row);
}

// Work item of pixel (col, row): the inverse of item_col_row.
__device__ __forceinline__ uint32_t pixel_item(uint32_t tiles_x, uint32_t col, uint32_t row) {
    return ((row / kItemTileH) * tiles_x + col / kItemTileW) * 64u +
           (row % kItemTileH) * kItemTileW + col % kItemTileW;
}

// Work items of one camera: its image padded to whole work-item tiles.
uint32_t camera_items(int32_t width, int32_t height);

// Launch-shape choices of the FP32 march (see bhs_tuning in include/bhs_trace.h).
struct Tuning {
    int32_t mode;            // 0 static 8x8-tile grid, 1 persistent work queue with lane refill
    int32_t block;           // threads per block: 64 or 256
    int32_t waves_per_simd;  // persistent: resident waves per SIMD (1..8)
    int32_t refill_min;      // persistent: refill when at least this many lanes are idle (1..64)
};

hipError_t launch_zero(void* p, size_t bytes, hipStream_t s);
// ev_begin/ev_end (may be null) are recorded on `s` around the main march kernel. In BAND64
// precision the FP64 band kernel runs on `aux` between ev_fork (recorded on s) and ev_join (waited
// for on s).
hipError_t launch_render(const MarchConst& c, const SoA& out, const Scratch& sc, const Tuning& t,
                         int num_cus, hipStream_t s, hipStream_t aux, hipEvent_t ev_fork,
                         hipEvent_t ev_join, hipEvent_t ev_begin, hipEvent_t ev_end);
// Batched launch over a device camera table (item_base: ncam + 1 prefix sums of camera_items).
hipError_t launch_batch(const MarchConst* table, const uint32_t* item_base, uint32_t ncam,
                        uint32_t n_items, bool any_band, bool any_root, const SoA& out,
                        const Scratch& sc, const Tuning& t, int num_cus, hipStream_t s,
                        hipStream_t aux, hipEvent_t ev_fork, hipEvent_t ev_join,
                        hipEvent_t ev_begin, hipEvent_t ev_end);
// FP32 persistent march of a listed subset of one camera's work items (ids[0 .. *count), count
// read on the device, at most cap).
hipError_t launch_list(const MarchConst& c, const uint32_t* ids, const uint32_t* count, uint32_t cap,
                       const SoA& out, const Scratch& sc, const Tuning& t, int num_cus,
                       hipStream_t s);
// FP64 march of the band list sc.band_ids[0 .. *sc.band_count) on aux, forked from s at ev_fork
// and joined by ev_join (which the caller waits for).
hipError_t launch_band_list(const MarchConst& c, const SoA& out, const Scratch& sc, hipStream_t s,
                            hipStream_t aux, hipEvent_t ev_fork, hipEvent_t ev_join);
// The integrator's band scan of one camera on s, then the FP64 march of the listed rays on aux
// between ev_fork and ev_join (as in launch_render).
hipError_t launch_band_fork(const MarchConst& c, const SoA& out, const Scratch& sc, hipStream_t s,
                            hipStream_t aux, hipEvent_t ev_fork, hipEvent_t ev_join);
// FP64 root refinement of the crossings of integrator-marched FP32 rays (root64 cameras).
hipError_t launch_root_refine(const MarchConst& c, const SoA& out, const Scratch& sc, hipStream_t s);
// Batch versions (device camera table and item prefix sums; items are batch-global).
hipError_t launch_list_batch(const MarchConst* table, const uint32_t* item_base, uint32_t ncam,
                             const uint32_t* ids, const uint32_t* count, uint32_t cap,
                             const SoA& out, const Scratch& sc, const Tuning& t, int num_cus,
                             hipStream_t s);
hipError_t launch_band_list_batch(const MarchConst* table, const uint32_t* item_base,
                                  uint32_t ncam, uint32_t n_items, const SoA& out,
                                  const Scratch& sc, hipStream_t s, hipStream_t aux,
                                  hipEvent_t ev_fork, hipEvent_t ev_join);
hipError_t launch_root_refine_batch(const MarchConst* table, const uint32_t* item_base,
                                    uint32_t ncam, uint32_t n_items, const SoA& out,
                                    const Scratch& sc, hipStream_t s);
hipError_t launch_march_rays(const float* pos3, const float* dir3, int64_t n, const MarchConst& c,
                             const SoA& out, const Scratch& sc, hipStream_t s, hipEvent_t ev_begin,
                             hipEvent_t ev_end);
// y[i] = rsqrt_rn(x[i]) (the march's correctly rounded reciprocal square root), n >= 1.
hipError_t launch_rsqrt_rn(const float* x, float* y, int64_t n, hipStream_t s);
// VALU issue-rate probe (k_valu_probe): blocks x 256 lanes, 16 chains each, `iters` rounds.
hipError_t launch_valu_probe(int kind, uint32_t blocks, int32_t iters, float* out, hipStream_t s);
hipError_t launch_trace_T(const double* pos3, const double* dir3, int64_t n, int32_t max_steps,
                          double step_scale, double esc_tol, double r_esc_min, void* out,
                          hipStream_t s);

}  // namespace bhs
