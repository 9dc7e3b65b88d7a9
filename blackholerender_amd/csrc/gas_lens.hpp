// gas_lens.hpp — launcher of the gas lens table kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bhs {

// One table: n_rays fan rays from a static camera at (cam_r, 0, 0); the table holds n_phis rows of
// n_rays floats (row j: log r at swept angle j * 2 pi / (n_phis - 1)).
struct LensArgs {
    float cam_r, alpha_crit, r_max;
    int32_t n_rays, n_phis, max_steps;
};

hipError_t launch_gas_lens(const LensArgs& a, float* table, hipStream_t s);

// Tables of up to kLensBatchMax camera radii (same n_rays, n_phis), table k at
// tables + k * n_rays * n_phis, in one launch.
constexpr int kLensBatchMax = 64;
struct LensBatch {
    float cam_r[kLensBatchMax], alpha_crit[kLensBatchMax];
    float r_max;
    int32_t n_rays, n_phis, max_steps, n_tables;
};

hipError_t launch_gas_lens_batch(const LensBatch& b, float* tables, hipStream_t s);

}  // namespace bhs
