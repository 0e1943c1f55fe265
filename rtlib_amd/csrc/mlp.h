// Host launchers of the learned guide's kernels (mlp.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp_dev.h"

namespace rtpg {

constexpr uint32_t kMlpTrainBlock = 256;   // 4 waves: 64 samples per iteration
constexpr uint32_t kMlpTrainGrid = 512;    // persistent work-groups of a training step
constexpr uint32_t kMlpSortCells = 64;     // cells per axis of the training windows' spatial order

// One training sample (four float4): (x, roughness), (n, bsdf pdf at w),
// (wo, weight = f Li |cos| / p_sample averaged over RGB), (w, 0).
hipError_t launch_mlp_init(float* w_master, uint16_t* w_bf16, float* table, uint32_t* table_bf16, size_t table_floats,
                           uint64_t seed, hipStream_t s);
hipError_t launch_mlp_train(const MlpView& m, const float4* samples, uint32_t n, float* partial, float* loss_partial,
                            uint32_t grid, hipStream_t s);
// Copies n samples to out in spatial order (buckets of the Morton code on a
// kMlpSortCells^3 grid over the guide's box); hist and offs hold
// kMlpSortCells^3 words, total one.
hipError_t launch_mlp_spatial_order(const MlpView& m, const float4* samples, uint32_t n, uint32_t* hist,
                                    uint32_t* offs, uint32_t* total, float4* out, hipStream_t s);
hipError_t launch_mlp_reduce(const float* partial, const float* loss_partial, uint32_t grid, float* grad_w,
                             float* loss, hipStream_t s);
// Adam on n fp32 parameters with gradient g * gscale; writes the bf16 copy
// (w_bf16 for weights, or packed pairs into table_bf16 for the table) and
// zeroes the gradient.  For the table, entries from pk_first on also take the
// packed bf16 gradient gpk[entry - pk_first] (and zero it); gpk may be NULL.
hipError_t launch_mlp_adam(float* p, float* m1, float* m2, float* g, float gscale, size_t n, float lr_t, float b1,
                           float b2, float eps, uint16_t* w_bf16, uint32_t* table_bf16, uint32_t* gpk,
                           size_t pk_first, hipStream_t s);
// Queries: q = 3 float4 per sample ((x, roughness), (n, 0), (wo, 0)).
hipError_t launch_mlp_forward(const MlpView& m, const float4* q, uint32_t n, float* raw, float* mix, hipStream_t s);
// Draws one direction per sample from the mixture of launch_mlp_forward's
// mix output (41 floats per sample); out: (w xyz, mixture pdf) per sample.
hipError_t launch_mlp_sample(const float* mix, uint32_t n, uint64_t seed, float4* out, hipStream_t s);
// Fills the n = cache_res^3 * 6 entries of the spatial cache (mlp_cache_entry).
hipError_t launch_mlp_cache_fill(const MlpView& m, uint32_t n, float4* mix, float* c, hipStream_t s);

}  // namespace rtpg
