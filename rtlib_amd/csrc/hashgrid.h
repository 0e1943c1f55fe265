// Hash grid + Morton quadtree launchers (hashgrid.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "views.h"

namespace rtpg {

hipError_t launch_grid_find(HashGridView h, const float* pos, uint32_t n, bool insert, uint32_t* out,
                            hipStream_t s);
hipError_t launch_morton_build(float* w, uint32_t level, uint32_t per_cell, uint32_t cells, hipStream_t s);
hipError_t launch_morton_record(MortonView m, const uint32_t* cells, const float* canon, const float* val,
                                uint32_t n, hipStream_t s);
hipError_t launch_morton_sample(MortonView m, const uint32_t* cells, uint32_t n, uint64_t seed, float* out,
                                hipStream_t s);
hipError_t launch_morton_pdf(MortonView m, const uint32_t* cells, const float* canon, uint32_t n, float* out,
                             hipStream_t s);
hipError_t launch_fill_zero(float* p, size_t n, hipStream_t s);

}  // namespace rtpg
