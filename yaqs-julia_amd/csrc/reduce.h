// Wavefront reductions on DPP lane moves (no LDS round trips): pair, quad,
// half-row and row stages inside each 16-lane row, then the four row sums
// read from lanes 0, 16, 32, 48.  The result is uniform across the wave.
#pragma once

#include <hip/hip_runtime.h>

namespace yq {

template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double lane_read(double v, int lane) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                            __builtin_amdgcn_readlane(__double2loint(v), lane));
}

__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_move<0xb1>(v);   // quad_perm [1,0,3,2]: pairs
    v += dpp_move<0x4e>(v);   // quad_perm [2,3,0,1]: quads
    v += dpp_move<0x141>(v);  // row_half_mirror: 8 lanes
    v += dpp_move<0x140>(v);  // row_mirror: 16 lanes
    return (lane_read(v, 0) + lane_read(v, 16)) + (lane_read(v, 32) + lane_read(v, 48));
}

__device__ __forceinline__ double wave_max(double v) {
    v = fmax(v, dpp_move<0xb1>(v));
    v = fmax(v, dpp_move<0x4e>(v));
    v = fmax(v, dpp_move<0x141>(v));
    v = fmax(v, dpp_move<0x140>(v));
    return fmax(fmax(lane_read(v, 0), lane_read(v, 16)), fmax(lane_read(v, 32), lane_read(v, 48)));
}

__device__ __forceinline__ double wave_min(double v) {
    v = fmin(v, dpp_move<0xb1>(v));
    v = fmin(v, dpp_move<0x4e>(v));
    v = fmin(v, dpp_move<0x141>(v));
    v = fmin(v, dpp_move<0x140>(v));
    return fmin(fmin(lane_read(v, 0), lane_read(v, 16)), fmin(lane_read(v, 32), lane_read(v, 48)));
}

}  // namespace yq
