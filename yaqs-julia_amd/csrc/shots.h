// Shot sampling of one MPS on the device (shots.hip).
#pragma once

#include <cstdint>
#include <vector>

#include "common.h"

namespace yq {

// single_shot_measure / measure_shots (src/MPS.jl:683-749) for `shots`
// shots at once.  A[i]: site i row-major [a][s][b] (dl[i] x 2 x dr[i],
// dl[0] = dr[L-1] = 1).  bits_out[t] (host) packs shot t's outcomes with
// site i in bit i (site 1 in the lowest bit, as the reference packs them);
// shot t draws the Philox uniforms (seed, t, i), i = 0..L-1.
void measure_shots(int L, const std::vector<std::vector<c64>> &A, const std::vector<int> &dl,
                   const std::vector<int> &dr, int64_t shots, uint64_t seed, int64_t *bits_out, hipStream_t s);

}  // namespace yq
