// Launch-variant tuner of the TJM engine: for each contraction shape the
// engine has several interchangeable kernels (the C tile shape of the
// matrix-core GEMM; the MPO step fused into the GEMM epilogue or run as its
// own kernel).  They give bit-identical results, so the choice only
// affects speed: the fastest is measured with HIP events on first use of a
// shape and cached for the process, per device (and in the file named by
// YAQS_TUNE_CACHE, when set, for later processes).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <functional>

namespace yq {

// Shape key: kind of site, element size, then up to eight shape integers
// (dimensions bucketed by the caller).
using TuneKey = std::array<int, 10>;

// Index of the fastest of n variants for key.  launch(v) enqueues variant v
// on stream s; it must compute a pure function of its inputs (each variant
// is run several times, the last run leaves the result).  Returns 0 without
// measuring when tuning is off (YAQS_TUNE=0 or tune_set_enabled(0)).
int tuned_variant(const TuneKey &key, int n, const std::function<void(int)> &launch, hipStream_t s);

// The cached choice for key, or 0 when none is cached or tuning is off;
// never measures (for launches queued ahead of a host decision).
int cached_variant(const TuneKey &key);

bool tune_enabled();
// While on, tuned_variant on the calling thread returns the cached choice
// (or 0) without measuring: no events are created, nothing extra launched.
void tune_freeze(bool on);
// True on the calling thread while a measurement runs (profilers skip it).
bool tune_measuring();

}  // namespace yq
