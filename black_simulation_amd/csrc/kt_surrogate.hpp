// kt_surrogate.hpp -- the learned surrogate (MLP 5-256-256-256-256-8, BF16 MFMA).
#pragma once

#include <hip/hip_runtime.h>

#include "kt_internal.hpp"

namespace kts {

struct Surrogate;

void destroy(Surrogate* s);

// Surrogate inference over every ray of the batch, gating, then the integrator
// over the rays the gate rejects (re-queued through L.ray_list).  ev[1], ev[2]
// bracket the integrator kernel.
hipError_t trace_with_surrogate(Surrogate* s, const kti::TraceLaunch& L, double gate_prob, double gate_ring,
                                hipStream_t stream, hipEvent_t* ev);

}  // namespace kts
