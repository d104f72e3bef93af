// kt_surrogate.hip -- surrogate inference/training kernels (placeholder until the MLP lands).
#include "kt_surrogate.hpp"

namespace kts {

struct Surrogate {
    int dummy;
};

void destroy(Surrogate* s) { delete s; }

hipError_t trace_with_surrogate(Surrogate*, const kti::TraceLaunch&, double, double, hipStream_t, hipEvent_t*) {
    return hipErrorNotSupported;
}

}  // namespace kts
