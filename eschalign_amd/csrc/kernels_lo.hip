// The local-mode instances of k_dp16 (every variant) and their launch.
#include "dp16.h"

namespace esch {

const void *dp16_fn_lo(int variant) { return dp16_fn_impl<true>(variant); }

hipError_t launch_dp16_lo(const KArgs &a, const Launch &g, hipStream_t st) {
    return launch_dp16_impl<true>(a, g, st);
}

}  // namespace esch
