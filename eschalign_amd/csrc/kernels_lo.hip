// The local-mode instances of k_dp16 on 16-row stripes (every variant)
// and their launch.
#include "dp16.h"

namespace esch {

const void *dp16_fn_lo16(int variant) { return dp16_fn_impl<true, 16>(variant); }

hipError_t launch_dp16_lo16(const KArgs &a, const Launch &g, hipStream_t st) {
    return launch_dp16_impl<true, 16>(a, g, st);
}

}  // namespace esch
