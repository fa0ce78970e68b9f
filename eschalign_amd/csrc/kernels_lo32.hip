// The local-mode instances of k_dp16 on 32-row stripes (the shifted schedules 6, 7 and 8)
// and their launch.
#include "dp16.h"

namespace esch {

const void *dp16_fn_lo32(int variant) { return dp16_fn_impl<true, 32>(variant); }

hipError_t launch_dp16_lo32(const KArgs &a, const Launch &g, hipStream_t st) {
    return launch_dp16_impl<true, 32>(a, g, st);
}

}  // namespace esch
