// The semiglobal-mode instances of k_dp16 on 24-row stripes (the schedules 4-8)
// and their launch.
#include "dp16.h"

namespace esch {

const void *dp16_fn_sg24(int variant) { return dp16_fn_impl<false, 24>(variant); }

hipError_t launch_dp16_sg24(const KArgs &a, const Launch &g, hipStream_t st) {
    return launch_dp16_impl<false, 24>(a, g, st);
}

}  // namespace esch
