// irc_screen.hip -- MFMA distance screen for the kNN pass (gfx950).
#include "irc_screen.h"
#include "irc_kernels.h"

namespace irc {

bool screen_supported(const DevPlan&) { return false; }
size_t screen_pack_bytes(const DevPlan&, int64_t, int) { return 0; }
size_t screen_aux_bytes(const DevPlan&) { return 0; }
hipError_t launch_screen_knn(hipStream_t, const DevPlan&, int, int, int, const ScreenBufs&, int, int*, double*,
                             double*) {
    return hipErrorNotSupported;
}

}  // namespace irc
