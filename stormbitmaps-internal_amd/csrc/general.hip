// General (mixed-representation) corpora: arena upload, device selection and
// the per-cell dispatch. (Filled in below the dense path.)
#include "device_common.h"
#include "internal.h"

namespace stormhip {

void default_device_model(DeviceCostModel& m) {
    // default_model (kernels/storm_cost.cpp:482-496) and the default knobs
    // STORM_PROBE_SCALE=0.05, STORM_BR_ORIENT=16, STORM_RR_SCALE=1.
    const double r[10] = {0.116, 0.30, 0.60, 0.55, 0.22, 0.35, 0.40, 0.30, 0.45, 0.50};
    for (int i = 0; i < 10; ++i) m.ns_per_unit[i] = r[i];
    m.ns_fixed = 2.0; m.ns_per_occ_word = 0.30; m.ns_probe_hot = 0; m.ns_probe_cold = 0;
    m.probe_scale = 0.05; m.br_margin = 16.0; m.rr_scale = 1.0;
}

int general_allpairs(storm_hip_ctx*, storm_hip_corpus*, const storm_hip_options*,
                     storm_hip_stats*, unsigned long long*) {
    set_error("general corpora not built yet");
    return STORM_HIP_ERR_UNSUPPORTED;
}
void general_free(storm_hip_corpus*) {}


}  // namespace stormhip
