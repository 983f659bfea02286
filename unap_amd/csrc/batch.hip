// batch.hip -- many independent LDU systems in one persistent launch.
#include "unap_hip.h"

extern "C" {
int unap_hip_batch_create(int64_t, const int64_t*, const int64_t*, const int64_t*, const int64_t*,
                          const double*, const double*, const double*, const int32_t*, unap_hip_batch*) {
    return UNAP_HIP_ERR_STATE;
}
int unap_hip_batch_destroy(unap_hip_batch) { return UNAP_HIP_OK; }
int unap_hip_batch_solve(unap_hip_batch, double*, const double*, int, int64_t, double, double, unap_hip_perf*) {
    return UNAP_HIP_ERR_STATE;
}
int unap_hip_batch_last_ms(unap_hip_batch, double*) { return UNAP_HIP_ERR_STATE; }
}
