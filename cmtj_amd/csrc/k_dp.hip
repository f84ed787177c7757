// Dormand-Prince kernels: single-layer instances and instances of 2-4
// layers, every feature (runtime flags), FP64.
#include "kernels_dp.h"
#include "registry.h"
namespace cmtj {
static const KEntry kD[] = {
    {CMTJ_FP64, 1, 1, CMTJ_DORMAND_PRINCE, F_ALL, false, (const void*)&k_dp<F_ALL>, "k_dp<double,1,F_ALL>"},
    {CMTJ_FP64, 2, 1, CMTJ_DORMAND_PRINCE, F_ALL, false, (const void*)&k_dp_ml<2, F_ALL>, "k_dp_ml<double,2,F_ALL>"},
    {CMTJ_FP64, 4, 1, CMTJ_DORMAND_PRINCE, F_ALL, false, (const void*)&k_dp_ml<4, F_ALL>, "k_dp_ml<double,4,F_ALL>"},
};
KTable kernels_dp() { return {kD, sizeof(kD) / sizeof(kD[0])}; }
const void* dp_queue_kernel() { return (const void*)&k_dp_q<F_ALL>; }
const void* clear_u32_kernel() { return (const void*)&k_clear_u32; }
}  // namespace cmtj
