// Dormand-Prince kernel: single-layer instances, every feature (runtime
// flags), FP64.
#include "kernels_dp.h"
#include "registry.h"
namespace cmtj {
static const KEntry kD[] = {
    {CMTJ_FP64, 1, 1, CMTJ_DORMAND_PRINCE, F_ALL, false, (const void*)&k_dp<F_ALL>, "k_dp<double,1,F_ALL>"},
};
KTable kernels_dp() { return {kD, sizeof(kD) / sizeof(kD[0])}; }
}  // namespace cmtj
