// One-component-per-lane kernels (anchored binary32 RK4): the exact feature
// sets of the PIMM sweep, the R(H) loop and the VSD sweep, and the generic
// lean set (anisotropy, IEC, field drives of any kind).
#include "registry.h"
namespace cmtj {
static const KEntry kC[] = {
    CMTJ_KC(2, FM_PIMM, 1),  // the PIMM sweep: one Oersted step record
    CMTJ_KCLOG(2, FM_PIMM, 1),  // the same with its log layout compiled in
    CMTJ_KC(2, FM_PIMM, -1), CMTJ_KC(4, FM_PIMM, -1), CMTJ_KC(1, FM_RH, 0), CMTJ_KC(1, FM_VSD, 1),
    CMTJ_KC(1, FM_LEAN, -1), CMTJ_KC(2, FM_LEAN, -1), CMTJ_KC(4, FM_LEAN, -1),
};
KTable kernels_cl() { return {kC, sizeof(kC) / sizeof(kC[0])}; }
}  // namespace cmtj
