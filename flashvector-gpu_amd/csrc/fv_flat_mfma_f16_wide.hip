// Scan instantiations for f16 codes of 769..1536 dims (16 queries per wave).
#include "fv_flat_mfma_scan.hpp"

namespace fv {

hipError_t scan_f16_wide(const ScanLaunch& L) {
    switch (L.p->ks) {
        case 32: return launch_scan_metric<FV_DTYPE_F16, 32>(L);
        case 48: return launch_scan_metric<FV_DTYPE_F16, 48>(L);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace fv
