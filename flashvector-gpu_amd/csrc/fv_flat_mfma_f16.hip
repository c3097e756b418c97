// Scan instantiations for f16 codes. The kernel is in fv_flat_mfma_scan.hpp.
#include "fv_flat_mfma_scan.hpp"

namespace fv {

hipError_t scan_f16_small(const ScanLaunch& L) {
    switch (L.p->ks) {
        case 1: return launch_scan_metric<FV_DTYPE_F16, 1>(L);
        case 2: return launch_scan_metric<FV_DTYPE_F16, 2>(L);
        case 3: return launch_scan_metric<FV_DTYPE_F16, 3>(L);
        case 4: return launch_scan_metric<FV_DTYPE_F16, 4>(L);
        case 6: return launch_scan_metric<FV_DTYPE_F16, 6>(L);
        case 8: return launch_scan_metric<FV_DTYPE_F16, 8>(L);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace fv
