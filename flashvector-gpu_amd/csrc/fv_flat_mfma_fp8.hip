// Scan instantiations for FP8 e4m3 codes with per-row scales. The kernel is
// in fv_flat_mfma_scan.hpp.
#include "fv_flat_mfma_scan.hpp"

namespace fv {

hipError_t scan_fp8(const ScanLaunch& L) {
    switch (L.p->ks) {
        case 1: return launch_scan_metric<FV_DTYPE_FP8_E4M3, 1>(L);
        case 2: return launch_scan_metric<FV_DTYPE_FP8_E4M3, 2>(L);
        case 4: return launch_scan_metric<FV_DTYPE_FP8_E4M3, 4>(L);
        case 6: return launch_scan_metric<FV_DTYPE_FP8_E4M3, 6>(L);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace fv
