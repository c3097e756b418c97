// Scan instantiations for FP8 e4m3 codes of 769..3072 dims (32 queries per
// wave up to 1536 dims, 16 beyond).
#include "fv_flat_mfma_scan.hpp"

namespace fv {

hipError_t scan_fp8_wide(const ScanLaunch& L) {
    switch (L.p->ks) {
        case 8: return launch_scan_metric<FV_DTYPE_FP8_E4M3, 8>(L);
        case 12: return launch_scan_metric<FV_DTYPE_FP8_E4M3, 12>(L);
        case 16: return launch_scan_metric<FV_DTYPE_FP8_E4M3, 16>(L);
        case 24: return launch_scan_metric<FV_DTYPE_FP8_E4M3, 24>(L);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace fv
