// Scan instantiations for bf16 codes of 384..768 dims (32 queries per wave).
#include "fv_flat_mfma_scan.hpp"

namespace fv {

hipError_t scan_bf16_large(const ScanLaunch& L) {
    switch (L.p->ks) {
        case 12: return launch_scan_metric<FV_DTYPE_BF16, 12>(L);
        case 16: return launch_scan_metric<FV_DTYPE_BF16, 16>(L);
        case 24: return launch_scan_metric<FV_DTYPE_BF16, 24>(L);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace fv
