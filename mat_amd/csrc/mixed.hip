// Mixed-precision least squares with iterative refinement (placeholder until
// the low-precision trailing update lands).
#include "mqr_internal.h"

namespace mqr {

int gels_mixed(mqr_handle_t hd, mqr_precision_t prec, int64_t m, int64_t n, const double *A,
               int64_t lda, const double *b, double *x, const mqr_ir_opts_t *opts,
               mqr_ir_info_t *info) {
    (void)hd; (void)prec; (void)m; (void)n; (void)A; (void)lda; (void)b; (void)x; (void)opts;
    (void)info;
    return MQR_ERR_UNSUPPORTED;
}

}  // namespace mqr
