// Split-FP16 path (placeholder until the limb kernels land).
#include "bdc_internal.hpp"
#include "kernels.hpp"

namespace bdc {
int split_prepare(bdc_ctx *c) { return fail(BDC_ERR_ARG, "split-FP16 path not built yet"); }
int split_output_law_g1(bdc_ctx *c, const double *X) { return fail(BDC_ERR_ARG, "split-FP16 path not built yet"); }
int split_cross_operands(bdc_ctx *c, const double *L) { return fail(BDC_ERR_ARG, "split-FP16 path not built yet"); }
int split_cross_term(bdc_ctx *c, int epi, const EpiArgs &ea, int *nb) { return fail(BDC_ERR_ARG, "split-FP16 path not built yet"); }
}  // namespace bdc
