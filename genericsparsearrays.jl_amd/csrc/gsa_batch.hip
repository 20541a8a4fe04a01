// gsa_batch.hip — batched uneven CSR SpMM (placeholder until the grouped kernel lands).
#include "gsa_common.h"

namespace gsa {
void batch_plan_release(gsa_plan_s* plan) { (void)plan; }
}  // namespace gsa

using namespace gsa;

extern "C" {
int gsa_batch_plan_create(void* stream, int vtype, int itype, int64_t count,
                          const gsa_batch_desc* host_desc, int rows_per_tile, gsa_plan* plan) {
    (void)stream; (void)vtype; (void)itype; (void)count; (void)host_desc; (void)rows_per_tile;
    (void)plan;
    set_error(GSA_ERR_UNSUPPORTED, "gsa_batch_plan_create: not built yet");
    return GSA_ERR_UNSUPPORTED;
}
int gsa_spmm_csr_bf16(void* stream, int itype, int64_t m, int64_t n, int64_t k,
                      const void* rowptr, const void* colval, const uint16_t* nzval_bf16,
                      const float* row_scale, int64_t nnz, const float* B, int64_t ldb, float* C,
                      int64_t ldc, float alpha, float beta, gsa_plan plan) {
    (void)stream; (void)itype; (void)m; (void)n; (void)k; (void)rowptr; (void)colval;
    (void)nzval_bf16; (void)row_scale; (void)nnz; (void)B; (void)ldb; (void)C; (void)ldc;
    (void)alpha; (void)beta; (void)plan;
    set_error(GSA_ERR_UNSUPPORTED, "gsa_spmm_csr_bf16: not built yet");
    return GSA_ERR_UNSUPPORTED;
}
int gsa_batch_spmm(void* stream, gsa_plan plan, float alpha, float beta) {
    (void)stream; (void)plan; (void)alpha; (void)beta;
    set_error(GSA_ERR_UNSUPPORTED, "gsa_batch_spmm: not built yet");
    return GSA_ERR_UNSUPPORTED;
}
}
