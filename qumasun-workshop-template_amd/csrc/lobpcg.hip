// Block LOBPCG driver (placeholder until the solver lands).
#include "qm_internal.h"

extern "C" {

int qm_lobpcg_create(qm_lobpcg **out, qm_ctx *ctx, int nbands, int ncomp, int precision) {
    (void)ctx; (void)nbands; (void)ncomp; (void)precision;
    if (out) *out = nullptr;
    qm::set_error("qm_lobpcg_create: LOBPCG solver not available in this build");
    return QM_ERR_STATE;
}

void qm_lobpcg_destroy(qm_lobpcg *s) { (void)s; }

int qm_lobpcg_solve(qm_lobpcg *s, void *d_x, double *eigenvalues, double *residual_norms, int iter_max, int ik,
                    int first_time, int *iters_done, void *stream) {
    (void)s; (void)d_x; (void)eigenvalues; (void)residual_norms; (void)iter_max; (void)ik; (void)first_time;
    (void)iters_done; (void)stream;
    qm::set_error("qm_lobpcg_solve: LOBPCG solver not available in this build");
    return QM_ERR_STATE;
}

int qm_lobpcg_timings(qm_lobpcg *s, double *t6) {
    (void)s; (void)t6;
    return QM_ERR_STATE;
}

}  // extern "C"
