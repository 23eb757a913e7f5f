/* f4_la.c -- test infrastructure: the oracle's linear algebra behind the
 * F4 host's callback contract (gbgpu_f4_la_t in include/gbgpu_f4.h), so
 * that the F4 host can run end to end on the CPU: la_option 1
 * (la_ff_32.c:1804-1851) for a round, interreduce_matrix_rows
 * (la_ff_32.c:1949-2017) for the final reduction.  user points to the
 * OpenMP thread count (or is NULL for one thread). */
#include <stdlib.h>

#include "../include/gbgpu.h"
#include "la_ff_oracle.h"

int orc_f4_la(void *user, const gbgpu_csr_t *m, int interreduce, gbgpu_rows_t *out) {
    const int nth = user ? *(const int *)user : 1;
    orc_mat_t om = {m->nr, m->nc, m->ncl, m->rp, m->ci, m->rcf, m->cp, m->cf};
    orc_rows_t r = {0, NULL, NULL, NULL};
    const int32_t np = interreduce ? orc_interreduce(&om, (int32_t)m->p, &r)
                                   : orc_la_sparse_dense(&om, (int32_t)m->p, nth, &r);
    if (np < 0) return GBGPU_ECONTRACT;
    out->np = r.np;
    out->rp = r.rp;
    out->ci = r.ci;
    out->cf = r.cf;
    return GBGPU_OK;
}
