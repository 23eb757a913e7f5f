/*
 * gb_adapter.c -- gb's linear_algebra / interreduce_matrix_rows entry
 * points (src/data.h:269-279) implemented on the flat C-ABI of libgbgpu.
 *
 * Marshalling: the rows of mat->r ([cf id, len % 4, len, columns]) become
 * one CSR; each coefficient array bs->cf_*[id] that some row uses is copied
 * once into a pool (rows built from one basis element keep sharing it).
 * The result rows are rebuilt as gb expects them (see include/gbgpu_gb.h).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <time.h>

#include "gb_types.h"
#include "../../include/gbgpu.h"
#include "../../include/gbgpu_gb.h"

enum { W8 = 8, W16 = 16, W32 = 32 };

static gbgpu_ctx *g_ctx = NULL;

static void die(const char *what, int rc)
{
    fprintf(stderr, "gbgpu: %s failed (%d): %s\n", what, rc,
            g_ctx ? gbgpu_last_error(g_ctx) : "no device context");
    abort();
}

static gbgpu_ctx *ctx_get(void)
{
    if (!g_ctx) {
        const char *dev = getenv("GBGPU_DEVICE");
        const int rc = gbgpu_init(dev ? atoi(dev) : 0, &g_ctx);
        if (rc) die("gbgpu_init", rc);
    }
    return g_ctx;
}

/* time.c:26-43 */
static double cpu_seconds(void)
{
    return ((double)CLOCKS_PER_SEC / 100000. + (double)clock()) / CLOCKS_PER_SEC;
}

static double wall_seconds(void)
{
    struct timeval t;
    gettimeofday(&t, NULL);
    return (double)t.tv_sec + (double)t.tv_usec / 1e6;
}

static uint32_t coef(const void *arr, int width, int64_t k)
{
    switch (width) {
        case W8:  return (uint32_t)(uint8_t)((const cf8_t *)arr)[k];
        case W16: return (uint32_t)(uint16_t)((const cf16_t *)arr)[k];
        default:  return (uint32_t)((const cf32_t *)arr)[k];
    }
}

static const void *basis_array(const struct bs_t *bs, int width, int32_t id)
{
    switch (width) {
        case W8:  return bs->cf_8[id];
        case W16: return bs->cf_16[id];
        default:  return bs->cf_32[id];
    }
}

typedef struct {
    gbgpu_csr_t csr;
    int64_t *rp, *cp;
    int32_t *ci, *rcf;
    uint32_t *cf;
} marshal_t;

/* rows of mat + their basis coefficient arrays -> flat CSR */
static void marshal(const struct mat_t *mat, const struct bs_t *bs, int width, uint32_t p,
                    int32_t ncl, marshal_t *m)
{
    const int32_t nr = mat->nr;
    int32_t maxid = -1;
    int64_t nnz = 0;
    for (int32_t i = 0; i < nr; ++i) {
        nnz += mat->r[i][2];
        if (mat->r[i][0] > maxid) maxid = mat->r[i][0];
    }
    int32_t *map = (int32_t *)malloc(((size_t)maxid + 2) * sizeof(int32_t));
    int32_t *alen = (int32_t *)calloc((size_t)nr + 1, sizeof(int32_t));
    int32_t *aid = (int32_t *)malloc(((size_t)nr + 1) * sizeof(int32_t));
    for (int32_t i = 0; i <= maxid; ++i) map[i] = -1;
    m->rp = (int64_t *)malloc(((size_t)nr + 1) * sizeof(int64_t));
    m->ci = (int32_t *)malloc(((size_t)nnz + 1) * sizeof(int32_t));
    m->rcf = (int32_t *)malloc(((size_t)nr + 1) * sizeof(int32_t));
    int32_t ncf = 0;
    int64_t w = 0;
    m->rp[0] = 0;
    for (int32_t i = 0; i < nr; ++i) {
        const hm_t *row = mat->r[i];
        const int32_t id = row[0], len = row[2];
        if (map[id] < 0) {
            map[id] = ncf;
            aid[ncf] = id;
            ++ncf;
        }
        if (len > alen[map[id]]) alen[map[id]] = len;
        m->rcf[i] = map[id];
        memcpy(m->ci + w, row + 3, (size_t)len * sizeof(int32_t));
        w += len;
        m->rp[i + 1] = w;
    }
    m->cp = (int64_t *)malloc(((size_t)ncf + 1) * sizeof(int64_t));
    m->cp[0] = 0;
    for (int32_t k = 0; k < ncf; ++k) m->cp[k + 1] = m->cp[k] + alen[k];
    m->cf = (uint32_t *)malloc(((size_t)m->cp[ncf] + 1) * sizeof(uint32_t));
    for (int32_t k = 0; k < ncf; ++k) {
        const void *src = basis_array(bs, width, aid[k]);
        for (int64_t j = 0; j < alen[k]; ++j) m->cf[m->cp[k] + j] = coef(src, width, j);
    }
    free(map);
    free(alen);
    free(aid);
    m->csr.nr = nr;
    m->csr.nc = mat->nc;
    m->csr.ncl = ncl;
    m->csr.ncf = ncf;
    m->csr.p = p;
    m->csr.rp = m->rp;
    m->csr.ci = m->ci;
    m->csr.rcf = m->rcf;
    m->csr.cp = m->cp;
    m->csr.cf = m->cf;
}

static void marshal_free(marshal_t *m)
{
    free(m->rp); free(m->cp); free(m->ci); free(m->rcf); free(m->cf);
}

/* result rows -> mat->r / mat->cf_* in gb's format (la_ff_32.c:1666-1736) */
static void adopt_rows(struct mat_t *mat, const gbgpu_rows_t *out, int width)
{
    const int32_t np = out->np;
    for (int32_t i = 0; i < mat->nr; ++i) free(mat->r[i]);  /* LA owns input rows */
    if (np == 0) {
        free(mat->r);
        mat->r = NULL;
    } else {
        mat->r = (hm_t **)realloc(mat->r, (size_t)np * sizeof(hm_t *));
    }
    void **cfa;
    switch (width) {
        case W8:  mat->cf_8 = (cf8_t **)realloc(mat->cf_8, ((size_t)np + 1) * sizeof(cf8_t *)); cfa = (void **)mat->cf_8; break;
        case W16: mat->cf_16 = (cf16_t **)realloc(mat->cf_16, ((size_t)np + 1) * sizeof(cf16_t *)); cfa = (void **)mat->cf_16; break;
        default:  mat->cf_32 = (cf32_t **)realloc(mat->cf_32, ((size_t)np + 1) * sizeof(cf32_t *)); cfa = (void **)mat->cf_32; break;
    }
    for (int32_t k = 0; k < np; ++k) {
        const int64_t b = out->rp[k];
        const int32_t len = (int32_t)(out->rp[k + 1] - b);
        hm_t *row = (hm_t *)malloc(((size_t)len + 3) * sizeof(hm_t));
        row[0] = k;
        row[1] = len % 4;
        row[2] = len;
        memcpy(row + 3, out->ci + b, (size_t)len * sizeof(int32_t));
        mat->r[k] = row;
        switch (width) {
            case W8: {
                cf8_t *c = (cf8_t *)malloc((size_t)len * sizeof(cf8_t));
                for (int32_t j = 0; j < len; ++j) c[j] = (cf8_t)out->cf[b + j];
                cfa[k] = c;
                break;
            }
            case W16: {
                cf16_t *c = (cf16_t *)malloc((size_t)len * sizeof(cf16_t));
                for (int32_t j = 0; j < len; ++j) c[j] = (cf16_t)out->cf[b + j];
                cfa[k] = c;
                break;
            }
            default: {
                cf32_t *c = (cf32_t *)malloc((size_t)len * sizeof(cf32_t));
                for (int32_t j = 0; j < len; ++j) c[j] = (cf32_t)out->cf[b + j];
                cfa[k] = c;
                break;
            }
        }
    }
    mat->np = mat->nr = mat->sz = np;
}

static int form_of(int32_t laopt) { return laopt == 101 ? GBGPU_FORM_PLAIN : GBGPU_FORM_MFMA; }

static void linear_algebra_common(struct mat_t *mat, const struct bs_t *bs, struct stat_t *st, int width)
{
    const double ct0 = cpu_seconds(), rt0 = wall_seconds();
    gbgpu_ctx *ctx = ctx_get();
    marshal_t m;
    marshal(mat, bs, width, (uint32_t)st->fc, mat->ncl, &m);
    gbgpu_rows_t out;
    memset(&out, 0, sizeof(out));
    int rc = 0;
    if (mat->nr > 0) {
        rc = gbgpu_la_ff(ctx, &m.csr, 1, form_of(st->laopt), &out);
        if (rc) die("linear algebra", rc);
    }
    marshal_free(&m);
    adopt_rows(mat, &out, width);
    gbgpu_rows_free(&out);
    st->la_ctime += cpu_seconds() - ct0;
    st->la_rtime += wall_seconds() - rt0;
    st->num_zerored += (mat->nrl - mat->np);
    if (st->info_level > 1) {
        printf("%7d new %7d zero", mat->np, mat->nrl - mat->np);
        fflush(stdout);
    }
}

void gbgpu_linear_algebra_ff_32(struct mat_t *mat, const struct bs_t *const bs, struct stat_t *st)
{
    linear_algebra_common(mat, bs, st, W32);
}

void gbgpu_linear_algebra_ff_16(struct mat_t *mat, const struct bs_t *const bs, struct stat_t *st)
{
    linear_algebra_common(mat, bs, st, W16);
}

void gbgpu_linear_algebra_ff_8(struct mat_t *mat, const struct bs_t *const bs, struct stat_t *st)
{
    linear_algebra_common(mat, bs, st, W8);
}

/* basis.c:25-60 (free_basis_elements) for the finite-field widths */
static void free_basis_elements(struct bs_t *bs, int width)
{
    for (int32_t i = 0; i < bs->ld; ++i) {
        switch (width) {
            case W8:  free(bs->cf_8[i]); break;
            case W16: free(bs->cf_16[i]); break;
            default:  free(bs->cf_32[i]); break;
        }
        free(bs->hm[i]);
    }
    bs->ld = 0;
}

static void interreduce_common(struct mat_t *mat, struct bs_t *bs, const struct stat_t *st, int width)
{
    if (st->info_level > 1) printf("                        ");
    gbgpu_ctx *ctx = ctx_get();
    marshal_t m;
    marshal(mat, bs, width, (uint32_t)st->fc, 0, &m);
    free_basis_elements(bs, width);  /* la_ff_32.c:1980 */
    gbgpu_rows_t out;
    memset(&out, 0, sizeof(out));
    if (mat->nr > 0) {
        gbgpu_mat *dm = NULL;
        int32_t np = 0;
        int rc = gbgpu_mat_upload(ctx, &m.csr, &dm);
        if (!rc) rc = gbgpu_interreduce(ctx, dm, form_of(st->laopt), &np);
        if (!rc) rc = gbgpu_rows_download(ctx, dm, &out);
        gbgpu_mat_free(dm);
        if (rc) die("interreduction", rc);
    }
    marshal_free(&m);
    adopt_rows(mat, &out, width);
    gbgpu_rows_free(&out);
}

void gbgpu_interreduce_matrix_rows_ff_32(struct mat_t *mat, struct bs_t *bs, const struct stat_t *st)
{
    interreduce_common(mat, bs, st, W32);
}

void gbgpu_interreduce_matrix_rows_ff_16(struct mat_t *mat, struct bs_t *bs, const struct stat_t *st)
{
    interreduce_common(mat, bs, st, W16);
}

void gbgpu_interreduce_matrix_rows_ff_8(struct mat_t *mat, struct bs_t *bs, const struct stat_t *st)
{
    interreduce_common(mat, bs, st, W8);
}
