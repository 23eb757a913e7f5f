/* f4_multi.c -- F4 of one system modulo several primes, in lockstep, with
 * every round's linear algebra of all images in one batched launch.
 *
 * The multimodular path over the rationals needs the same system's images
 * modulo many primes.  Their F4 runs are independent and, for the lucky
 * primes, go through matrices of one shape round after round; each matrix
 * alone is too small to fill an MI355X.  Here every image runs the F4 host
 * (gbgpu_f4) in a host thread of its own; its linear-algebra callback hands
 * the round's matrix to a rendezvous and waits.  When every image still
 * running has handed in its matrix, the last one to arrive uploads them as
 * one batch (gbgpu_batch_upload) and reduces them with one graph launch
 * (gbgpu_batch_la, la_option 1 on every matrix, each modulo its own prime),
 * then wakes the others with their rows.  An image that finishes leaves
 * the rendezvous, so the remaining ones never wait for it.  The final
 * interreduction of a reduced basis goes to the device one image at a time.
 * All device work is serialised on the one context. */
#include <omp.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/gbgpu_f4.h"

typedef struct {
    pthread_mutex_t mu;
    pthread_cond_t cv;
    pthread_mutex_t gpu;      /* one caller of the context at a time */
    gbgpu_ctx *ctx;
    int32_t form;
    int32_t nimg, active, nsub;
    int64_t gen;              /* rounds handed out so far */
    const gbgpu_csr_t **mat;  /* nimg: the matrix an image handed in, or NULL */
    gbgpu_rows_t *res;        /* nimg: its result rows */
    int *rc;                  /* nimg: its result code */
    int64_t batches, batched; /* launches, matrices through them */
    double device_ms;
} rdv_t;

typedef struct {
    rdv_t *r;
    int32_t i;
    const gbgpu_f4_in_t *in;
    gbgpu_f4_out_t *out;
    int nthreads;
    int rc;
} img_t;

/* with r->mu held: every active image has handed in a matrix -- run them
 * as one batch and publish the results */
static void run_batch(rdv_t *r) {
    const gbgpu_csr_t **sub = (const gbgpu_csr_t **)malloc((size_t)r->nimg * sizeof(*sub));
    gbgpu_csr_t *mats = (gbgpu_csr_t *)malloc((size_t)r->nimg * sizeof(gbgpu_csr_t));
    int32_t *who = (int32_t *)malloc((size_t)r->nimg * sizeof(int32_t));
    int32_t n = 0;
    int rc = (sub && mats && who) ? GBGPU_OK : GBGPU_ENOMEM;
    for (int32_t i = 0; i < r->nimg && rc == GBGPU_OK; ++i)
        if (r->mat[i]) {
            mats[n] = *r->mat[i];
            who[n++] = i;
        }
    gbgpu_batch *b = NULL;
    pthread_mutex_lock(&r->gpu);
    if (rc == GBGPU_OK) rc = gbgpu_batch_upload(r->ctx, mats, n, &b);
    if (rc == GBGPU_OK) rc = gbgpu_batch_la(r->ctx, b, r->form, NULL);
    if (rc == GBGPU_OK) r->device_ms += gbgpu_phase_ms(r->ctx, 2);
    for (int32_t k = 0; k < n; ++k) {
        const int32_t i = who[k];
        memset(&r->res[i], 0, sizeof(gbgpu_rows_t));
        r->rc[i] = rc == GBGPU_OK ? gbgpu_batch_rows_download(r->ctx, b, k, &r->res[i]) : rc;
        r->mat[i] = NULL;
    }
    gbgpu_batch_free(b);
    pthread_mutex_unlock(&r->gpu);
    r->batches++;
    r->batched += n;
    r->nsub = 0;
    r->gen++;
    pthread_cond_broadcast(&r->cv);
    free(sub);
    free(mats);
    free(who);
}

static int la_multi(void *user, const gbgpu_csr_t *m, int interreduce, gbgpu_rows_t *out) {
    img_t *im = (img_t *)user;
    rdv_t *r = im->r;
    if (interreduce) {
        /* the reduced basis: one image at a time */
        gbgpu_f4_gpu_t g = {r->ctx, r->form, 0.0};
        pthread_mutex_lock(&r->gpu);
        const int rc = gbgpu_f4_gpu_la(&g, m, 1, out);
        pthread_mutex_unlock(&r->gpu);
        pthread_mutex_lock(&r->mu);
        r->device_ms += g.device_ms;
        pthread_mutex_unlock(&r->mu);
        return rc;
    }
    pthread_mutex_lock(&r->mu);
    const int64_t gen = r->gen;
    r->mat[im->i] = m;
    r->nsub++;
    if (r->nsub == r->active) run_batch(r);
    else
        while (r->gen == gen) pthread_cond_wait(&r->cv, &r->mu);
    *out = r->res[im->i];
    memset(&r->res[im->i], 0, sizeof(gbgpu_rows_t));
    const int rc = r->rc[im->i];
    pthread_mutex_unlock(&r->mu);
    return rc;
}

static void *img_main(void *arg) {
    img_t *im = (img_t *)arg;
    omp_set_num_threads(im->nthreads);
    im->rc = gbgpu_f4(im->in, la_multi, im, im->out);
    rdv_t *r = im->r;
    pthread_mutex_lock(&r->mu);
    r->active--;
    /* the others may all be waiting for this image to hand in a matrix */
    if (r->active > 0 && r->nsub == r->active) run_batch(r);
    pthread_mutex_unlock(&r->mu);
    return NULL;
}

int gbgpu_f4_multimod(gbgpu_ctx *ctx, const gbgpu_f4_in_t *ins, int32_t nimg, int32_t form,
                      gbgpu_f4_out_t *outs, gbgpu_f4_multi_stats_t *stats) {
    if (!ctx || !ins || !outs || nimg < 0) return GBGPU_EINVAL;
    if (form != GBGPU_FORM_PLAIN && form != GBGPU_FORM_MFMA) return GBGPU_EINVAL;
    for (int32_t i = 0; i < nimg; ++i) memset(&outs[i], 0, sizeof(gbgpu_f4_out_t));
    if (stats) memset(stats, 0, sizeof(*stats));
    if (nimg == 0) return GBGPU_OK;
    rdv_t r;
    memset(&r, 0, sizeof(r));
    pthread_mutex_init(&r.mu, NULL);
    pthread_cond_init(&r.cv, NULL);
    pthread_mutex_init(&r.gpu, NULL);
    r.ctx = ctx;
    r.form = form;
    r.nimg = r.active = nimg;
    r.mat = (const gbgpu_csr_t **)calloc((size_t)nimg, sizeof(*r.mat));
    r.res = (gbgpu_rows_t *)calloc((size_t)nimg, sizeof(gbgpu_rows_t));
    r.rc = (int *)calloc((size_t)nimg, sizeof(int));
    img_t *im = (img_t *)calloc((size_t)nimg, sizeof(img_t));
    pthread_t *th = (pthread_t *)calloc((size_t)nimg, sizeof(pthread_t));
    int rc = (r.mat && r.res && r.rc && im && th) ? GBGPU_OK : GBGPU_ENOMEM;
    /* the host's OpenMP threads shared out over the images */
    const int share = omp_get_max_threads() / nimg;
    int32_t started = 0;
    for (int32_t i = 0; i < nimg && rc == GBGPU_OK; ++i) {
        im[i].r = &r;
        im[i].i = i;
        im[i].in = &ins[i];
        im[i].out = &outs[i];
        im[i].nthreads = share > 1 ? share : 1;
        if (pthread_create(&th[i], NULL, img_main, &im[i]) != 0) {
            rc = GBGPU_ENOMEM;
            /* the images that did start must not wait for this one */
            pthread_mutex_lock(&r.mu);
            r.active -= nimg - i;
            if (r.active > 0 && r.nsub == r.active) run_batch(&r);
            pthread_mutex_unlock(&r.mu);
            break;
        }
        ++started;
    }
    for (int32_t i = 0; i < started; ++i) {
        pthread_join(th[i], NULL);
        if (rc == GBGPU_OK && im[i].rc != GBGPU_OK) rc = im[i].rc;
    }
    if (stats) {
        stats->batches = r.batches;
        stats->matrices = r.batched;
        stats->device_ms = r.device_ms;
    }
    free(r.mat);
    free(r.res);
    free(r.rc);
    free(im);
    free(th);
    pthread_mutex_destroy(&r.mu);
    pthread_cond_destroy(&r.cv);
    pthread_mutex_destroy(&r.gpu);
    return rc;
}
