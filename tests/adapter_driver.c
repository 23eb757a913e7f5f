/*
 * adapter_driver.c -- calls libgbgpu's gb adapters the way gb's F4 loop
 * does (through function pointers with gb's signatures) on a matrix read
 * from a file, checks the ownership/stat contract and writes the rows.
 *
 * usage: adapter_driver <in.bin> <out.bin>
 * in:  int32 mode (0 = linear_algebra, 1 = interreduce_matrix_rows), width
 *      (8/16/32), nr, nc, ncl, ncf, p, laopt, nrl; int64 rp[nr+1];
 *      int32 ci[nnz]; int32 rcf[nr]; int64 cp[ncf+1]; uint32 cf[cp[ncf]]
 * out: int32 np; int64 rp[np+1]; int32 ci[nnz]; uint32 cf[nnz]
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../gb_amd/csrc/gb_types.h"
#include "../include/gbgpu_gb.h"

#define CHECK(c, msg) do { if (!(c)) { fprintf(stderr, "contract: %s\n", msg); return 3; } } while (0)

static void *rd(FILE *f, size_t n) {
    void *p = malloc(n ? n : 1);
    if (n && fread(p, 1, n, f) != n) { fprintf(stderr, "short read\n"); exit(2); }
    return p;
}

int main(int argc, char **argv) {
    if (argc != 3) return 1;
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 1;
    int32_t h[9];
    if (fread(h, 4, 9, f) != 9) return 2;
    const int32_t mode = h[0], width = h[1], nr = h[2], nc = h[3], ncl = h[4], ncf = h[5];
    const int32_t p = h[6], laopt = h[7], nrl = h[8];
    int64_t *rp = rd(f, ((size_t)nr + 1) * 8);
    int32_t *ci = rd(f, (size_t)rp[nr] * 4);
    int32_t *rcf = rd(f, (size_t)nr * 4);
    int64_t *cp = rd(f, ((size_t)ncf + 1) * 8);
    uint32_t *cf = rd(f, (size_t)cp[ncf] * 4);
    fclose(f);

    struct bs_t bs;
    memset(&bs, 0, sizeof(bs));
    bs.ld = bs.sz = ncf;
    bs.hm = calloc((size_t)ncf + 1, sizeof(hm_t *));
    if (width == 8) bs.cf_8 = calloc((size_t)ncf + 1, sizeof(cf8_t *));
    else if (width == 16) bs.cf_16 = calloc((size_t)ncf + 1, sizeof(cf16_t *));
    else bs.cf_32 = calloc((size_t)ncf + 1, sizeof(cf32_t *));
    for (int32_t k = 0; k < ncf; ++k) {
        const int64_t len = cp[k + 1] - cp[k];
        bs.hm[k] = malloc(((size_t)len + 3) * sizeof(hm_t));  /* basis element rows */
        bs.hm[k][0] = k; bs.hm[k][1] = (int32_t)(len % 4); bs.hm[k][2] = (int32_t)len;
        if (width == 8) { bs.cf_8[k] = malloc((size_t)len + 1); for (int64_t j = 0; j < len; ++j) bs.cf_8[k][j] = (cf8_t)cf[cp[k] + j]; }
        else if (width == 16) { bs.cf_16[k] = malloc(((size_t)len + 1) * 2); for (int64_t j = 0; j < len; ++j) bs.cf_16[k][j] = (cf16_t)cf[cp[k] + j]; }
        else { bs.cf_32[k] = malloc(((size_t)len + 1) * 4); for (int64_t j = 0; j < len; ++j) bs.cf_32[k][j] = (cf32_t)cf[cp[k] + j]; }
    }
    struct mat_t mat;
    memset(&mat, 0, sizeof(mat));
    mat.r = malloc(((size_t)nr + 1) * sizeof(hm_t *));
    for (int32_t i = 0; i < nr; ++i) {
        const int32_t len = (int32_t)(rp[i + 1] - rp[i]);
        hm_t *row = malloc(((size_t)len + 3) * sizeof(hm_t));
        row[0] = rcf[i]; row[1] = len % 4; row[2] = len;
        memcpy(row + 3, ci + rp[i], (size_t)len * 4);
        mat.r[i] = row;
    }
    mat.nr = mat.sz = nr; mat.nc = nc; mat.ncl = ncl; mat.ncr = nc - ncl;
    mat.nru = nr - nrl; mat.nrl = nrl;
    struct stat_t st;
    memset(&st, 0, sizeof(st));
    st.fc = p; st.laopt = laopt; st.ff_bits = width; st.nthrds = 1; st.info_level = 0;

    if (mode == 0) {
        void (*linear_algebra)(struct mat_t *, const struct bs_t *const, struct stat_t *) =
            width == 8 ? gbgpu_linear_algebra_ff_8 : width == 16 ? gbgpu_linear_algebra_ff_16
                                                                 : gbgpu_linear_algebra_ff_32;
        linear_algebra(&mat, &bs, &st);
        CHECK(st.num_zerored == nrl - mat.np, "num_zerored");
        CHECK(st.la_rtime > 0, "la_rtime");
        CHECK(bs.ld == ncf, "basis untouched");
    } else {
        void (*interreduce_matrix_rows)(struct mat_t *, struct bs_t *, const struct stat_t *) =
            width == 8 ? gbgpu_interreduce_matrix_rows_ff_8
                       : width == 16 ? gbgpu_interreduce_matrix_rows_ff_16
                                     : gbgpu_interreduce_matrix_rows_ff_32;
        interreduce_matrix_rows(&mat, &bs, &st);
        CHECK(bs.ld == 0, "basis elements freed");
    }
    CHECK(mat.nr == mat.np, "nr == np");
    FILE *o = fopen(argv[2], "wb");
    fwrite(&mat.np, 4, 1, o);
    int64_t w = 0;
    fwrite(&w, 8, 1, o);
    for (int32_t k = 0; k < mat.np; ++k) {
        CHECK(mat.r[k][0] == k, "row[0] indexes mat->cf");
        CHECK(mat.r[k][1] == mat.r[k][2] % 4, "row[1] = len % 4");
        w += mat.r[k][2];
        fwrite(&w, 8, 1, o);
    }
    for (int32_t k = 0; k < mat.np; ++k) fwrite(mat.r[k] + 3, 4, (size_t)mat.r[k][2], o);
    for (int32_t k = 0; k < mat.np; ++k)
        for (int32_t j = 0; j < mat.r[k][2]; ++j) {
            uint32_t v = width == 8 ? (uint32_t)(uint8_t)mat.cf_8[k][j]
                       : width == 16 ? (uint32_t)(uint16_t)mat.cf_16[k][j] : (uint32_t)mat.cf_32[k][j];
            fwrite(&v, 4, 1, o);
        }
    fclose(o);
    printf("adapter ok: %d rows\n", mat.np);
    return 0;
}
