/*
 * C test driver for the C ABI (include/fqxv_rans_hip.h), linked the way a
 * Rust `-sys` crate links libfqxv_rans_hip.so: plain C, no Python, no torch.
 *
 *   abi_driver --probe         link check only: no device calls
 *   abi_driver CASES_FILE      run every case on HIP device 0
 *
 * CASES_FILE: "FQXVDRV1", u32 count, then per case
 *   u32 order, u64 raw_len, raw bytes, u64 enc_len, enc bytes
 * where enc is the stream the reference encoder (or the oracle) wrote for
 * raw at that order.  Checked per case, through the host-buffer calls
 * (lib.rs:111-127, 186-200, 226-231):
 *   fqxv_rans_encode(raw, order)           == enc
 *   fqxv_rans_decode_bounded(enc, raw_len) == raw
 *   fqxv_rans_decode_bounded(enc, raw_len - 1) fails with FQXV_RANS_ERR_BOUND
 *   fqxv_rans_decode_with(enc, GFX950_BATCHED) == raw
 * and then all cases at once through the device-resident batched calls:
 *   fqxv_rans_decode_batch over an arena of every enc   == every raw
 *   fqxv_rans_encode_batch over an arena of every raw   == every enc
 * Exit status 0 when everything matches; the first mismatch is printed.
 */
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/fqxv_rans_hip.h"

typedef struct {
    uint32_t order;
    uint64_t raw_len, enc_len;
    uint8_t *raw, *enc;
} drv_case;

static int fail(const char *what, uint32_t i, int rc)
{
    fprintf(stderr, "abi_driver: case %u: %s (rc %d: %s)\n", i, what, rc, fqxv_rans_strerror(rc));
    return 1;
}

static int read_exact(FILE *f, void *p, size_t n) { return n == 0 || fread(p, 1, n, f) == n; }

static drv_case *load_cases(const char *path, uint32_t *count)
{
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    char magic[8];
    drv_case *cs = NULL;
    if (!read_exact(f, magic, 8) || memcmp(magic, "FQXVDRV1", 8) || !read_exact(f, count, 4)) goto bad;
    cs = (drv_case *)calloc(*count ? *count : 1, sizeof *cs);
    if (!cs) goto bad;
    for (uint32_t i = 0; i < *count; i++) {
        drv_case *c = &cs[i];
        if (!read_exact(f, &c->order, 4) || !read_exact(f, &c->raw_len, 8)) goto bad;
        c->raw = (uint8_t *)malloc(c->raw_len + 1);
        if (!c->raw || !read_exact(f, c->raw, c->raw_len) || !read_exact(f, &c->enc_len, 8)) goto bad;
        c->enc = (uint8_t *)malloc(c->enc_len + 1);
        if (!c->enc || !read_exact(f, c->enc, c->enc_len)) goto bad;
    }
    fclose(f);
    return cs;
bad:
    fclose(f);
    free(cs);
    return NULL;
}

static int host_calls(fqxv_rans_ctx *ctx, const drv_case *c, uint32_t i)
{
    size_t cap = fqxv_rans_encode_bound(c->raw_len, (int)c->order), ol = 0;
    uint8_t *buf = (uint8_t *)malloc(cap + c->raw_len + 1);
    uint8_t *out = buf + cap;
    int rc = fqxv_rans_encode(ctx, c->raw, c->raw_len, (int)c->order, buf, cap, &ol);
    if (rc) return free(buf), fail("encode", i, rc);
    if (ol != c->enc_len || memcmp(buf, c->enc, ol)) return free(buf), fail("encode differs", i, 0);
    rc = fqxv_rans_decode_bounded(ctx, c->enc, c->enc_len, c->raw_len, out, c->raw_len, &ol);
    if (rc) return free(buf), fail("decode_bounded", i, rc);
    if (ol != c->raw_len || memcmp(out, c->raw, ol)) return free(buf), fail("decode differs", i, 0);
    if (c->raw_len) {
        rc = fqxv_rans_decode_bounded(ctx, c->enc, c->enc_len, c->raw_len - 1, out, c->raw_len, &ol);
        if (rc != FQXV_RANS_ERR_BOUND || ol != 0) return free(buf), fail("bound not enforced", i, rc);
    }
    memset(out, 0, c->raw_len);
    rc = fqxv_rans_decode_with(ctx, FQXV_RANS_BACKEND_GFX950_BATCHED, c->enc, c->enc_len, out, c->raw_len, &ol);
    if (rc) return free(buf), fail("decode_with(batched)", i, rc);
    if (ol != c->raw_len || memcmp(out, c->raw, ol)) return free(buf), fail("decode_with differs", i, 0);
    free(buf);
    return 0;
}

/* Every case in one device arena: batched decode of the streams, batched
 * encode of the raw blocks. */
static int batched_calls(fqxv_rans_ctx *ctx, const drv_case *cs, uint32_t n)
{
    uint64_t enc_total = 64, raw_total = 64, bound_total = 64;
    fqxv_rans_job *dj = (fqxv_rans_job *)calloc(n, sizeof *dj), *ej = (fqxv_rans_job *)calloc(n, sizeof *ej);
    for (uint32_t i = 0; i < n; i++) {
        dj[i].src_off = enc_total;
        dj[i].src_len = cs[i].enc_len;
        dj[i].dst_off = raw_total;
        dj[i].dst_cap = cs[i].raw_len;
        dj[i].max_len = UINT64_MAX;
        ej[i].src_off = raw_total;
        ej[i].src_len = cs[i].raw_len;
        ej[i].dst_off = bound_total;
        ej[i].dst_cap = fqxv_rans_encode_bound(cs[i].raw_len, (int)cs[i].order);
        ej[i].order = (int32_t)cs[i].order;
        enc_total += cs[i].enc_len + 7; /* odd gaps: no alignment assumed */
        raw_total += cs[i].raw_len + 3;
        bound_total += ej[i].dst_cap;
    }
    uint8_t *h_enc = (uint8_t *)calloc(enc_total, 1), *h_raw = (uint8_t *)calloc(raw_total, 1);
    uint8_t *h_back = (uint8_t *)malloc(bound_total > raw_total ? bound_total : raw_total);
    for (uint32_t i = 0; i < n; i++) {
        memcpy(h_enc + dj[i].src_off, cs[i].enc, cs[i].enc_len);
        memcpy(h_raw + dj[i].dst_off, cs[i].raw, cs[i].raw_len);
    }
    uint8_t *d_enc = NULL, *d_raw = NULL, *d_out = NULL;
    fqxv_rans_job *d_dj = NULL, *d_ej = NULL;
    int32_t *d_st = NULL;
    uint64_t *d_ol = NULL;
    int32_t *st = (int32_t *)malloc(4 * (size_t)n);
    uint64_t *ol = (uint64_t *)malloc(8 * (size_t)n);
    int bad = 1;
    if (hipMalloc((void **)&d_enc, enc_total) || hipMalloc((void **)&d_raw, raw_total) ||
        hipMalloc((void **)&d_out, bound_total > raw_total ? bound_total : raw_total) ||
        hipMalloc((void **)&d_dj, n * sizeof *dj) || hipMalloc((void **)&d_ej, n * sizeof *ej) ||
        hipMalloc((void **)&d_st, 4 * (size_t)n) || hipMalloc((void **)&d_ol, 8 * (size_t)n) ||
        hipMemcpy(d_enc, h_enc, enc_total, hipMemcpyHostToDevice) ||
        hipMemcpy(d_raw, h_raw, raw_total, hipMemcpyHostToDevice) ||
        hipMemcpy(d_dj, dj, n * sizeof *dj, hipMemcpyHostToDevice) ||
        hipMemcpy(d_ej, ej, n * sizeof *ej, hipMemcpyHostToDevice) ||
        hipMemset(d_out, 0xEE, raw_total)) {
        fprintf(stderr, "abi_driver: HIP allocation or copy failed\n");
        goto out;
    }
    int rc = fqxv_rans_decode_batch(ctx, d_enc, d_out, d_dj, n, NULL, d_st, d_ol, NULL);
    if (rc || hipDeviceSynchronize() || hipMemcpy(h_back, d_out, raw_total, hipMemcpyDeviceToHost) ||
        hipMemcpy(st, d_st, 4 * (size_t)n, hipMemcpyDeviceToHost) ||
        hipMemcpy(ol, d_ol, 8 * (size_t)n, hipMemcpyDeviceToHost)) {
        fprintf(stderr, "abi_driver: decode_batch failed (rc %d)\n", rc);
        goto out;
    }
    for (uint32_t i = 0; i < n; i++) {
        if (st[i] || ol[i] != cs[i].raw_len || memcmp(h_back + dj[i].dst_off, cs[i].raw, cs[i].raw_len)) {
            fail("decode_batch differs", i, st[i]);
            goto out;
        }
        if (h_back[dj[i].dst_off + cs[i].raw_len] != 0xEE) {
            fail("decode_batch wrote past its region", i, 0);
            goto out;
        }
    }
    rc = fqxv_rans_encode_batch(ctx, d_raw, d_out, d_ej, n, NULL, d_st, d_ol, NULL);
    if (rc || hipDeviceSynchronize() || hipMemcpy(h_back, d_out, bound_total, hipMemcpyDeviceToHost) ||
        hipMemcpy(st, d_st, 4 * (size_t)n, hipMemcpyDeviceToHost) ||
        hipMemcpy(ol, d_ol, 8 * (size_t)n, hipMemcpyDeviceToHost)) {
        fprintf(stderr, "abi_driver: encode_batch failed (rc %d)\n", rc);
        goto out;
    }
    for (uint32_t i = 0; i < n; i++) {
        if (st[i] || ol[i] != cs[i].enc_len || memcmp(h_back + ej[i].dst_off, cs[i].enc, cs[i].enc_len)) {
            fail("encode_batch differs", i, st[i]);
            goto out;
        }
    }
    bad = 0;
out:
    (void)hipFree(d_enc);
    (void)hipFree(d_raw);
    (void)hipFree(d_out);
    (void)hipFree(d_dj);
    (void)hipFree(d_ej);
    (void)hipFree(d_st);
    (void)hipFree(d_ol);
    free(dj);
    free(ej);
    free(h_enc);
    free(h_raw);
    free(h_back);
    free(st);
    free(ol);
    return bad;
}

int main(int argc, char **argv)
{
    if (argc == 2 && !strcmp(argv[1], "--probe")) {
        /* Link check without a device: pure host functions only. */
        uint64_t n = 0;
        const uint8_t hdr[9] = {1, 0x39, 0x30, 0, 0, 0, 0, 0, 0};
        const int ok = !strcmp(fqxv_rans_strerror(FQXV_RANS_ERR_TRUNCATED), "truncated stream") &&
                       fqxv_rans_encode_bound(0, 0) == 9 && fqxv_rans_peek_len(hdr, 9, &n) == 0 && n == 12345 &&
                       fqxv_rans_backend_detect() == FQXV_RANS_BACKEND_GFX950;
        printf("abi_driver probe %s\n", ok ? "ok" : "FAILED");
        return ok ? 0 : 1;
    }
    if (argc != 2) {
        fprintf(stderr, "usage: abi_driver --probe | abi_driver CASES_FILE\n");
        return 2;
    }
    uint32_t n = 0;
    drv_case *cs = load_cases(argv[1], &n);
    if (!cs) {
        fprintf(stderr, "abi_driver: cannot read %s\n", argv[1]);
        return 2;
    }
    fqxv_rans_ctx *ctx = NULL;
    int rc = fqxv_rans_ctx_create(0, &ctx);
    if (rc) return fail("ctx_create", 0, rc);
    int bad = 0;
    for (uint32_t i = 0; i < n && !bad; i++) bad = host_calls(ctx, &cs[i], i);
    if (!bad) bad = batched_calls(ctx, cs, n);
    fqxv_rans_ctx_destroy(ctx);
    if (!bad) printf("abi_driver: %u cases ok (host calls, decode_with, batched decode and encode)\n", n);
    return bad;
}
