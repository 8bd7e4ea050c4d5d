/*
 * CPU oracle for the fqxv-rans Nx16 coder -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the reference's scalar coder.  Each function cites
 * the file:line of crates/fqxv-rans it follows.  Only tests/, the smoke entry
 * point and the bench's CPU-baseline leg use this code, as the checker; the
 * shipped HIP library is independent of it.
 *
 * Pinning (see tests/test_oracle.py):
 *   - closed forms: empty input, single-symbol input, n <= 32 inputs;
 *   - re-encode identity on the rANS streams embedded in the reference's
 *     golden .fqxv archives (tests/golden/fixture_streams.bin);
 *   - EncSym reciprocal == exact division sweep (model.rs:313-365);
 *   - normalize invariants (model.rs:270-310) and brute force on tiny inputs.
 */
#include "rans_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- model.rs */

/* model.rs:213-264 `normalize` */
void orc_normalize(const uint32_t counts[256], uint16_t freq[256])
{
    uint64_t total = 0;
    for (int s = 0; s < 256; s++) total += counts[s];
    memset(freq, 0, 256 * sizeof(uint16_t));
    if (total == 0) return;

    int64_t sum = 0;
    int max_s = 0;
    uint32_t max_f = 0;
    for (int s = 0; s < 256; s++) {
        if (counts[s] == 0) continue;
        uint32_t f = (uint32_t)(((uint64_t)counts[s] * ORC_TOTFREQ) / total);
        if (f == 0) f = 1;
        freq[s] = (uint16_t)f;
        sum += f;
        if (f > max_f) { /* strictly greater: the first maximum wins */
            max_f = f;
            max_s = s;
        }
    }
    int64_t diff = (int64_t)ORC_TOTFREQ - sum;
    while (diff != 0) {
        if (diff > 0) {
            freq[max_s] = (uint16_t)(freq[max_s] + diff);
            diff = 0;
        } else {
            int64_t take = (int64_t)freq[max_s] - 1;
            if (take < 0) take = 0;
            if (take > -diff) take = -diff;
            freq[max_s] = (uint16_t)(freq[max_s] - take);
            diff += take;
            if (diff != 0) {
                int nm = -1;
                int64_t nf = 1;
                for (int s = 0; s < 256; s++) {
                    if ((int64_t)freq[s] > nf) {
                        nf = freq[s];
                        nm = s;
                    }
                }
                if (nm < 0) return; /* unreachable: sum > TOTFREQ implies a bucket > 1 */
                max_s = nm;
            }
        }
    }
}

/* model.rs:70-119 `EncSym::new` */
void orc_encsym_new(uint32_t freq, uint32_t cum, orc_encsym *e)
{
    e->x_max = ((uint64_t)(ORC_RANS_L >> ORC_SCALE_BITS) << 16) * (uint64_t)freq;
    e->freq = freq;
    e->cum = cum;
    if (freq < 2) {
        e->rcp_freq = 0xffffffffu;
        e->rcp_shift = 0;
        e->bias = cum + (1u << ORC_SCALE_BITS) - 1;
        e->cmpl_freq = (1u << ORC_SCALE_BITS) - freq;
        e->use_div = 0;
    } else if (freq > ORC_RCP_FREQ_MAX) {
        e->rcp_freq = 0;
        e->rcp_shift = 0;
        e->bias = 0;
        e->cmpl_freq = 0;
        e->use_div = 1;
    } else {
        uint32_t shift = 0;
        while (freq > (1u << shift)) shift++;
        uint64_t num = 1ull << (shift + 31);
        e->rcp_freq = (uint32_t)((num + freq - 1) / freq);
        e->rcp_shift = shift - 1;
        e->bias = cum;
        e->cmpl_freq = (1u << ORC_SCALE_BITS) - freq;
        e->use_div = 0;
    }
}

/* model.rs:125-133 `EncSym::apply` (u32 wrapping arithmetic) */
uint32_t orc_encsym_apply(const orc_encsym *e, uint32_t v)
{
    if (e->use_div)
        return ((v / e->freq) << ORC_SCALE_BITS) + (v % e->freq) + e->cum;
    uint32_t q = (uint32_t)(((uint64_t)v * e->rcp_freq) >> 32) >> e->rcp_shift;
    return v + e->bias + q * e->cmpl_freq;
}

/* model.rs:160-174 `Model::from_freqs` validation */
int orc_validate_freqs(const uint16_t freq[256])
{
    uint32_t acc = 0;
    for (int s = 0; s < 256; s++) {
        acc += freq[s];
        if (acc > ORC_TOTFREQ) return ORC_ERR_FREQ_EXCEEDS;
    }
    return acc == ORC_TOTFREQ ? ORC_OK : ORC_ERR_FREQ_SUM;
}

typedef struct {
    uint16_t freq[256];
    uint16_t cum[257];
    uint8_t slot2sym[ORC_TOTFREQ];
} orc_model;

/* model.rs:180-194 `from_freqs_unchecked` */
static void model_build(orc_model *m, const uint16_t freq[256])
{
    memcpy(m->freq, freq, sizeof(m->freq));
    m->cum[0] = 0;
    for (int s = 0; s < 256; s++) m->cum[s + 1] = (uint16_t)(m->cum[s] + freq[s]);
    memset(m->slot2sym, 0, sizeof(m->slot2sym));
    for (int s = 0; s < 256; s++)
        for (uint32_t k = m->cum[s]; k < m->cum[s + 1]; k++) m->slot2sym[k] = (uint8_t)s;
}

/* --------------------------------------------------------------- scalar.rs */

static void put16(uint8_t *p, uint16_t v) { p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); }
static void put32(uint8_t *p, uint32_t v) { for (int i = 0; i < 4; i++) p[i] = (uint8_t)(v >> (8 * i)); }
static void put64(uint8_t *p, uint64_t v) { for (int i = 0; i < 8; i++) p[i] = (uint8_t)(v >> (8 * i)); }

size_t orc_encode_bound(size_t n, int order)
{
    if (n == 0) return 9;
    size_t hdr = 9 + 2 * 256 + 4 * ORC_N_STATES;
    if (order == 1) {
        size_t k = n < 256 ? n : 256;
        hdr = 9 + 32 + 512 * k + 4 * ORC_N_STATES;
    }
    return hdr + 2 * n;
}

/* scalar.rs:120-127 `encode_symbol`: words are collected in emission order. */
static inline uint32_t enc_step(uint32_t x, const orc_encsym *e, uint16_t *ren, size_t *nw)
{
    uint32_t v = x;
    while ((uint64_t)v >= e->x_max) {
        ren[(*nw)++] = (uint16_t)(v & 0xffff);
        v >>= 16;
    }
    return orc_encsym_apply(e, v);
}

/* scalar.rs:130-140 `finish`: states, then the renorm words reversed. */
static void finish(uint8_t *out, const uint32_t *states, const uint16_t *ren, size_t nw)
{
    for (uint32_t j = 0; j < ORC_N_STATES; j++) put32(out + 4 * j, states[j]);
    uint8_t *w = out + 4 * ORC_N_STATES;
    for (size_t k = 0; k < nw; k++) put16(w + 2 * k, ren[nw - 1 - k]);
}

/* scalar.rs:29-55 `encode_order0` */
static int encode_o0(const uint8_t *src, size_t n, uint8_t *dst, size_t cap, size_t *out_len)
{
    uint32_t counts[256] = {0};
    for (size_t i = 0; i < n; i++) counts[src[i]]++;
    uint16_t freq[256];
    orc_normalize(counts, freq);
    orc_encsym enc[256];
    uint32_t cum = 0;
    for (int s = 0; s < 256; s++) {
        orc_encsym_new(freq[s], cum, &enc[s]);
        cum += freq[s];
    }
    uint32_t states[ORC_N_STATES];
    for (uint32_t j = 0; j < ORC_N_STATES; j++) states[j] = ORC_RANS_L;
    uint16_t *ren = n ? (uint16_t *)malloc(n * sizeof(uint16_t)) : NULL;
    if (n && !ren) return ORC_ERR_OUTPUT_ALLOC;
    size_t nw = 0;
    for (size_t i = n; i-- > 0;) {
        uint32_t *x = &states[i % ORC_N_STATES];
        *x = enc_step(*x, &enc[src[i]], ren, &nw);
    }
    size_t total = n ? 9 + 512 + 4 * ORC_N_STATES + 2 * nw : 9;
    if (total > cap) {
        free(ren);
        return ORC_ERR_DST_TOO_SMALL;
    }
    dst[0] = 0;
    put64(dst + 1, (uint64_t)n);
    if (n) {
        for (int s = 0; s < 256; s++) put16(dst + 9 + 2 * s, freq[s]);
        finish(dst + 9 + 512, states, ren, nw);
    }
    free(ren);
    *out_len = total;
    return ORC_OK;
}

/* scalar.rs:61-111 `encode_order1` */
static int encode_o1(const uint8_t *src, size_t n, uint8_t *dst, size_t cap, size_t *out_len)
{
    uint32_t(*counts)[256] = (uint32_t(*)[256])calloc(256, sizeof(*counts));
    orc_encsym(*enc)[256] = (orc_encsym(*)[256])malloc(256 * sizeof(*enc));
    uint16_t(*freq)[256] = (uint16_t(*)[256])malloc(256 * sizeof(*freq));
    uint16_t *ren = n ? (uint16_t *)malloc(n * sizeof(uint16_t)) : NULL;
    int rc = ORC_OK;
    if (!counts || !enc || !freq || (n && !ren)) {
        rc = ORC_ERR_OUTPUT_ALLOC;
        goto out;
    }
    uint8_t prev = 0;
    for (size_t i = 0; i < n; i++) {
        counts[prev][src[i]]++;
        prev = src[i];
    }
    int present[256];
    int k = 0;
    for (int c = 0; c < 256; c++) {
        present[c] = 0;
        for (int s = 0; s < 256; s++)
            if (counts[c][s]) { present[c] = 1; break; }
        if (!present[c]) continue;
        k++;
        orc_normalize(counts[c], freq[c]);
        uint32_t cum = 0;
        for (int s = 0; s < 256; s++) {
            orc_encsym_new(freq[c][s], cum, &enc[c][s]);
            cum += freq[c][s];
        }
    }
    uint32_t states[ORC_N_STATES];
    for (uint32_t j = 0; j < ORC_N_STATES; j++) states[j] = ORC_RANS_L;
    size_t nw = 0;
    for (size_t i = n; i-- > 0;) {
        int ctx = i == 0 ? 0 : src[i - 1];
        uint32_t *x = &states[i % ORC_N_STATES];
        *x = enc_step(*x, &enc[ctx][src[i]], ren, &nw);
    }
    size_t hdr = 9 + 32 + 512 * (size_t)k + 4 * ORC_N_STATES;
    size_t total = n ? hdr + 2 * nw : 9;
    if (total > cap) {
        rc = ORC_ERR_DST_TOO_SMALL;
        goto out;
    }
    dst[0] = 1;
    put64(dst + 1, (uint64_t)n);
    if (n) {
        uint8_t *p = dst + 9;
        memset(p, 0, 32);
        for (int c = 0; c < 256; c++)
            if (present[c]) p[c / 8] |= (uint8_t)(1u << (c % 8));
        p += 32;
        for (int c = 0; c < 256; c++) {
            if (!present[c]) continue;
            for (int s = 0; s < 256; s++) put16(p + 2 * s, freq[c][s]);
            p += 512;
        }
        finish(p, states, ren, nw);
    }
    *out_len = total;
out:
    free(counts);
    free(enc);
    free(freq);
    free(ren);
    return rc;
}

/* lib.rs:111-127 `encode` */
int orc_encode(const uint8_t *src, size_t n, int order, uint8_t *dst, size_t cap, size_t *out_len)
{
    *out_len = 0;
    if (order == 0) return encode_o0(src, n, dst, cap, out_len);
    if (order == 1) return encode_o1(src, n, dst, cap, out_len);
    return ORC_ERR_INVALID_ARG;
}

/* scalar.rs:241-277 `Reader` */
typedef struct {
    const uint8_t *buf;
    size_t len, pos;
} reader;

static int rd(reader *r, size_t k, uint64_t *v)
{
    if (r->len - r->pos < k) return ORC_ERR_TRUNCATED;
    uint64_t x = 0;
    for (size_t i = 0; i < k; i++) x |= (uint64_t)r->buf[r->pos + i] << (8 * i);
    r->pos += k;
    *v = x;
    return ORC_OK;
}

/* scalar.rs:229-238 `step_state` */
static inline int step_state(uint32_t *x, uint32_t f, uint32_t c, reader *r)
{
    uint32_t v = f * (*x >> ORC_SCALE_BITS) + (*x & (ORC_TOTFREQ - 1)) - c;
    while (v < ORC_RANS_L) {
        uint64_t w;
        int rc = rd(r, 2, &w);
        if (rc) return rc;
        v = (v << 16) | (uint32_t)w;
    }
    *x = v;
    return ORC_OK;
}

/* scalar.rs:219-225 `read_states` */
static int read_states(reader *r, uint32_t *states)
{
    for (uint32_t j = 0; j < ORC_N_STATES; j++) {
        uint64_t v;
        int rc = rd(r, 4, &v);
        if (rc) return rc;
        states[j] = (uint32_t)v;
    }
    return ORC_OK;
}

static int read_freqs(reader *r, uint16_t *freq)
{
    for (int s = 0; s < 256; s++) {
        uint64_t v;
        int rc = rd(r, 2, &v);
        if (rc) return rc;
        freq[s] = (uint16_t)v;
    }
    return ORC_OK;
}

/* scalar.rs:157-178 `decode_order0` */
static int decode_o0(reader *r, size_t n, uint8_t *dst, size_t cap)
{
    uint16_t freq[256];
    int rc = read_freqs(r, freq);
    if (rc) return rc;
    if ((rc = orc_validate_freqs(freq))) return rc;
    orc_model *m = (orc_model *)malloc(sizeof(orc_model));
    if (!m) return ORC_ERR_OUTPUT_ALLOC;
    model_build(m, freq);
    uint32_t states[ORC_N_STATES];
    if ((rc = read_states(r, states))) goto out;
    if (n > cap) { /* model.rs:199-205 alloc_output */
        rc = ORC_ERR_OUTPUT_ALLOC;
        goto out;
    }
    for (size_t i = 0; i < n; i++) {
        uint32_t *x = &states[i % ORC_N_STATES];
        uint8_t s = m->slot2sym[*x & (ORC_TOTFREQ - 1)];
        dst[i] = s;
        if ((rc = step_state(x, m->freq[s], m->cum[s], r))) goto out;
    }
out:
    free(m);
    return rc;
}

/* scalar.rs:180-217 `decode_order1` */
static int decode_o1(reader *r, size_t n, uint8_t *dst, size_t cap)
{
    uint8_t bitmap[32];
    int rc;
    for (int b = 0; b < 32; b++) {
        uint64_t v;
        if ((rc = rd(r, 1, &v))) return rc;
        bitmap[b] = (uint8_t)v;
    }
    orc_model *models[256] = {0};
    for (int c = 0; c < 256; c++) {
        if (!((bitmap[c / 8] >> (c % 8)) & 1)) continue;
        uint16_t freq[256];
        if ((rc = read_freqs(r, freq))) goto out;
        if ((rc = orc_validate_freqs(freq))) goto out;
        models[c] = (orc_model *)malloc(sizeof(orc_model));
        if (!models[c]) {
            rc = ORC_ERR_OUTPUT_ALLOC;
            goto out;
        }
        model_build(models[c], freq);
    }
    uint32_t states[ORC_N_STATES];
    if ((rc = read_states(r, states))) goto out;
    if (n > cap) {
        rc = ORC_ERR_OUTPUT_ALLOC;
        goto out;
    }
    uint8_t prev = 0;
    for (size_t i = 0; i < n; i++) {
        int ctx = i == 0 ? 0 : prev;
        const orc_model *m = models[ctx];
        if (!m) {
            rc = ORC_ERR_ABSENT_CONTEXT;
            goto out;
        }
        uint32_t *x = &states[i % ORC_N_STATES];
        uint8_t s = m->slot2sym[*x & (ORC_TOTFREQ - 1)];
        dst[i] = s;
        if ((rc = step_state(x, m->freq[s], m->cum[s], r))) goto out;
        prev = s;
    }
out:
    for (int c = 0; c < 256; c++) free(models[c]);
    return rc;
}

/* scalar.rs:143-155 `decode` */
int orc_decode(const uint8_t *src, size_t len, uint8_t *dst, size_t cap, size_t *out_len,
               size_t *consumed)
{
    reader r = {src, len, 0};
    uint64_t order, n;
    int rc;
    *out_len = 0;
    if ((rc = rd(&r, 1, &order))) goto out;
    if ((rc = rd(&r, 8, &n))) goto out;
    if (n == 0) goto out;
    if (order == 0)
        rc = decode_o0(&r, (size_t)n, dst, cap);
    else if (order == 1)
        rc = decode_o1(&r, (size_t)n, dst, cap);
    else
        rc = ORC_ERR_UNKNOWN_ORDER;
    if (rc == ORC_OK) *out_len = (size_t)n;
out:
    if (consumed) *consumed = r.pos;
    return rc;
}

/* lib.rs:207-210 `peek_output_len` */
int orc_peek_len(const uint8_t *src, size_t len, uint64_t *n)
{
    if (len < 9) return ORC_ERR_TRUNCATED;
    uint64_t v = 0;
    for (int i = 0; i < 8; i++) v |= (uint64_t)src[1 + i] << (8 * i);
    *n = v;
    return ORC_OK;
}

/* lib.rs:226-231 `decode_bounded` */
int orc_decode_bounded(const uint8_t *src, size_t len, size_t max_len, uint8_t *dst,
                       size_t cap, size_t *out_len)
{
    uint64_t n;
    *out_len = 0;
    if (orc_peek_len(src, len, &n) == ORC_OK && n > (uint64_t)max_len) return ORC_ERR_BOUND;
    return orc_decode(src, len, dst, cap, out_len, NULL);
}

/* ------------------------------------------------ block-parallel CPU helpers */

typedef struct {
    int enc;
    const uint8_t *src;
    const uint64_t *src_off, *src_len;
    const int *order;
    uint8_t *dst;
    const uint64_t *dst_off, *dst_cap;
    uint64_t *out_len;
    int *status;
    size_t count;
    size_t next;
    pthread_mutex_t mu;
} many_job;

static void *many_worker(void *arg)
{
    many_job *j = (many_job *)arg;
    for (;;) {
        pthread_mutex_lock(&j->mu);
        size_t b = j->next++;
        pthread_mutex_unlock(&j->mu);
        if (b >= j->count) break;
        size_t ol = 0;
        int rc;
        if (j->enc)
            rc = orc_encode(j->src + j->src_off[b], j->src_len[b], j->order[b],
                            j->dst + j->dst_off[b], j->dst_cap[b], &ol);
        else
            rc = orc_decode(j->src + j->src_off[b], j->src_len[b], j->dst + j->dst_off[b],
                            j->dst_cap[b], &ol, NULL);
        j->status[b] = rc;
        j->out_len[b] = ol;
    }
    return NULL;
}

static int run_many(many_job *j, int threads)
{
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t tid[256];
    pthread_mutex_init(&j->mu, NULL);
    j->next = 0;
    int started = 0;
    for (int t = 1; t < threads; t++)
        if (pthread_create(&tid[started], NULL, many_worker, j) == 0) started++;
    many_worker(j);
    for (int t = 0; t < started; t++) pthread_join(tid[t], NULL);
    pthread_mutex_destroy(&j->mu);
    return ORC_OK;
}

int orc_encode_many(const uint8_t *src, const uint64_t *src_off, const uint64_t *src_len,
                    const int *order, size_t count, uint8_t *dst, const uint64_t *dst_off,
                    const uint64_t *dst_cap, uint64_t *out_len, int *status, int threads)
{
    many_job j = {.enc = 1, .src = src, .src_off = src_off, .src_len = src_len, .order = order,
                  .dst = dst, .dst_off = dst_off, .dst_cap = dst_cap, .out_len = out_len,
                  .status = status, .count = count};
    return run_many(&j, threads);
}

int orc_decode_many(const uint8_t *src, const uint64_t *src_off, const uint64_t *src_len,
                    size_t count, uint8_t *dst, const uint64_t *dst_off, const uint64_t *dst_cap,
                    uint64_t *out_len, int *status, int threads)
{
    many_job j = {.enc = 0, .src = src, .src_off = src_off, .src_len = src_len, .order = NULL,
                  .dst = dst, .dst_off = dst_off, .dst_cap = dst_cap, .out_len = out_len,
                  .status = status, .count = count};
    return run_many(&j, threads);
}

/* ---- CRC-32C (crates/fqxv/src/crc.rs) ------------------------------------
 * Reflected Castagnoli polynomial 0x82F63B78, init and xor-out 0xFFFFFFFF,
 * byte-wise table (crc.rs:15-36, update crc.rs:51-58), and the zlib-style
 * combine over GF(2) operator matrices (crc.rs:84-160). */
static uint32_t crc_table[256];
static int crc_table_ready;

static void crc_init_table(void)
{
    for (uint32_t i = 0; i < 256; i++) {
        uint32_t c = i;
        for (int b = 0; b < 8; b++) c = (c & 1) ? (c >> 1) ^ 0x82F63B78u : c >> 1;
        crc_table[i] = c;
    }
    crc_table_ready = 1;
}

uint32_t orc_crc32c_update(uint32_t state, const uint8_t *p, size_t n)
{
    if (!crc_table_ready) crc_init_table();
    for (size_t i = 0; i < n; i++) state = crc_table[(state ^ p[i]) & 0xff] ^ (state >> 8);
    return state;
}

uint32_t orc_crc32c(const uint8_t *p, size_t n) { return orc_crc32c_update(0xFFFFFFFFu, p, n) ^ 0xFFFFFFFFu; }

static uint32_t gf2_times(const uint32_t mat[32], uint32_t vec)
{
    uint32_t sum = 0;
    for (int i = 0; vec; i++, vec >>= 1)
        if (vec & 1) sum ^= mat[i];
    return sum;
}

static void gf2_square(uint32_t sq[32], const uint32_t mat[32])
{
    for (int n = 0; n < 32; n++) sq[n] = gf2_times(mat, mat[n]);
}

uint32_t orc_crc32c_combine(uint32_t crc_a, uint32_t crc_b, uint64_t len_b)
{
    if (len_b == 0) return crc_a;
    uint32_t odd[32], even[32];
    odd[0] = 0x82F63B78u;
    for (int i = 1; i < 32; i++) odd[i] = 1u << (i - 1);
    gf2_square(even, odd);
    gf2_square(odd, even);
    uint32_t crc = crc_a;
    for (;;) {
        gf2_square(even, odd);
        if (len_b & 1) crc = gf2_times(even, crc);
        len_b >>= 1;
        if (!len_b) break;
        gf2_square(odd, even);
        if (len_b & 1) crc = gf2_times(odd, crc);
        len_b >>= 1;
        if (!len_b) break;
    }
    return crc ^ crc_b;
}

void orc_crc32c_many(const uint8_t *src, const uint64_t *off, const uint64_t *len, uint32_t count,
                     uint32_t *crc_out)
{
    for (uint32_t i = 0; i < count; i++) crc_out[i] = orc_crc32c(src + off[i], (size_t)len[i]);
}
