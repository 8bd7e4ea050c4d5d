/*
 * CPU oracle for the fqxv-rans Nx16 coder -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of crates/fqxv-rans/src/{model.rs,scalar.rs,lib.rs}
 * used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
 * the checker.  The product library (libfqxv_rans_hip.so) never links, loads
 * or calls anything in this directory.
 *
 * Error codes deliberately carry the same numeric values as the product's
 * include/fqxv_rans_hip.h so tests can compare outcomes directly.
 */
#ifndef FQXV_RANS_ORACLE_H
#define FQXV_RANS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_OK = 0,
    ORC_ERR_TRUNCATED = 1,        /* "truncated stream"                       */
    ORC_ERR_UNKNOWN_ORDER = 2,    /* "unknown order tag"                      */
    ORC_ERR_FREQ_EXCEEDS = 3,     /* "rANS freq table exceeds TOTFREQ"        */
    ORC_ERR_FREQ_SUM = 4,         /* "rANS freq table does not sum to TOTFREQ"*/
    ORC_ERR_OUTPUT_ALLOC = 5,     /* "rANS output length too large to allocate"*/
    ORC_ERR_ABSENT_CONTEXT = 6,   /* "symbol references absent context"       */
    ORC_ERR_BOUND = 7,            /* "rANS output length exceeds caller bound"*/
    ORC_ERR_DST_TOO_SMALL = 8,    /* encode: destination capacity too small   */
    ORC_ERR_INVALID_ARG = 9,
};

/* model.rs:10-19, :36 */
#define ORC_SCALE_BITS 12u
#define ORC_TOTFREQ (1u << ORC_SCALE_BITS)
#define ORC_RANS_L (1u << 16)
#define ORC_N_STATES 32u
#define ORC_RCP_FREQ_MAX (1u << 11)

/* model.rs:51-64 EncSym */
typedef struct {
    uint64_t x_max;
    uint32_t rcp_freq, rcp_shift, bias, cmpl_freq, freq, cum;
    int use_div;
} orc_encsym;

void orc_normalize(const uint32_t counts[256], uint16_t freq[256]);
void orc_encsym_new(uint32_t freq, uint32_t cum, orc_encsym *out);
uint32_t orc_encsym_apply(const orc_encsym *e, uint32_t v);
/* model.rs:160-174: 0 / ORC_ERR_FREQ_EXCEEDS / ORC_ERR_FREQ_SUM */
int orc_validate_freqs(const uint16_t freq[256]);

size_t orc_encode_bound(size_t n, int order);
/* lib.rs:111-127 (scalar.rs:29-111) */
int orc_encode(const uint8_t *src, size_t n, int order, uint8_t *dst,
               size_t cap, size_t *out_len);
/* scalar.rs:143-155.  cap plays the role of the fallible allocation in
 * model.rs:199-205: a declared n larger than cap is ORC_ERR_OUTPUT_ALLOC.
 * *consumed (may be NULL) receives the number of stream bytes read. */
int orc_decode(const uint8_t *src, size_t len, uint8_t *dst, size_t cap,
               size_t *out_len, size_t *consumed);
/* lib.rs:207-231 */
int orc_peek_len(const uint8_t *src, size_t len, uint64_t *n);
int orc_decode_bounded(const uint8_t *src, size_t len, size_t max_len,
                       uint8_t *dst, size_t cap, size_t *out_len);

/* Multi-threaded helpers for the CPU baseline: code `count` independent
 * blocks over `threads` pthreads (block-level parallelism as in the
 * reference's rayon par_iter).  Offsets index the given arenas. */
int orc_encode_many(const uint8_t *src, const uint64_t *src_off,
                    const uint64_t *src_len, const int *order, size_t count,
                    uint8_t *dst, const uint64_t *dst_off,
                    const uint64_t *dst_cap, uint64_t *out_len, int *status,
                    int threads);
int orc_decode_many(const uint8_t *src, const uint64_t *src_off,
                    const uint64_t *src_len, size_t count, uint8_t *dst,
                    const uint64_t *dst_off, const uint64_t *dst_cap,
                    uint64_t *out_len, int *status, int threads);

#ifdef __cplusplus
}
#endif
/* CRC-32C of crates/fqxv/src/crc.rs (container integrity checksums). */
uint32_t orc_crc32c_update(uint32_t state, const uint8_t *p, size_t n);
uint32_t orc_crc32c(const uint8_t *p, size_t n);
uint32_t orc_crc32c_combine(uint32_t crc_a, uint32_t crc_b, uint64_t len_b);
void orc_crc32c_many(const uint8_t *src, const uint64_t *off, const uint64_t *len, uint32_t count,
                     uint32_t *crc_out);

#endif
