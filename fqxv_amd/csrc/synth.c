/*
 * Synthetic block generators for the benchmark workloads (host C, threads).
 *
 * Every byte is a pure function of (seed, stream, index) through a
 * counter-based generator (splitmix64 finaliser over a mixed key), so any
 * block can be regenerated alone, on any host, bit for bit.
 *
 * Workloads:
 *   binned  : i.i.d. over {35,45,56,70} with p = .02/.05/.08/.85
 *   normal  : 33 + q, q = round(N(30, 8^2)) clipped to 2..41 (inverse-CDF
 *             table of integer thresholds, so no float work per byte)
 *   markov  : 150-symbol quality strings; q0 ~ U{30..40}; each next q stays
 *             with p = 0.7, else moves by +-U{1..5}; clipped to 2..41;
 *             byte 33 + q.  Strings are laid end to end from global
 *             position 0 (a block boundary may cut a string).
 *   dna     : ACGT uniform, N with p = 0.001
 *   names   : ':'-terminated decimal fields of width U{2..6}, uniform
 *             digits, restarted at every 256-byte segment (the last field
 *             of a segment may be cut).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static inline uint64_t rng(uint64_t seed, uint64_t stream, uint64_t i)
{
    return mix64(mix64(seed * 0x9e3779b97f4a7c15ull + stream) + i * 0xd1b54a32d192ed03ull);
}

/* P(x) thresholds on 32-bit draws. */
#define P32(p) ((uint64_t)((p) * 4294967296.0 + 0.5))

void synth_binned(uint8_t *out, uint64_t n, uint64_t seed, uint64_t stream, uint64_t start)
{
    const uint64_t t1 = P32(0.02), t2 = P32(0.07), t3 = P32(0.15);
    for (uint64_t i = 0; i < n; i++) {
        uint64_t r = rng(seed, stream, start + i) & 0xffffffffu;
        out[i] = r < t1 ? 35 : r < t2 ? 45 : r < t3 ? 56 : 70;
    }
}

static uint32_t normal_thr[40];
static int normal_init;

static void init_normal(void)
{
    if (normal_init) return;
    /* P(q <= k) for q = round(N(30,64)) clipped to [2,41]. */
    for (int k = 2; k <= 41; k++) {
        double c = k == 41 ? 1.0 : 0.5 * (1.0 + erf(((k + 0.5) - 30.0) / (8.0 * sqrt(2.0))));
        double t = c * 4294967296.0;
        normal_thr[k - 2] = t >= 4294967295.0 ? 0xffffffffu : (uint32_t)t;
    }
    normal_init = 1;
}

void synth_normal(uint8_t *out, uint64_t n, uint64_t seed, uint64_t stream, uint64_t start)
{
    init_normal();
    for (uint64_t i = 0; i < n; i++) {
        uint32_t r = (uint32_t)rng(seed, stream, start + i);
        int k = 0;
        while (k < 39 && r >= normal_thr[k]) k++;
        out[i] = (uint8_t)(33 + 2 + k);
    }
}

#define MARKOV_LEN 150

void synth_markov(uint8_t *out, uint64_t n, uint64_t seed, uint64_t stream, uint64_t start)
{
    const uint64_t stay = P32(0.7);
    uint64_t i = 0;
    while (i < n) {
        uint64_t g = (start + i) / MARKOV_LEN;
        uint64_t k0 = (start + i) % MARKOV_LEN;
        /* replay string g from its start up to k0 */
        int q = 30 + (int)(rng(seed, stream, g * 2 * MARKOV_LEN) % 11);
        for (uint64_t k = 0; k < MARKOV_LEN && i < n; k++) {
            if (k > 0) {
                uint64_t r = rng(seed, stream, g * 2 * MARKOV_LEN + k);
                if ((r & 0xffffffffu) >= stay) {
                    int step = 1 + (int)((r >> 32) % 5);
                    if ((r >> 63) & 1) step = -step;
                    q += step;
                    if (q < 2) q = 2;
                    if (q > 41) q = 41;
                }
            }
            if (k >= k0) out[i++] = (uint8_t)(33 + q);
        }
    }
}

void synth_dna(uint8_t *out, uint64_t n, uint64_t seed, uint64_t stream, uint64_t start)
{
    static const uint8_t b[4] = {'A', 'C', 'G', 'T'};
    const uint64_t pn = P32(0.001);
    for (uint64_t i = 0; i < n; i++) {
        uint64_t r = rng(seed, stream, start + i);
        out[i] = (r & 0xffffffffu) < pn ? 'N' : b[r >> 62];
    }
}

#define SEG 256

void synth_names(uint8_t *out, uint64_t n, uint64_t seed, uint64_t stream, uint64_t start)
{
    uint64_t i = 0;
    while (i < n) {
        uint64_t seg = (start + i) / SEG, off = (start + i) % SEG;
        uint64_t p = 0, d = 0;
        while (p < SEG && i < n) {
            uint64_t r = rng(seed, stream, seg * 4096 + d++);
            int w = 2 + (int)(r % 5);
            for (int k = 0; k <= w && p < SEG; k++, p++) {
                uint8_t c = k == w ? ':' : (uint8_t)('0' + (rng(seed, stream, seg * 4096 + d++) % 10));
                if (p >= off && i < n) out[i++] = c;
            }
        }
    }
}

enum { K_BINNED = 0, K_NORMAL = 1, K_MARKOV = 2, K_DNA = 3, K_NAMES = 4 };

void synth_fill(int kind, uint8_t *out, uint64_t n, uint64_t seed, uint64_t stream, uint64_t start)
{
    switch (kind) {
    case K_BINNED: synth_binned(out, n, seed, stream, start); break;
    case K_NORMAL: synth_normal(out, n, seed, stream, start); break;
    case K_MARKOV: synth_markov(out, n, seed, stream, start); break;
    case K_DNA: synth_dna(out, n, seed, stream, start); break;
    default: synth_names(out, n, seed, stream, start); break;
    }
}

/* Mixed corpus layout: sizes log-uniform on [lo, hi] bytes, order 0/1 at
 * 50/50, kind uniform over {markov, dna, names}. */
void synth_mixed_layout(uint64_t seed, uint32_t count, uint64_t lo, uint64_t hi, uint64_t *sizes,
                        int32_t *orders, int32_t *kinds)
{
    const double a = log((double)lo), b = log((double)hi);
    for (uint32_t j = 0; j < count; j++) {
        uint64_t r = rng(seed, 0xabcdef, j);
        double u = (double)(r >> 11) * (1.0 / 9007199254740992.0);
        uint64_t s = (uint64_t)floor(exp(a + u * (b - a)));
        if (s < lo) s = lo;
        if (s > hi) s = hi;
        sizes[j] = s;
        uint64_t r2 = rng(seed, 0xabcdee, j);
        orders[j] = (int32_t)(r2 & 1);
        static const int32_t kk[3] = {K_MARKOV, K_DNA, K_NAMES};
        kinds[j] = kk[(r2 >> 32) % 3];
    }
}

typedef struct {
    const int32_t *kinds;
    const uint64_t *sizes, *offs, *starts;
    uint64_t seed;
    uint8_t *arena;
    uint32_t count;
    uint32_t next;
    pthread_mutex_t mu;
} fill_job;

static void *fill_worker(void *arg)
{
    fill_job *f = (fill_job *)arg;
    for (;;) {
        pthread_mutex_lock(&f->mu);
        uint32_t j = f->next++;
        pthread_mutex_unlock(&f->mu);
        if (j >= f->count) break;
        /* With `starts`, blocks are consecutive cuts of one stream (id 0). */
        synth_fill(f->kinds[j], f->arena + f->offs[j], f->sizes[j], f->seed, f->starts ? 0 : j,
                   f->starts ? f->starts[j] : 0);
    }
    return NULL;
}

/* Fill `count` blocks into `arena` (block j at offs[j], stream id j). */
void synth_fill_blocks(const int32_t *kinds, const uint64_t *sizes, const uint64_t *offs,
                       const uint64_t *starts, uint32_t count, uint64_t seed, uint8_t *arena,
                       int threads)
{
    fill_job f = {.kinds = kinds, .sizes = sizes, .offs = offs, .starts = starts, .seed = seed,
                  .arena = arena, .count = count, .next = 0};
    pthread_mutex_init(&f.mu, NULL);
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    pthread_t tid[64];
    int started = 0;
    for (int t = 1; t < threads; t++)
        if (pthread_create(&tid[started], NULL, fill_worker, &f) == 0) started++;
    fill_worker(&f);
    for (int t = 0; t < started; t++) pthread_join(tid[t], NULL);
    pthread_mutex_destroy(&f.mu);
}
