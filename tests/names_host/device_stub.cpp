// CPU build of the names-stream host code (fqxv_amd/csrc/rans_names.cpp) --
// TEST INFRASTRUCTURE.  The device batch that the names code hands its rANS
// sub-streams to is replaced here by the oracle (oracle/rans_oracle.c), job
// by job, with the same contract as fqxv_jobs_host_packed: per-job status
// and length, only successful outputs packed back.  A batch whose output
// regions add up to more than FQXV_STUB_DEVICE_BYTES (default 4 GiB) fails
// with FQXV_RANS_ERR_DEVICE, as a device allocation of that size would.
// Built with AddressSanitizer and UBSan (tests/names_host/Makefile) so the
// framing walk, the host header checks and the replay run under both.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../fqxv_amd/csrc/rans_internal.h"
#include "../../oracle/rans_oracle.h"

static uint64_t device_limit()
{
    const char *e = getenv("FQXV_STUB_DEVICE_BYTES");
    return e ? strtoull(e, nullptr, 10) : (4ull << 30);
}

extern "C" int fqxv_jobs_host_packed(fqxv_rans_ctx *c, int mode, const uint8_t *h_src, uint64_t src_len,
                                     const fqxv_rans_job *h_jobs, uint32_t count, uint64_t dst_total,
                                     uint8_t **h_packed, uint64_t *h_pack_off, int32_t *h_status,
                                     uint64_t *h_out_len)
{
    if (!c || !h_packed || (mode != 0 && mode != 2)) return FQXV_RANS_ERR_INVALID_ARG;
    *h_packed = nullptr;
    if (dst_total > device_limit()) return FQXV_RANS_ERR_DEVICE;
    std::vector<std::vector<uint8_t>> outs(count);
    uint64_t total = 0;
    for (uint32_t i = 0; i < count; i++) {
        const fqxv_rans_job &j = h_jobs[i];
        if (j.src_off + j.src_len > src_len || j.dst_off + j.dst_cap > dst_total) return FQXV_RANS_ERR_INVALID_ARG;
        const uint8_t *s = h_src + j.src_off;
        size_t ol = 0;
        int st;
        if (mode == 0) {
            outs[i].resize(j.dst_cap);
            st = orc_decode(s, j.src_len, outs[i].data(), j.dst_cap, &ol, nullptr);
        } else {
            // Both orders, the smaller kept, j.order on a tie.
            std::vector<uint8_t> o0(orc_encode_bound(j.src_len, 0)), o1(orc_encode_bound(j.src_len, 1));
            size_t l0 = 0, l1 = 0;
            st = orc_encode(s, j.src_len, 0, o0.data(), o0.size(), &l0);
            if (!st) st = orc_encode(s, j.src_len, 1, o1.data(), o1.size(), &l1);
            const bool one = l1 < l0 || (l1 == l0 && j.order == 1);
            outs[i] = one ? o1 : o0;
            ol = one ? l1 : l0;
        }
        h_status[i] = st;
        h_out_len[i] = st ? 0 : ol;
        h_pack_off[i] = total;
        if (!st) total += ol;
    }
    uint8_t *h = (uint8_t *)malloc(total ? total : 1);
    if (!h) return FQXV_RANS_ERR_DEVICE;
    for (uint32_t i = 0; i < count; i++)
        if (!h_status[i] && h_out_len[i]) memcpy(h + h_pack_off[i], outs[i].data(), h_out_len[i]);
    *h_packed = h;
    return FQXV_RANS_OK;
}

extern "C" int fqxv_xxh3_host_ranges(fqxv_rans_ctx *, const uint8_t *const *, const uint64_t *, uint32_t,
                                     uint64_t *)
{
    return FQXV_RANS_ERR_DEVICE; // digests are not part of the CPU build
}

// The product's region size for an encode-best job (fqxv_rans_encode_bound of
// both orders: the oracle's bound plus 16 bytes of slack, 9 for n = 0).
extern "C" size_t fqxv_rans_encode_best_bound(size_t n)
{
    if (n == 0) return 18;
    return orc_encode_bound(n, 0) + orc_encode_bound(n, 1) + 32;
}
