// Driver of the CPU build of the names host code (see device_stub.cpp).
//   names_host decode IN OUT : IN = u32 count, then per stream u64 len + bytes;
//                              OUT = per stream i32 status, u64 size, records
//   names_host encode IN OUT : IN = per block the records buffer, same framing;
//                              OUT = per block i32 status, u64 size, stream
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/fqxv_rans_hip.h"

static bool read_all(const char *path, std::vector<uint8_t> &buf)
{
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    uint8_t tmp[1 << 16];
    size_t n;
    while ((n = fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    fclose(f);
    return true;
}

int main(int argc, char **argv)
{
    if (argc != 4 || (strcmp(argv[1], "decode") && strcmp(argv[1], "encode"))) {
        fprintf(stderr, "usage: names_host decode|encode IN OUT\n");
        return 2;
    }
    std::vector<uint8_t> in;
    if (!read_all(argv[2], in) || in.size() < 4) return 2;
    uint32_t count;
    memcpy(&count, in.data(), 4);
    std::vector<uint64_t> off(count), len(count);
    size_t pos = 4;
    for (uint32_t i = 0; i < count; i++) {
        if (pos + 8 > in.size()) return 2;
        memcpy(&len[i], in.data() + pos, 8);
        off[i] = pos + 8;
        pos += 8 + len[i];
        if (pos > in.size()) return 2;
    }
    // Any non-null context: the stub does not use it.
    fqxv_rans_ctx *ctx = reinterpret_cast<fqxv_rans_ctx *>(&count);
    fqxv_names_batch *b = nullptr;
    const bool dec = !strcmp(argv[1], "decode");
    int rc = dec ? fqxv_names_decode_batch(ctx, in.data(), off.data(), len.data(), count, &b)
                 : fqxv_names_encode_batch(ctx, in.data(), off.data(), len.data(), count, &b);
    if (rc) {
        fprintf(stderr, "names_host: batch failed with %d\n", rc);
        return 1;
    }
    FILE *f = fopen(argv[3], "wb");
    if (!f) return 2;
    for (uint32_t i = 0; i < count; i++) {
        const uint8_t *p = nullptr;
        uint64_t size = 0, nn = 0;
        int32_t st = dec ? fqxv_names_records(b, i, &p, &size, &nn) : fqxv_names_output(b, i, &p, &size);
        if (st) size = 0;
        fwrite(&st, 4, 1, f);
        fwrite(&size, 8, 1, f);
        if (size) fwrite(p, 1, size, f);
    }
    fclose(f);
    fqxv_names_batch_destroy(b);
    return 0;
}
