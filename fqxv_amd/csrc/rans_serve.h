// Shared layout of the persistent decode server's queue (device kernel
// k_rans_serve in rans_decode.hip, host side in rans_server.cpp).
//
// The host publishes block jobs into a ring of `ring` records in pinned,
// coherent host memory and then raises `tail`; the kernel's dispatcher wave
// mirrors `tail` and `stop` into device memory, decode waves claim published
// jobs with a compare-and-swap on `head`, and each finished job writes its
// result record to pinned host memory, `tag` (= job index + 1) last.
#pragma once
#include <cstdint>

struct fqxv_serve_ctl {   // pinned host memory, written by the host
    uint64_t tail;        // jobs published so far (monotonic)
    uint32_t stop;        // set once: no more jobs, exit when idle
    uint32_t pad;
};

struct fqxv_serve_res {   // pinned host memory, one per ring slot, written by the device
    uint64_t tag;         // job index + 1 once the job is done
    int32_t status;
    uint32_t pad;
    uint64_t out_len;
};

struct fqxv_serve_dev {   // device memory; the claim counter and the read-mostly
                          // published count sit on separate 128-byte lines
    uint64_t head;        // jobs claimed so far (compare-and-swap)
    uint32_t alive;       // decode waves still running
    uint32_t disp_cu;     // the dispatcher's compute unit (set once, 0 before)
    uint32_t pad0[28];
    uint64_t dtail;       // device copy of the published count
    uint32_t dstop;       // device copy of stop
    uint32_t pad1[29];
};
