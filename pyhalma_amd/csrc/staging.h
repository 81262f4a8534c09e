// Stream-ordered host->device upload of small host-built tables (work plans, offsets).
// The bytes are first copied into a pinned staging slot; the slot is reused only after
// the event recorded behind its previous copy has completed, so the caller's host
// buffer may be freed as soon as upload_async returns.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace halma {

int upload_async(void* dst_device, const void* src_host, size_t bytes, hipStream_t stream);

// Several host tables in one copy: piece q lands at dst_device + offset; the bytes of the
// region between pieces are unspecified.
struct UploadPiece {
    size_t offset;
    const void* src;
    size_t bytes;
};
int upload_pieces(void* dst_device, size_t region_bytes, const UploadPiece* pieces,
                  int n_pieces, hipStream_t stream);

// A pinned host word array for small device->host results (counters) read after a
// stream synchronisation.  One per host thread and device; grows on demand.
long long* pinned_words(int count);

}  // namespace halma
