// Library-internal entry points used by the batching server (server.cpp).
// They are exported from libwhirgpu.so but are not part of the public C ABI
// of include/whir_gpu.h.  server.cpp depends on the prover only through these
// and the public ABI, so its threading can be checked on a host without a GPU
// against a stand-in library.
#ifndef WHIRGPU_INTERNAL_H
#define WHIRGPU_INTERNAL_H

#include <stddef.h>

#include "../../include/whir_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Device bytes a context holds right now: its fixed buffers plus every
 * allocation of its buffer pool (pooled buffers are kept between proofs). */
size_t whirgpu_internal_ctx_bytes(const whirgpu_ctx* ctx);

/* Free and total device memory of `device` (hipMemGetInfo). */
int whirgpu_internal_mem_info(int device, size_t* free_bytes, size_t* total_bytes);

#ifdef __cplusplus
}
#endif

#endif /* WHIRGPU_INTERNAL_H */
