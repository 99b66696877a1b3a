// Vulkan <-> HIP interop of the drop-in: the reference's output image and scene
// buffers are Vulkan memory (Raytracing.cpp:96-147 binds them,
// S_VkBuffer.cpp:5-22 allocates them).  Exported by the Vulkan side with
// VK_KHR_external_memory_fd (vkGetMemoryFdKHR, OPAQUE_FD) and
// VK_KHR_external_semaphore_fd, they are imported here: memory as a device
// pointer the other entry points accept, semaphores as waits / signals
// enqueued on the caller's stream around the pass.
#include "snkpt_host.h"

struct snkpt_external {
    hipExternalMemory_t mem = nullptr;
    void* ptr = nullptr;
};
struct snkpt_semaphore {
    hipExternalSemaphore_t sem = nullptr;
    bool timeline = false;
};

using namespace snkpt;

extern "C" {

snkpt_status snkpt_import_memory_fd(snkpt_ctx* ctx, int fd, uint64_t allocation_size, uint64_t offset, uint64_t size,
                                    void** out_dev_ptr, snkpt_external** out) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    if (fd < 0 || !out_dev_ptr || !out || size == 0 || offset > allocation_size || size > allocation_size - offset)
        return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_import_memory_fd: bad fd, range or output pointers");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, SNKPT_ERR_HIP, "hipSetDevice");
    hipExternalMemoryHandleDesc hd{};
    hd.type = hipExternalMemoryHandleTypeOpaqueFd;
    hd.handle.fd = fd;
    hd.size = allocation_size;
    snkpt_external* x = new snkpt_external();
    hipError_t e = hipImportExternalMemory(&x->mem, &hd);
    if (e != hipSuccess) {
        delete x;
        return hip_fail(ctx, e, "hipImportExternalMemory");
    }
    hipExternalMemoryBufferDesc bd{};
    bd.offset = offset;
    bd.size = size;
    e = hipExternalMemoryGetMappedBuffer(&x->ptr, x->mem, &bd);
    if (e != hipSuccess) {
        (void)hipDestroyExternalMemory(x->mem);
        delete x;
        return hip_fail(ctx, e, "hipExternalMemoryGetMappedBuffer");
    }
    *out_dev_ptr = x->ptr;
    *out = x;
    return SNKPT_OK;
}

snkpt_status snkpt_release_memory(snkpt_ctx* ctx, snkpt_external* x) {
    if (!ctx || !x) return SNKPT_ERR_INVALID_ARG;
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, SNKPT_ERR_HIP, "hipSetDevice");
    // the mapped range is released with the external memory object
    hipError_t e = hipDestroyExternalMemory(x->mem);
    delete x;
    return e == hipSuccess ? SNKPT_OK : hip_fail(ctx, e, "hipDestroyExternalMemory");
}

snkpt_status snkpt_import_semaphore_fd(snkpt_ctx* ctx, int fd, int timeline, snkpt_semaphore** out) {
    if (!ctx) return SNKPT_ERR_INVALID_ARG;
    if (fd < 0 || !out) return fail(ctx, SNKPT_ERR_INVALID_ARG, "snkpt_import_semaphore_fd: bad fd or output");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, SNKPT_ERR_HIP, "hipSetDevice");
    hipExternalSemaphoreHandleDesc hd{};
    hd.type = timeline ? hipExternalSemaphoreHandleTypeTimelineSemaphoreFd : hipExternalSemaphoreHandleTypeOpaqueFd;
    hd.handle.fd = fd;
    snkpt_semaphore* s = new snkpt_semaphore();
    s->timeline = timeline != 0;
    hipError_t e = hipImportExternalSemaphore(&s->sem, &hd);
    if (e != hipSuccess) {
        delete s;
        return hip_fail(ctx, e, "hipImportExternalSemaphore");
    }
    *out = s;
    return SNKPT_OK;
}

snkpt_status snkpt_semaphore_wait(snkpt_ctx* ctx, snkpt_semaphore* s, uint64_t value, snkpt_stream stream) {
    if (!ctx || !s) return SNKPT_ERR_INVALID_ARG;
    hipExternalSemaphoreWaitParams p{};
    p.params.fence.value = s->timeline ? value : 0;
    hipError_t e = hipWaitExternalSemaphoresAsync(&s->sem, &p, 1, (hipStream_t)stream);
    return e == hipSuccess ? SNKPT_OK : hip_fail(ctx, e, "hipWaitExternalSemaphoresAsync");
}

snkpt_status snkpt_semaphore_signal(snkpt_ctx* ctx, snkpt_semaphore* s, uint64_t value, snkpt_stream stream) {
    if (!ctx || !s) return SNKPT_ERR_INVALID_ARG;
    hipExternalSemaphoreSignalParams p{};
    p.params.fence.value = s->timeline ? value : 0;
    hipError_t e = hipSignalExternalSemaphoresAsync(&s->sem, &p, 1, (hipStream_t)stream);
    return e == hipSuccess ? SNKPT_OK : hip_fail(ctx, e, "hipSignalExternalSemaphoresAsync");
}

snkpt_status snkpt_release_semaphore(snkpt_ctx* ctx, snkpt_semaphore* s) {
    if (!ctx || !s) return SNKPT_ERR_INVALID_ARG;
    hipError_t e = hipDestroyExternalSemaphore(s->sem);
    delete s;
    return e == hipSuccess ? SNKPT_OK : hip_fail(ctx, e, "hipDestroyExternalSemaphore");
}

}  // extern "C"
