// hs_noalloc.hip - read-back that allocates nothing (hs_read_strands_noalloc).
//
// hs_read_strands stages the groom's Strand-layout copy in device memory it
// allocates per call.  Callers that cannot allocate (tight serving loops,
// harnesses that forbid hipMalloc) use this variant instead: the staging lives
// in a file-scope device array, `scratch`, of kNoAllocBytes per GPU, held
// from module load until the process ends.  Its device address is looked up
// once per device (hs_create) and only that address is used from host code.
// Every call on one GPU shares that GPU's array: calls that overlap on one GPU
// (two host threads, two contexts) race and must be serialised by the caller.
// Nothing in the array needs clearing: the unpack kernel writes every byte of
// the S*N*48 it stages before the copy reads them.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "hs_internal.h"

namespace {

constexpr size_t kNoAllocBytes = HS_NOALLOC_STAGING_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

}  // namespace

namespace hsi {

// The device address of `scratch` on the current device (looked up once per
// device; the device must be current).  Null with the HIP error on failure.
hipError_t noalloc_scratch(int device, void** out)
{
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_ptr[device] = p;
    }
    *out = g_ptr[device];
    return hipSuccess;
}

}  // namespace hsi

extern "C" int hs_read_strands_noalloc(hs_ctx* c, uint32_t id, float* aos, hs_draw_indirect* ind)
{
    if (!c) return HS_ERR_INVALID;
    auto it = c->grooms.find(id);
    if (it == c->grooms.end()) {
        c->err = "hs_read_strands_noalloc: unknown groom";
        return HS_ERR_NOT_FOUND;
    }
    const hsi::Groom& g = it->second;
    const size_t bytes = (size_t)g.S * g.N * 48;
    if (aos && bytes > kNoAllocBytes) {
        c->err = "hs_read_strands_noalloc: the groom's " + std::to_string(bytes) +
                 " bytes exceed the fixed staging array of " + std::to_string(kNoAllocBytes) +
                 " bytes (use hs_read_strands)";
        return HS_ERR_INVALID;
    }
    if (ind) {
        // the count compute.comp:302 accumulates: one per strand processed
        ind->vertexCount = g.S;
        ind->instanceCount = 1;
        ind->firstVertex = 0;
        ind->firstInstance = 0;
    }
    if (!aos) return HS_OK;
    hipError_t e = hipSetDevice(c->device);
    void* stage = nullptr;
    if (e == hipSuccess) e = hsi::noalloc_scratch(c->device, &stage);
    if (e == hipSuccess) e = hsi::groom_unpack(g, static_cast<float*>(stage), c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(aos, stage, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        c->err = std::string("hs_read_strands_noalloc: ") + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? HS_ERR_OUT_OF_MEMORY : HS_ERR_HIP;
    }
    return HS_OK;
}
