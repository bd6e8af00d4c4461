// Library-wide C entry points: version, error text, raw Philox draws.
#include <cstddef>
#include <cstring>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "philox.hpp"

namespace cm {

static thread_local char g_error[512] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof g_error, fmt, ap);
    va_end(ap);
}

const char* last_error() { return g_error; }

namespace {
// Pinned staging for small host->device uploads.  A ring of slots per device,
// each guarded by the event recorded after its copy, so an upload on one
// stream waits only for the copy that last used the same slot, not for work
// queued on other streams.  The ring grows (up to kMaxSlots) instead of
// waiting when its oldest slot's copy is still queued behind other work, so a
// caller that keeps several streams busy is not held to their pace.
struct Staging {
    void* host = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
    bool pending = false;
};
constexpr int kMaxDevices = 64;
constexpr size_t kMaxSlots = 256;
thread_local std::vector<Staging> g_ring[kMaxDevices];
thread_local size_t g_next[kMaxDevices];  // the oldest slot in use order

// A slot whose previous copy has completed: the oldest if it has, else a new
// slot placed just before it (it becomes the newest), else the oldest after
// waiting for its copy.
hipError_t take_slot(int dev, Staging** out)
{
    std::vector<Staging>& ring = g_ring[dev];
    size_t& next = g_next[dev];
    if (!ring.empty()) {
        Staging& old = ring[next % ring.size()];
        hipError_t q = hipSuccess;
        if (old.pending) {
            q = hipEventQuery(old.done);
            if (q != hipSuccess && q != hipErrorNotReady) return q;
        }
        if (!old.pending || q == hipSuccess || ring.size() >= kMaxSlots) {
            if (old.pending && q != hipSuccess) {
                const hipError_t e = hipEventSynchronize(old.done);
                if (e != hipSuccess) return e;
            }
            old.pending = false;
            *out = &old;
            next = (next % ring.size()) + 1;
            return hipSuccess;
        }
    }
    const size_t pos = ring.empty() ? 0 : next % ring.size();
    ring.insert(ring.begin() + (std::ptrdiff_t)pos, Staging{});
    next = pos + 1;
    *out = &ring[pos];
    return hipSuccess;
}
}  // namespace

hipError_t upload_async(void* dst, const void* src, size_t bytes, hipStream_t stream)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    Staging* slot = nullptr;
    e = take_slot(dev, &slot);
    if (e != hipSuccess) return e;
    Staging& st = *slot;
    if (st.done == nullptr) {
        e = hipEventCreateWithFlags(&st.done, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    if (st.cap < bytes) {  // grown in powers of two from 64 KiB, kept for the process
        if (st.host) (void)hipHostFree(st.host);
        st.host = nullptr;
        st.cap = 0;
        size_t cap = 64 * 1024;
        while (cap < bytes) cap *= 2;
        e = hipHostMalloc(&st.host, cap, hipHostMallocDefault);
        if (e != hipSuccess) return e;
        st.cap = cap;
    }
    std::memcpy(st.host, src, bytes);
    e = hipMemcpyAsync(dst, st.host, bytes, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return e;
    e = hipEventRecord(st.done, stream);
    if (e != hipSuccess) return e;
    st.pending = true;
    return hipSuccess;
}

__global__ void philox_fill_kernel(uint32_t* __restrict__ out, int64_t n_blocks, uint32_t k0,
                                   uint32_t k1, uint32_t x0, uint32_t y, uint32_t z, uint32_t w)
{
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < n_blocks;
         b += (int64_t)gridDim.x * blockDim.x) {
        const u32x4 r = philox4x32_10(u32x4{x0 + (uint32_t)b, y, z, w}, k0, k1);
        out[4 * b + 0] = r.x;
        out[4 * b + 1] = r.y;
        out[4 * b + 2] = r.z;
        out[4 * b + 3] = r.w;
    }
}

// One workgroup per span: copies spans[3 s + 2] bytes (a multiple of 4) from
// spans[3 s + 1] to spans[3 s], both 4-byte aligned, in 4-byte words.
__global__ __launch_bounds__(256) void copy_spans_kernel(const uint64_t* __restrict__ spans, int n)
{
    for (int sp = blockIdx.x; sp < n; sp += gridDim.x) {
        uint32_t* dst = (uint32_t*)spans[3 * sp];
        const uint32_t* src = (const uint32_t*)spans[3 * sp + 1];
        const uint64_t words = spans[3 * sp + 2] >> 2;
        for (uint64_t w = threadIdx.x; w < words; w += blockDim.x) dst[w] = src[w];
    }
}

}  // namespace cm

using namespace cm;

extern "C" {

int cm_copy_spans(const void* spans, int32_t n, void* stream)
{
    CM_REQUIRE(n >= 0, "n must not be negative");
    if (n == 0) return CM_OK;
    const int blocks = std::min(n, 65535);
    hipLaunchKernelGGL(copy_spans_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       (const uint64_t*)spans, n);
    CM_HIP_CHECK(hipGetLastError());
    return CM_OK;
}

int32_t cm_version(void) { return 1; }

const char* cm_last_error(void) { return last_error(); }

int cm_philox_fill(uint32_t* out, int64_t n_blocks, uint32_t k0, uint32_t k1, uint32_t x0,
                   uint32_t y, uint32_t z, uint32_t w, void* stream)
{
    CM_REQUIRE(n_blocks >= 0, "n_blocks must not be negative");
    if (n_blocks == 0) return CM_OK;
    const int blocks = (int)std::min<int64_t>((n_blocks + 255) / 256, 4096);
    hipLaunchKernelGGL(philox_fill_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, out,
                       n_blocks, k0, k1, x0, y, z, w);
    CM_HIP_CHECK(hipGetLastError());
    return CM_OK;
}

void cm_philox_host(const uint32_t* ctr, const uint32_t* key, uint32_t* out)
{
    const u32x4 r = philox4x32_10(u32x4{ctr[0], ctr[1], ctr[2], ctr[3]}, key[0], key[1]);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
}

int cm_device_info(int32_t* cu_count, int64_t* lds_per_block, int32_t* arch_major,
                   int32_t* arch_minor)
{
    int dev = 0;
    CM_HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    CM_HIP_CHECK(hipGetDeviceProperties(&p, dev));
    *cu_count = p.multiProcessorCount;
    *lds_per_block = (int64_t)p.sharedMemPerBlock;
    *arch_major = p.major;
    *arch_minor = p.minor;
    return CM_OK;
}

}  // extern "C"
