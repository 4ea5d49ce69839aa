// noalloc.hip - the stable radix sort for callers that cannot allocate.
//
// mpm_dev_radix_sort takes a caller-allocated temp buffer.  This variant
// takes none: its keys and values ping-pong halves, the digit spine and the
// spine scan's tile sums live in `scratch`, a device array of fixed size that
// each GPU holds from the library's load until the process ends.  The entry
// point makes no hipMalloc, hipFree or other allocating call.
#include <mutex>

#include "internal.hpp"

namespace {

// Sorts of at most 2^22 keys; radix_sort_temp_bytes(2^22) is 32 MiB of
// ping-pong halves plus 1 MiB of digit spine and its scan sums.
constexpr uint64_t NOALLOC_MAX_KEYS = 1ull << 22;
constexpr size_t SCRATCH_BYTES = 34ull << 20;

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

// The array's address on each device: the host-side name `scratch` is not a
// device address, so every use on the host goes through this table.
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void *g_scratch[MAX_DEVICES] = {};

// The current device's scratch address, looked up once per device with that
// device current.
hipError_t scratch_for_current_device(void **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

}  // namespace

extern "C" {

uint64_t mpm_dev_radix_sort_noalloc_max_keys(void) { return NOALLOC_MAX_KEYS; }

int mpm_dev_radix_sort_noalloc(const uint32_t *keys_in, const uint32_t *values_in, uint32_t *keys_out,
                               uint32_t *values_out, uint64_t n, int32_t key_bits, void *stream) {
    if (n > 0 && (!keys_in || !keys_out || !values_out)) return MPM_ERR_INVALID_ARGUMENT;
    if (key_bits < 0 || key_bits > 32) return MPM_ERR_INVALID_ARGUMENT;
    if (keys_in == keys_out || (values_in && values_in == values_out)) return MPM_ERR_INVALID_ARGUMENT;
    if (n > NOALLOC_MAX_KEYS || svamd::radix_sort_temp_bytes(n) > SCRATCH_BYTES) return MPM_ERR_CAPACITY;
    if (n == 0) return MPM_OK;
    void *temp = nullptr;
    if (scratch_for_current_device(&temp) != hipSuccess) return MPM_ERR_HIP;
    // Nothing in the scratch has to start at zero: each pass writes the
    // digit spine and the ping-pong half it reads next before reading them.
    if (svamd::radix_sort(keys_in, values_in, keys_out, values_out, nullptr, nullptr, n, key_bits, temp,
                          static_cast<hipStream_t>(stream)) != hipSuccess)
        return MPM_ERR_HIP;
    return MPM_OK;
}

}  // extern "C"
