// noalloc.hip - qcs_sample_noalloc: qcs_sample with its per-shot device
// buffers in a fixed file-scope array instead of per-call allocations, for
// callers that cannot allocate (code under stream capture, serving loops,
// harnesses that forbid allocation calls).
//
// `scratch` holds kNoAllocShots shot values followed by kNoAllocShots
// outcomes (16 B per shot, 1 MiB). Each GPU the library runs on holds its own
// copy until the process ends. Host code never passes the array's name to a
// kernel or a copy: on the first call on a device, with that device current,
// its address on that device is read with hipGetSymbolAddress and kept per
// device, and only that address is used. Every call on one GPU shares that
// GPU's array: calls that overlap on one GPU (different registers sampled
// from different threads at once) race, so callers serialise them. Nothing
// in the array needs to start at zero: a call writes every entry it reads.
#include <mutex>

#include "common.h"

namespace qcs {

constexpr uint64_t kNoAllocShots = 65536;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocShots * (sizeof(double) + sizeof(uint64_t))];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device`, read once per device.
static int scratch_on(int device, void** out) {
    if (device < 0 || device >= kMaxDevices) return set_error(QCS_EDEVICE, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int prev = 0;
        if (hipGetDevice(&prev) != hipSuccess) return set_error(QCS_EDEVICE, "hipGetDevice failed");
        if (prev != device && hipSetDevice(device) != hipSuccess)
            return set_error(QCS_EDEVICE, "hipSetDevice(%d) failed", device);
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev != device) (void)hipSetDevice(prev);
        if (e != hipSuccess || !p)
            return set_error(QCS_EDEVICE, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return QCS_OK;
}

}  // namespace qcs

using namespace qcs;

extern "C" int qcs_sample_noalloc(qcs_register* r, int all, uint32_t first, uint32_t second, uint64_t nshots,
                                  const double* u, uint64_t* outcomes) {
    if (!r) return set_error(QCS_EINVAL, "null register");
    if (nshots > kNoAllocShots)
        return set_error(QCS_EINVAL, "qcs_sample_noalloc: %llu shots exceed the maximum of %llu",
                         (unsigned long long)nshots, (unsigned long long)kNoAllocShots);
    void* base = nullptr;
    if (const int e = scratch_on(r->device, &base)) return e;
    ShotDev dev;
    dev.in = static_cast<double*>(base);
    dev.out = reinterpret_cast<uint64_t*>(static_cast<char*>(base) + kNoAllocShots * sizeof(double));
    dev.cap = kNoAllocShots;
    return sample_impl(r, all, first, second, nshots, u, outcomes, &dev);
}
