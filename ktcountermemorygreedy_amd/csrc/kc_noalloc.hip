// kc_table_histogram_noalloc: the histogram of kc_table_histogram with its
// device bins in a file-scope device array instead of a per-call hipMalloc,
// for callers that cannot allocate (code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation).  No counterpart in the
// reference.
//
// The array, `scratch`, holds KC_NOALLOC_HIST_BINS 64-bit bins (512 KiB)
// and lives on every device the library runs on until the process ends.
// Host code never passes the array's name to a kernel or a copy (on the host
// it names a host-side stand-in, not device memory): on a device's first
// call its device address is looked up once with hipGetSymbolAddress, with
// that device current, and kept per device.  Every call on one device uses
// that device's array and leaves its contents behind; the bins are cleared
// by a fill kernel at the start of each call, and calls that overlap on one
// device (from several host threads or streams) race on the array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kc_internal.h"

namespace {

static __device__ __align__(256) unsigned long long scratch[KC_NOALLOC_HIST_BINS];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned long long *g_scratch[kMaxDevices];  // device address of `scratch`, per device

// The device address of this device's `scratch` (looked up on first use).
int scratch_of(int device, unsigned long long **out)
{
    if (device < 0 || device >= kMaxDevices) return KC_EINVAL;
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_scratch[device]) {
        KC_HIP(hipSetDevice(device));
        void *p = nullptr;
        KC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return KC_EDEVICE;
        g_scratch[device] = static_cast<unsigned long long *>(p);
    }
    *out = g_scratch[device];
    return KC_OK;
}

}  // namespace

extern "C" int kc_table_histogram_noalloc(kc_table *t, uint32_t min_freq, uint32_t max_freq, int64_t *h_bins)
{
    if (!t || !h_bins || max_freq < min_freq) return KC_EINVAL;
    if ((uint64_t)max_freq - min_freq + 1 > KC_NOALLOC_HIST_BINS) return KC_EINVAL;
    unsigned long long *bins = nullptr;
    int rc = scratch_of(t->ctx->device, &bins);
    if (rc != KC_OK) return rc;
    KC_HIP(hipSetDevice(t->ctx->device));
    return kc_histogram_into(t, min_freq, max_freq, h_bins, bins);
}
