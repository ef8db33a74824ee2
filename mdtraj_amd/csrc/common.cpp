// Host-side state of the hiprmsd library: thread-local error text and cached
// device properties.
#include <atomic>
#include <mutex>
#include <string>

#include "common.h"

namespace hr {

static thread_local std::string g_error;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
}

void clear_error() { g_error.clear(); }

static std::atomic<int> g_tuning[HR_TUNE_COUNT];

int tuning(int knob) { return knob >= 0 && knob < HR_TUNE_COUNT ? g_tuning[knob].load() : 0; }

static std::mutex g_dev_mutex;
static DeviceInfo g_devices[64];

const DeviceInfo* current_device() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        set_error("hipGetDevice failed or device index out of range");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_dev_mutex);
    DeviceInfo& d = g_devices[dev];
    if (d.device < 0) {
        int cu = 0;
        int lds = 0;
        int lds_cu = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
                hipSuccess ||
            hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                                  dev) != hipSuccess) {
            set_error("hipDeviceGetAttribute failed on device %d", dev);
            return nullptr;
        }
        d.cu_count = cu > 0 ? cu : 1;
        d.lds_per_block = lds > 0 ? (size_t)lds : 65536;
        d.lds_per_cu = lds_cu > 0 ? (size_t)lds_cu : d.lds_per_block;
        d.device = dev;
    }
    return &d;
}

}  // namespace hr

extern "C" {

int hr_version(void) { return 100; }

const char* hr_last_error(void) { return hr::g_error.c_str(); }

int hr_init(int device) {
    hr::clear_error();
    HR_CHECK_HIP(hipSetDevice(device));
    return hr::current_device() ? HR_OK : HR_ERR_HIP;
}

int hr_set_tuning(int knob, int value) {
    hr::clear_error();
    if (knob < 0 || knob >= HR_TUNE_COUNT) {
        hr::set_error("unknown tuning knob %d", knob);
        return HR_ERR_INVALID;
    }
    if (knob == HR_TUNE_FRAMES_WAVES_PER_SIMD && value != 0 && value != 3 && value != 4) {
        hr::set_error("HR_TUNE_FRAMES_WAVES_PER_SIMD must be 0, 3 or 4 (got %d)", value);
        return HR_ERR_INVALID;
    }
    if (knob == HR_TUNE_FRAMES_GRID && value != 0 && value != 1 && value != 2) {
        hr::set_error("HR_TUNE_FRAMES_GRID must be 0, 1 or 2 (got %d)", value);
        return HR_ERR_INVALID;
    }
    if (knob == HR_TUNE_FRAMES_ROW_LANES && value != 0 && value != 4 && value != 8 &&
        value != 16 && value != 32 && value != 64) {
        hr::set_error("HR_TUNE_FRAMES_ROW_LANES must be 0, 4, 8, 16, 32 or 64 (got %d)", value);
        return HR_ERR_INVALID;
    }
    if (knob == HR_TUNE_BATCH_SCHEDULE && (value < 0 || value > 4)) {
        hr::set_error("HR_TUNE_BATCH_SCHEDULE must be 0 .. 4 (got %d)", value);
        return HR_ERR_INVALID;
    }
    if (knob == HR_TUNE_PAIRS_TILE && (value < 0 || value > 5)) {
        hr::set_error("HR_TUNE_PAIRS_TILE must be 0 .. 5 (got %d)", value);
        return HR_ERR_INVALID;
    }
    if (knob == HR_TUNE_FRAMES_STATIC_PCT && (value < 0 || value > 100)) {
        hr::set_error("HR_TUNE_FRAMES_STATIC_PCT must be 0 .. 100 (got %d)", value);
        return HR_ERR_INVALID;
    }
    if (knob == HR_TUNE_FRAMES_CHUNK_KIB && value < 0) {
        hr::set_error("HR_TUNE_FRAMES_CHUNK_KIB must be 0 or positive (got %d)", value);
        return HR_ERR_INVALID;
    }
    if (knob == HR_TUNE_HOST_DIRECT_KIB && value < -1) {
        hr::set_error("HR_TUNE_HOST_DIRECT_KIB must be -1, 0 or positive (got %d)", value);
        return HR_ERR_INVALID;
    }
    if (knob == HR_TUNE_SUPERPOSE_CHUNK_MIB && value < 0) {
        hr::set_error("HR_TUNE_SUPERPOSE_CHUNK_MIB must be 0 or positive (got %d)", value);
        return HR_ERR_INVALID;
    }
    hr::g_tuning[knob].store(value);
    return HR_OK;
}

int hr_get_tuning(int knob) { return hr::tuning(knob); }

int hr_device_cu_count(void) {
    const hr::DeviceInfo* d = hr::current_device();
    return d ? d->cu_count : 0;
}

}  // extern "C"
