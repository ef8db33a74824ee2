// Host-memory side of the hiprmsd C ABI: device and pinned allocations,
// copies, streams, and the host-pointer one-vs-all entry point.
//
// hr_rmsd_one_vs_all_host takes host pointers, as the reference's C ABI does
// (rmsd/include/theobald_rmsd.h:16-18 reads host frames; the Cython driver
// rmsd/_rmsd.pyx:186-224 gathers the selection on the host first).  It keeps
// a per-device context of cached buffers and a worker pool:
//   * the reference selection and every chunk of frames are gathered on the
//     host straight into pinned staging slots, so only the selected atoms
//     cross PCIe (75 % of the bytes for a 75 % selection) and the caller's
//     pageable array is read once, by several threads;
//   * a slot's host-to-device copy and the per-frame kernel over that chunk
//     are queued on the context's stream while the next chunk is gathered
//     into the next slot (three slots: the gather runs at most two chunks
//     ahead of the copies);
//   * the RMSDs come back with one device-to-host copy into pinned memory.
// Small inputs (HR_TUNE_HOST_DIRECT_KIB, 1 MiB by default) skip both copies:
// the kernels read the gathered frames straight from the pinned slot and
// write the RMSDs straight to pinned memory, over PCIe, which saves two
// round trips of copy latency on a call whose work is a few microseconds.
// The call returns when `out` holds the result.  Calls on one device are
// serialised by the context's mutex.
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "common.h"

namespace hr {
namespace {

constexpr int kSlots = 3;
constexpr size_t kChunkBytes = 64u << 20;     // gathered bytes per chunk
constexpr size_t kParallelBytes = 1u << 20;   // below this, gather on the calling thread
constexpr size_t kDirectBytes = 1u << 20;     // default HR_TUNE_HOST_DIRECT_KIB

// Fixed set of workers that run one function over [0, n) indices at a time.
class WorkerPool {
public:
    void start(int n) {
        if (!threads_.empty()) return;
        for (int i = 0; i < n; ++i) threads_.emplace_back([this, i] { loop(i); });
        n_ = n;
    }
    ~WorkerPool() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : threads_) t.join();
    }
    int size() const { return n_; }
    // Runs f(i) for i in [0, size()) on the workers; returns when all are done.
    void run(const std::function<void(int)>& f) {
        std::unique_lock<std::mutex> g(m_);
        job_ = &f;
        remaining_ = n_;
        ++generation_;
        cv_.notify_all();
        done_cv_.wait(g, [this] { return remaining_ == 0; });
        job_ = nullptr;
    }

private:
    void loop(int i) {
        long long seen = 0;
        for (;;) {
            const std::function<void(int)>* job;
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return stop_ || generation_ != seen; });
                if (stop_) return;
                seen = generation_;
                job = job_;
            }
            (*job)(i);
            {
                std::lock_guard<std::mutex> g(m_);
                if (--remaining_ == 0) done_cv_.notify_all();
            }
        }
    }
    std::vector<std::thread> threads_;
    std::mutex m_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(int)>* job_ = nullptr;
    long long generation_ = 0;
    int remaining_ = 0;
    int n_ = 0;
    bool stop_ = false;
};

// A device or pinned buffer that only grows.
struct Buffer {
    void* p = nullptr;
    size_t bytes = 0;
    bool pinned = false;
    void* dev = nullptr;  // pinned: the device-visible address, looked up once
};

int grow(Buffer* b, size_t need, bool pinned) {
    if (b->p && b->bytes >= need) return HR_OK;
    if (b->p) {
        hipError_t e = b->pinned ? hipHostFree(b->p) : hipFree(b->p);
        b->p = nullptr;
        b->bytes = 0;
        b->dev = nullptr;
        if (e != hipSuccess) {
            set_error("freeing a staging buffer failed: %s", hipGetErrorString(e));
            return HR_ERR_HIP;
        }
    }
    need = (need + 4095) & ~(size_t)4095;
    hipError_t e = pinned ? hipHostMalloc(&b->p, need, hipHostMallocDefault) : hipMalloc(&b->p, need);
    if (e != hipSuccess) {
        b->p = nullptr;
        set_error("%s of %zu bytes failed: %s", pinned ? "hipHostMalloc" : "hipMalloc", need,
                  hipGetErrorString(e));
        return HR_ERR_HIP;
    }
    b->bytes = need;
    b->pinned = pinned;
    b->dev = nullptr;
    if (pinned) {
        e = hipHostGetDevicePointer(&b->dev, b->p, 0);
        if (e != hipSuccess) {
            set_error("hipHostGetDevicePointer failed: %s", hipGetErrorString(e));
            return HR_ERR_HIP;
        }
    }
    return HR_OK;
}

struct HostContext {
    std::mutex m;
    bool ready = false;
    hipStream_t stream = nullptr;
    hipEvent_t slot_done[kSlots] = {};
    bool slot_busy[kSlots] = {};
    Buffer pinned_slot[kSlots], dev_slot[kSlots];
    Buffer pinned_ref, dev_ref, dev_ws, dev_out, pinned_out;
    WorkerPool workers;
};

HostContext g_ctx[64];

int context_for_current_device(HostContext** out) {
    int dev = 0;
    HR_CHECK_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) {
        set_error("device index %d out of range", dev);
        return HR_ERR_HIP;
    }
    *out = &g_ctx[dev];
    return HR_OK;
}

int init_context(HostContext* c) {
    if (c->ready) return HR_OK;
    HR_CHECK_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (int s = 0; s < kSlots; ++s)
        HR_CHECK_HIP(hipEventCreateWithFlags(&c->slot_done[s], hipEventDisableTiming));
    unsigned hw = std::thread::hardware_concurrency();
    int n = hw ? (int)hw : 4;
    if (n > 16) n = 16;
    c->workers.start(n);
    c->ready = true;
    return HR_OK;
}

// Copies frames [f0, f1) of the selection into dst (frame-major, n_sel atoms
// of 3 floats each), over the workers when the chunk is large.
void gather_frames(HostContext* c, const float* target, int64_t n_atoms_total,
                   const int32_t* idx, int32_t n_sel, int64_t f0, int64_t f1, float* dst) {
    const size_t frame_out = (size_t)n_sel * 3;
    auto range = [&](int64_t a, int64_t b) {
        if (!idx) {
            // whole frames: the range is one contiguous block
            std::memcpy(dst + (size_t)(a - f0) * frame_out, target + (size_t)a * frame_out,
                        (size_t)(b - a) * frame_out * sizeof(float));
            return;
        }
        for (int64_t f = a; f < b; ++f) {
            const float* src = target + (size_t)f * n_atoms_total * 3;
            float* o = dst + (size_t)(f - f0) * frame_out;
            for (int32_t i = 0; i < n_sel; ++i) {
                const float* a3 = src + (size_t)idx[i] * 3;
                o[3 * i] = a3[0];
                o[3 * i + 1] = a3[1];
                o[3 * i + 2] = a3[2];
            }
        }
    };
    const int64_t n = f1 - f0;
    if ((size_t)n * frame_out * sizeof(float) < kParallelBytes || c->workers.size() < 2) {
        range(f0, f1);
        return;
    }
    const int w = c->workers.size();
    c->workers.run([&](int i) {
        const int64_t a = f0 + n * i / w, b = f0 + n * (i + 1) / w;
        range(a, b);
    });
}

// Gathers the reference selection into pinned_ref (axis-agnostic atom-major).
void gather_reference(HostContext* c, const float* ref_frame, const int32_t* ref_idx,
                      int32_t n_sel) {
    float* pr = static_cast<float*>(c->pinned_ref.p);
    for (int32_t i = 0; i < n_sel; ++i) {
        const float* a3 = ref_frame + (size_t)(ref_idx ? ref_idx[i] : i) * 3;
        pr[3 * i] = a3[0];
        pr[3 * i + 1] = a3[1];
        pr[3 * i + 2] = a3[2];
    }
}

// The whole selection in one pinned slot, read by the kernels in place; the
// RMSDs written to pinned memory by the per-frame kernel.
int one_vs_all_direct(HostContext* c, const float* target, int64_t n_frames,
                      int32_t n_atoms_total, const int32_t* atom_idx, int32_t n_sel,
                      const float* ref_frame, const int32_t* ref_idx, int64_t self_frame,
                      float* out) {
    int rc;
    const size_t frame_bytes = (size_t)n_sel * 3 * sizeof(float);
    if ((rc = grow(&c->pinned_ref, frame_bytes, true))) return rc;
    if ((rc = grow(&c->dev_ws, hr_rmsd_workspace_bytes(n_sel), false))) return rc;
    if ((rc = grow(&c->pinned_out, (size_t)n_frames * sizeof(float), true))) return rc;
    if ((rc = grow(&c->pinned_slot[0], (size_t)n_frames * frame_bytes, true))) return rc;
    if (c->slot_busy[0]) {
        HR_CHECK_HIP(hipEventSynchronize(c->slot_done[0]));
        c->slot_busy[0] = false;
    }
    gather_reference(c, ref_frame, ref_idx, n_sel);
    gather_frames(c, target, n_atoms_total, atom_idx, n_sel, 0, n_frames,
                  static_cast<float*>(c->pinned_slot[0].p));
    const float* d_ref = static_cast<const float*>(c->pinned_ref.dev);
    const float* d_frames = static_cast<const float*>(c->pinned_slot[0].dev);
    float* d_out = static_cast<float*>(c->pinned_out.dev);
    // one launch when the reference fits the per-frame kernel's LDS (the
    // kernel prepares it with the prep kernel's arithmetic: the same bits)
    bool done = false;
    if ((rc = one_vs_all_fused_ref(c->stream, d_frames, n_frames, n_sel, d_ref, self_frame, d_out,
                                   &done)))
        return rc;
    if (!done) {
        if ((rc = hr_rmsd_prepare_reference(d_ref, nullptr, n_sel, c->dev_ws.p, c->stream)))
            return rc;
        if ((rc = hr_rmsd_one_vs_all_prepared(d_frames, n_frames, n_sel, nullptr, n_sel,
                                              self_frame, d_out, c->dev_ws.p, c->stream)))
            return rc;
    }
    HR_CHECK_HIP(hipStreamSynchronize(c->stream));
    std::memcpy(out, c->pinned_out.p, (size_t)n_frames * sizeof(float));
    return HR_OK;
}

int one_vs_all_host(HostContext* c, const float* target, int64_t n_frames, int32_t n_atoms_total,
                    const int32_t* atom_idx, int32_t n_sel, const float* ref_frame,
                    const int32_t* ref_idx, int64_t self_frame, float* out) {
    int rc;
    const size_t frame_bytes = (size_t)n_sel * 3 * sizeof(float);
    int64_t chunk = (int64_t)(kChunkBytes / frame_bytes);
    if (chunk < 1) chunk = 1;
    if (chunk > n_frames) chunk = n_frames;
    const size_t slot_bytes = (size_t)chunk * frame_bytes;
    if ((rc = grow(&c->pinned_ref, frame_bytes, true))) return rc;
    if ((rc = grow(&c->dev_ref, frame_bytes, false))) return rc;
    if ((rc = grow(&c->dev_ws, hr_rmsd_workspace_bytes(n_sel), false))) return rc;
    if ((rc = grow(&c->dev_out, (size_t)n_frames * sizeof(float), false))) return rc;
    if ((rc = grow(&c->pinned_out, (size_t)n_frames * sizeof(float), true))) return rc;
    const int direct_kib = tuning(HR_TUNE_HOST_DIRECT_KIB);
    const size_t direct_bytes = direct_kib == 0 ? kDirectBytes
                                : direct_kib < 0 ? 0 : (size_t)direct_kib << 10;
    if ((size_t)n_frames * frame_bytes <= direct_bytes)
        return one_vs_all_direct(c, target, n_frames, n_atoms_total, atom_idx, n_sel, ref_frame,
                                 ref_idx, self_frame, out);
    const int n_chunks = (int)((n_frames + chunk - 1) / chunk);
    const int used = n_chunks < kSlots ? n_chunks : kSlots;
    for (int s = 0; s < used; ++s) {
        if ((rc = grow(&c->pinned_slot[s], slot_bytes, true))) return rc;
        if ((rc = grow(&c->dev_slot[s], slot_bytes, false))) return rc;
    }
    gather_reference(c, ref_frame, ref_idx, n_sel);
    HR_CHECK_HIP(hipMemcpyAsync(c->dev_ref.p, c->pinned_ref.p, frame_bytes, hipMemcpyHostToDevice,
                                c->stream));
    if ((rc = hr_rmsd_prepare_reference(static_cast<const float*>(c->dev_ref.p), nullptr, n_sel,
                                        c->dev_ws.p, c->stream)))
        return rc;
    float* d_out = static_cast<float*>(c->dev_out.p);
    for (int k = 0; k < n_chunks; ++k) {
        const int s = k % kSlots;
        const int64_t f0 = (int64_t)k * chunk;
        const int64_t f1 = f0 + chunk < n_frames ? f0 + chunk : n_frames;
        if (c->slot_busy[s]) {
            HR_CHECK_HIP(hipEventSynchronize(c->slot_done[s]));
            c->slot_busy[s] = false;
        }
        float* ps = static_cast<float*>(c->pinned_slot[s].p);
        gather_frames(c, target, n_atoms_total, atom_idx, n_sel, f0, f1, ps);
        const size_t bytes = (size_t)(f1 - f0) * frame_bytes;
        HR_CHECK_HIP(hipMemcpyAsync(c->dev_slot[s].p, ps, bytes, hipMemcpyHostToDevice, c->stream));
        HR_CHECK_HIP(hipEventRecord(c->slot_done[s], c->stream));
        c->slot_busy[s] = true;
        const int64_t rel = (self_frame >= f0 && self_frame < f1) ? self_frame - f0 : -1;
        if ((rc = hr_rmsd_one_vs_all_prepared(static_cast<const float*>(c->dev_slot[s].p), f1 - f0,
                                              n_sel, nullptr, n_sel, rel, d_out + f0, c->dev_ws.p,
                                              c->stream)))
            return rc;
    }
    HR_CHECK_HIP(hipMemcpyAsync(c->pinned_out.p, d_out, (size_t)n_frames * sizeof(float),
                                hipMemcpyDeviceToHost, c->stream));
    HR_CHECK_HIP(hipStreamSynchronize(c->stream));
    for (int s = 0; s < kSlots; ++s) c->slot_busy[s] = false;
    std::memcpy(out, c->pinned_out.p, (size_t)n_frames * sizeof(float));
    return HR_OK;
}

}  // namespace

// Whether p lies in host memory the runtime knows as pinned; pageable
// memory, for which copies wait until they are done.
bool is_pinned(const void* p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

}  // namespace hr

extern "C" {

int hr_device_count(int* count) {
    hr::clear_error();
    HR_REQUIRE(count != nullptr, "count is NULL");
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e == hipErrorNoDevice) {
        *count = 0;
        return HR_OK;
    }
    HR_CHECK_HIP(e);
    return HR_OK;
}

int hr_alloc(void** ptr, size_t nbytes) {
    hr::clear_error();
    HR_REQUIRE(ptr != nullptr, "ptr is NULL");
    *ptr = nullptr;
    if (nbytes == 0) return HR_OK;
    HR_CHECK_HIP(hipMalloc(ptr, nbytes));
    return HR_OK;
}

int hr_free(void* ptr) {
    hr::clear_error();
    if (ptr) HR_CHECK_HIP(hipFree(ptr));
    return HR_OK;
}

int hr_alloc_host(void** ptr, size_t nbytes) {
    hr::clear_error();
    HR_REQUIRE(ptr != nullptr, "ptr is NULL");
    *ptr = nullptr;
    if (nbytes == 0) return HR_OK;
    HR_CHECK_HIP(hipHostMalloc(ptr, nbytes, hipHostMallocDefault));
    return HR_OK;
}

int hr_free_host(void* ptr) {
    hr::clear_error();
    if (ptr) HR_CHECK_HIP(hipHostFree(ptr));
    return HR_OK;
}

int hr_memcpy_h2d(void* dst, const void* src, size_t nbytes, void* stream) {
    hr::clear_error();
    if (nbytes == 0) return HR_OK;
    HR_REQUIRE(dst && src, "NULL pointer argument");
    HR_CHECK_HIP(hipMemcpyAsync(dst, src, nbytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    if (!hr::is_pinned(src)) HR_CHECK_HIP(hipStreamSynchronize((hipStream_t)stream));
    return HR_OK;
}

int hr_memcpy_d2h(void* dst, const void* src, size_t nbytes, void* stream) {
    hr::clear_error();
    if (nbytes == 0) return HR_OK;
    HR_REQUIRE(dst && src, "NULL pointer argument");
    HR_CHECK_HIP(hipMemcpyAsync(dst, src, nbytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    if (!hr::is_pinned(dst)) HR_CHECK_HIP(hipStreamSynchronize((hipStream_t)stream));
    return HR_OK;
}

int hr_memcpy_d2d(void* dst, const void* src, size_t nbytes, void* stream) {
    hr::clear_error();
    if (nbytes == 0) return HR_OK;
    HR_REQUIRE(dst && src, "NULL pointer argument");
    HR_CHECK_HIP(hipMemcpyAsync(dst, src, nbytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return HR_OK;
}

int hr_memset(void* dst, int value, size_t nbytes, void* stream) {
    hr::clear_error();
    if (nbytes == 0) return HR_OK;
    HR_REQUIRE(dst != nullptr, "dst is NULL");
    HR_CHECK_HIP(hipMemsetAsync(dst, value, nbytes, (hipStream_t)stream));
    return HR_OK;
}

int hr_stream_create(void** stream) {
    hr::clear_error();
    HR_REQUIRE(stream != nullptr, "stream is NULL");
    hipStream_t s = nullptr;
    HR_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return HR_OK;
}

int hr_stream_destroy(void* stream) {
    hr::clear_error();
    if (stream) HR_CHECK_HIP(hipStreamDestroy((hipStream_t)stream));
    return HR_OK;
}

int hr_stream_synchronize(void* stream) {
    hr::clear_error();
    HR_CHECK_HIP(hipStreamSynchronize((hipStream_t)stream));
    return HR_OK;
}

int hr_rmsd_one_vs_all_host(const float* target, int64_t n_frames, int32_t n_atoms_total,
                            const int32_t* atom_idx, int32_t n_sel, const float* ref_frame,
                            int32_t n_ref_atoms_total, const int32_t* ref_idx,
                            int64_t self_frame, float* out) {
    hr::clear_error();
    HR_REQUIRE(n_frames >= 0 && n_atoms_total > 0 && n_sel > 0 && n_ref_atoms_total > 0,
               "invalid sizes n_frames=%lld n_atoms=%d n_sel=%d n_ref_atoms=%d",
               (long long)n_frames, n_atoms_total, n_sel, n_ref_atoms_total);
    HR_REQUIRE(atom_idx != nullptr || n_sel == n_atoms_total,
               "n_sel (%d) must equal n_atoms_total (%d) without atom_idx", n_sel, n_atoms_total);
    HR_REQUIRE(ref_idx != nullptr || n_sel <= n_ref_atoms_total,
               "n_sel (%d) exceeds the reference's %d atoms without ref_idx", n_sel,
               n_ref_atoms_total);
    if (n_frames == 0) return HR_OK;
    HR_REQUIRE(target && ref_frame && out, "NULL pointer argument");
    for (int32_t i = 0; i < (atom_idx ? n_sel : 0); ++i)
        HR_REQUIRE(atom_idx[i] >= 0 && atom_idx[i] < n_atoms_total,
                   "atom_idx[%d] = %d outside [0, %d)", i, atom_idx[i], n_atoms_total);
    for (int32_t i = 0; i < (ref_idx ? n_sel : 0); ++i)
        HR_REQUIRE(ref_idx[i] >= 0 && ref_idx[i] < n_ref_atoms_total,
                   "ref_idx[%d] = %d outside [0, %d)", i, ref_idx[i], n_ref_atoms_total);
    hr::HostContext* c = nullptr;
    int rc = hr::context_for_current_device(&c);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(c->m);
    if ((rc = hr::init_context(c))) return rc;
    rc = hr::one_vs_all_host(c, target, n_frames, n_atoms_total, atom_idx, n_sel, ref_frame,
                             ref_idx, self_frame, out);
    if (rc) {
        // leave nothing in flight that reads the staging buffers
        (void)hipStreamSynchronize(c->stream);
        for (int s = 0; s < hr::kSlots; ++s) c->slot_busy[s] = false;
    }
    return rc;
}

}  // extern "C"
