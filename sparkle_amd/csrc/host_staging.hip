// Closest hit of rays held in pageable host memory: a pipeline over pinned staging buffers.
//
// The caller's arrays are ordinary pageable memory, which the GPU cannot read. hipMemcpy from and
// to pageable memory measured 28 GB/s up and 17 GB/s down on MI355X against 57 GB/s each way from
// pinned memory (97 GB/s with both directions in flight), and a multi-threaded host memcpy into
// pinned memory runs at 75-87 GB/s on 4-16 threads. Here the batch is cut into
// chunks that flow through a ring of pinned buffers: the host copies chunk c into pinned memory
// with several threads while the DMA engine uploads chunk c-1, the GPU traces the chunk before that
// and the hits of earlier chunks come back on a second copy engine. The upload is the bound (32 B
// per ray in, 8 B per ray out, both directions of the link in flight at once).
//
// Every chunk is a spk_trace_closest_async call on the context stream over its slice of the
// device ray and hit buffers, so the hits are those of one call over the whole batch (a ray's
// closest hit does not depend on the other rays of its call).
#include <hip/hip_runtime.h>
#include <sched.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/spk_pt.h"
#include "spk_context.h"

namespace spk {

namespace {

// Large memcpy split over a few persistent threads and the calling thread.
class CopyPool {
  public:
    explicit CopyPool(unsigned workers) {
        for (unsigned k = 0; k < workers; ++k) threads_.emplace_back([this, k] { worker(k + 1u); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (std::thread &t : threads_) t.join();
    }
    void copy(void *dst, const void *src, size_t bytes) {
        if (threads_.empty() || bytes < (1u << 20)) {
            std::memcpy(dst, src, bytes);
            return;
        }
        {
            std::lock_guard<std::mutex> g(m_);
            dst_ = static_cast<char *>(dst);
            src_ = static_cast<const char *>(src);
            bytes_ = bytes;
            pending_ = (unsigned)threads_.size();
            ++gen_;
        }
        cv_.notify_all();
        part(0u);
        std::unique_lock<std::mutex> l(m_);
        done_.wait(l, [this] { return pending_ == 0u; });
    }

  private:
    // Part k of parts() pieces, cut at 4 KiB multiples.
    void part(unsigned k) {
        const size_t parts = threads_.size() + 1u;
        const size_t piece = ((bytes_ + parts - 1u) / parts + 4095u) & ~size_t(4095u);
        const size_t lo = std::min(bytes_, piece * k), hi = std::min(bytes_, lo + piece);
        if (hi > lo) std::memcpy(dst_ + lo, src_ + lo, hi - lo);
    }
    void worker(unsigned k) {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> l(m_);
                cv_.wait(l, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            part(k);
            std::lock_guard<std::mutex> g(m_);
            if (--pending_ == 0u) done_.notify_one();
        }
    }

    std::vector<std::thread> threads_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    uint64_t gen_ = 0;
    unsigned pending_ = 0;
    bool stop_ = false;
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t bytes_ = 0;
};

unsigned host_copy_threads() {
    cpu_set_t set;
    CPU_ZERO(&set);
    int cpus = sched_getaffinity(0, sizeof(set), &set) == 0 ? CPU_COUNT(&set) : 1;
    return (unsigned)std::clamp(cpus, 1, 8);
}

}  // namespace

// Pinned ring buffers, copy streams and events of spk_trace_closest_host (one per context).
struct HostStaging {
    static constexpr int kSlots = 3;
    void *pin_in[kSlots] = {}, *pin_out[kSlots] = {};
    uint64_t slot_rays = 0;  // rays one slot holds
    hipStream_t h2d = nullptr, d2h = nullptr;
    hipEvent_t start = nullptr, in_done[kSlots] = {}, traced[kSlots] = {}, out_done[kSlots] = {};
    CopyPool *pool = nullptr;

    void free_pinned() {
        for (int s = 0; s < kSlots; ++s) {
            if (pin_in[s]) (void)hipHostFree(pin_in[s]);
            if (pin_out[s]) (void)hipHostFree(pin_out[s]);
            pin_in[s] = pin_out[s] = nullptr;
        }
        slot_rays = 0;
    }
    ~HostStaging() {
        if (h2d) (void)hipStreamSynchronize(h2d);
        if (d2h) (void)hipStreamSynchronize(d2h);
        free_pinned();
        for (hipEvent_t e : {start}) if (e) (void)hipEventDestroy(e);
        for (int s = 0; s < kSlots; ++s)
            for (hipEvent_t e : {in_done[s], traced[s], out_done[s]}) if (e) (void)hipEventDestroy(e);
        if (h2d) (void)hipStreamDestroy(h2d);
        if (d2h) (void)hipStreamDestroy(d2h);
        delete pool;
    }
    hipError_t init() {
        hipError_t e;
        // default-priority streams (the context stream's priority)
        if ((e = hipStreamCreateWithFlags(&h2d, hipStreamNonBlocking)) != hipSuccess) return e;
        if ((e = hipStreamCreateWithFlags(&d2h, hipStreamNonBlocking)) != hipSuccess) return e;
        if ((e = hipEventCreateWithFlags(&start, hipEventDisableTiming)) != hipSuccess) return e;
        for (int s = 0; s < kSlots; ++s)
            for (hipEvent_t *ev : {&in_done[s], &traced[s], &out_done[s]})
                if ((e = hipEventCreateWithFlags(ev, hipEventDisableTiming)) != hipSuccess) return e;
        pool = new CopyPool(host_copy_threads() - 1u);
        return hipSuccess;
    }
    hipError_t ensure(uint64_t rays) {
        if (rays <= slot_rays) return hipSuccess;
        free_pinned();
        hipError_t e;
        for (int s = 0; s < kSlots; ++s) {
            if ((e = hipHostMalloc(&pin_in[s], rays * sizeof(spk_ray), hipHostMallocDefault)) != hipSuccess ||
                (e = hipHostMalloc(&pin_out[s], rays * sizeof(spk_hit), hipHostMallocDefault)) != hipSuccess) {
                free_pinned();
                return e;
            }
        }
        slot_rays = rays;
        return hipSuccess;
    }
};

void destroy_host_staging(HostStaging *hs) { delete hs; }

namespace {

// Rays per chunk: a sixteenth of the batch in multiples of 65,536, at least 524,288 and at most
// 1,048,576 (the ring then holds 120 MiB of pinned memory). Each chunk is a traversal call of its own
// with a fixed latency (measured on MI355X with the coherence sort: 1,048,576 rays in 16 chunks took
// 5.2 ms, in two 1.5 ms); 8,388,608 rays in 16 chunks took 6.1-7.0 ms, bound by the upload.
uint64_t host_chunk_rays(uint64_t n) {
    constexpr uint64_t kStep = 1u << 16, kMin = 1u << 19, kMax = 1u << 20;
    const uint64_t c = ((n + 15u) / 16u + kStep - 1u) / kStep * kStep;
    return std::clamp(c, kMin, kMax);
}

}  // namespace

int trace_closest_host_staged(spk_context *ctx, const spk_ray *rays, uint64_t n, spk_hit *hits) {
    if (!ctx->scene.valid) return set_error(ctx, SPK_ERR_NO_SCENE, "no scene uploaded");
    (void)hipSetDevice(ctx->device);
    hipError_t e;
    if (!ctx->staging) {
        ctx->staging = new HostStaging();
        if ((e = ctx->staging->init()) != hipSuccess) {
            delete ctx->staging;
            ctx->staging = nullptr;
            return hip_fail(ctx, e, "host staging set-up");
        }
    }
    HostStaging &st = *ctx->staging;
    const uint64_t chunk = std::min(n, host_chunk_rays(n));
    const uint64_t chunks = (n + chunk - 1u) / chunk;
    if ((e = st.ensure(chunk)) != hipSuccess) return hip_fail(ctx, e, "pinned staging buffers");
    if ((e = ctx->tmp_rays.ensure(n * sizeof(spk_ray))) != hipSuccess ||
        (e = ctx->tmp_hits.ensure(n * sizeof(spk_hit))) != hipSuccess)
        return hip_fail(ctx, e, "ray buffers");
    spk_ray *dev_rays = ctx->tmp_rays.as<spk_ray>();
    spk_hit *dev_hits = ctx->tmp_hits.as<spk_hit>();
    auto drain_all = [&] {
        (void)hipStreamSynchronize(st.h2d);
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamSynchronize(st.d2h);
    };
    // work already queued on the context stream may still use the device ray and hit buffers
    if ((e = hipEventRecord(st.start, ctx->stream)) != hipSuccess ||
        (e = hipStreamWaitEvent(st.h2d, st.start, 0)) != hipSuccess)
        return hip_fail(ctx, e, "host staging order");
    auto slice = [&](uint64_t c) { return std::min(chunk, n - c * chunk); };
    // hits of chunk c (its D2H enqueued): wait for them and copy them out of the ring
    auto drain = [&](uint64_t c) -> hipError_t {
        const int s = (int)(c % HostStaging::kSlots);
        hipError_t r = hipEventSynchronize(st.out_done[s]);
        if (r != hipSuccess) return r;
        st.pool->copy(hits + c * chunk, st.pin_out[s], slice(c) * sizeof(spk_hit));
        return hipSuccess;
    };
    uint64_t drained = 0;
    for (uint64_t c = 0; c < chunks; ++c) {
        const int s = (int)(c % HostStaging::kSlots);
        const uint64_t off = c * chunk, m = slice(c);
        // the slot's previous upload (chunk c - kSlots) has finished: its hits were drained
        if (c >= (uint64_t)HostStaging::kSlots && (e = hipEventSynchronize(st.in_done[s])) != hipSuccess) break;
        st.pool->copy(st.pin_in[s], rays + off, m * sizeof(spk_ray));
        if ((e = hipMemcpyAsync(dev_rays + off, st.pin_in[s], m * sizeof(spk_ray), hipMemcpyHostToDevice, st.h2d)) !=
                hipSuccess ||
            (e = hipEventRecord(st.in_done[s], st.h2d)) != hipSuccess ||
            (e = hipStreamWaitEvent(ctx->stream, st.in_done[s], 0)) != hipSuccess)
            break;
        const int rc = spk_trace_closest_async(ctx, dev_rays + off, m, dev_hits + off);
        if (rc != SPK_OK) {
            drain_all();
            return rc;
        }
        if ((e = hipEventRecord(st.traced[s], ctx->stream)) != hipSuccess ||
            (e = hipStreamWaitEvent(st.d2h, st.traced[s], 0)) != hipSuccess ||
            (e = hipMemcpyAsync(st.pin_out[s], dev_hits + off, m * sizeof(spk_hit), hipMemcpyDeviceToHost, st.d2h)) !=
                hipSuccess ||
            (e = hipEventRecord(st.out_done[s], st.d2h)) != hipSuccess)
            break;
        // copy out two chunks behind, so the wait rarely blocks the next chunk's staging copy
        while (e == hipSuccess && drained + 2u <= c) e = drain(drained++);
        if (e != hipSuccess) break;
    }
    while (e == hipSuccess && drained < chunks) e = drain(drained++);
    if (e != hipSuccess) {
        drain_all();
        return hip_fail(ctx, e, "host-buffer trace pipeline");
    }
    return spk_synchronize(ctx);
}

}  // namespace spk
