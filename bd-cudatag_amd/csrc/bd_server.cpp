// Serving: continuous batching of device-resident frames (no counterpart in
// the reference, whose calls are synchronous and serialised per handle,
// gpu_detector.cu:1839-1841).  Two servers share one C ABI:
//
// HostServer.  Per frame-size class, `lanes` batch engines of `lane_frames`
// positions each form the preallocated device arena.  Requests wait in a
// per-class host queue of bounded depth (beyond it they are dropped).  One
// scheduler thread busy-polls: a finished lane hands its results over and
// becomes free; a free lane takes up to `lane_frames` queued requests of its
// class and launches them at once on its own stream.
//
// ResidentServer.  The same lanes, but admission runs on the device.
// Requests are published into a per-class ring in pinned host memory; every
// lane launch starts with k_admit, which waits on the device (bounded) for
// published requests and claims a run of them for the lane's positions, and
// ends with k_publish, which writes each request's records, count and
// completion flag into the ring.  The host thread only re-launches each
// lane's captured graph as it finishes and timestamps completions; it never
// reads a request or chooses a batch.
//
// Either way every request runs exactly the kernels of a single-frame call,
// so its result does not depend on the batch it rides in.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "bd_engine.h"
#include "bdcudatag/bdcudatag.h"

namespace bd {
namespace {

using Clock = std::chrono::steady_clock;

inline void cpu_relax() { __builtin_ia32_pause(); }

struct Request {
    int64_t id;
    int cls;
    const uint8_t* gray;
    int32_t stride;
    Clock::time_point t_arrival;
};

struct Completed {
    bd_server_result result;
    std::vector<bd_detection> dets;
};

struct Lane {
    std::unique_ptr<Engine> engine;
    int cls = 0;
    bool busy = false;
    std::vector<Request> reqs;
};

}  // namespace

// What the C ABI needs of a server.
class ServerBase {
public:
    virtual ~ServerBase() = default;
    virtual int64_t submit(int cls, const uint8_t* gray, int32_t stride, Clock::time_point t_arrival) = 0;
    virtual int poll(bd_server_result* results, int max_results, bd_detection* dets, int dets_per_result) = 0;
    virtual bool wait_idle(int timeout_ms) = 0;
    virtual void set_keep_detections(bool keep) = 0;
    virtual std::vector<bd_server_result> take_all() = 0;
    virtual int64_t dropped() const = 0;
    virtual int64_t launches() const = 0;
    virtual int64_t launched_frames() const = 0;
    virtual const std::vector<bd_server_class>& classes() const = 0;
    virtual bool failed() const = 0;
};

class HostServer : public ServerBase {
public:
    HostServer() = default;
    ~HostServer() override { stop(); }

    bool init(const Params& params, const std::vector<bd_server_class>& classes, std::string* error) {
        classes_ = classes;
        max_det_ = std::min(std::max(params.max_detections, 1), 256);
        pending_.resize(classes.size());
        for (size_t c = 0; c < classes.size(); ++c) {
            const bd_server_class& k = classes[c];
            for (int l = 0; l < k.lanes; ++l) {
                Lane lane;
                lane.cls = static_cast<int>(c);
                lane.engine.reset(Engine::create(params, std::vector<int>(k.lane_frames, k.width),
                                                 std::vector<int>(k.lane_frames, k.height), error));
                if (!lane.engine) return false;
                lanes_.push_back(std::move(lane));
            }
        }
        running_ = true;
        thread_ = std::thread([this] { loop(); });
        return true;
    }

    void stop() {
        if (running_.exchange(false) && thread_.joinable()) thread_.join();
    }

    int64_t submit(int cls, const uint8_t* gray, int32_t stride, Clock::time_point t_arrival) override {
        if (cls < 0 || cls >= static_cast<int>(classes_.size())) return BD_ERROR_INVALID_ARGUMENT;
        const std::lock_guard<std::mutex> lock(mutex_);
        if (static_cast<int>(pending_[cls].size()) >= classes_[cls].queue_depth) {
            ++dropped_;
            return BD_ERROR_OUT_OF_MEMORY;
        }
        const int64_t id = next_id_++;
        pending_[cls].push_back(Request{id, cls, gray, stride, t_arrival});
        ++outstanding_;
        return id;
    }

    int poll(bd_server_result* results, int max_results, bd_detection* dets, int dets_per_result) override {
        const std::lock_guard<std::mutex> lock(mutex_);
        int n = 0;
        while (n < max_results && !completed_.empty()) {
            Completed& c = completed_.front();
            bd_server_result r = c.result;
            r.written = 0;
            if (dets && dets_per_result > 0) {
                r.written = std::min(static_cast<int>(c.dets.size()), dets_per_result);
                std::memcpy(dets + static_cast<size_t>(n) * dets_per_result, c.dets.data(),
                            static_cast<size_t>(r.written) * sizeof(bd_detection));
            }
            results[n++] = r;
            completed_.pop_front();
        }
        return n;
    }

    bool wait_idle(int timeout_ms) override {
        const auto deadline = Clock::now() + std::chrono::milliseconds(timeout_ms);
        while (Clock::now() < deadline) {
            {
                const std::lock_guard<std::mutex> lock(mutex_);
                if (outstanding_ == 0) return true;
            }
            std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
        return false;
    }

    void set_keep_detections(bool keep) override { keep_dets_ = keep; }
    std::vector<bd_server_result> take_all() override {
        const std::lock_guard<std::mutex> lock(mutex_);
        std::vector<bd_server_result> out;
        out.reserve(completed_.size());
        for (const Completed& c : completed_) out.push_back(c.result);
        completed_.clear();
        return out;
    }
    int64_t dropped() const override { return dropped_; }
    int64_t launches() const override { return launches_; }
    int64_t launched_frames() const override { return launched_frames_; }
    const std::vector<bd_server_class>& classes() const override { return classes_; }
    bool failed() const override { return failed_; }

private:
    void loop() {
        std::vector<const uint8_t*> grays;
        std::vector<int32_t> strides;
        std::vector<bd_detection> scratch(static_cast<size_t>(max_det_));
        while (running_) {
            bool progress = false;
            for (Lane& lane : lanes_) {
                if (!lane.busy) continue;
                const int q = lane.engine->query_async();
                if (q == 0) continue;
                const auto t_done = Clock::now();
                std::vector<Completed> done;
                done.reserve(lane.reqs.size());
                for (size_t i = 0; i < lane.reqs.size(); ++i) {
                    Completed c{};
                    const Request& r = lane.reqs[i];
                    int found = -1;
                    if (q > 0) found = lane.engine->collect_async(static_cast<int>(i), scratch.data(), max_det_);
                    c.result.request = r.id;
                    c.result.class_index = r.cls;
                    c.result.count = found;
                    c.result.written = 0;
                    c.result.batch_frames = static_cast<int32_t>(lane.reqs.size());
                    c.result.latency_ms = std::chrono::duration<double, std::milli>(t_done - r.t_arrival).count();
                    if (keep_dets_ && found > 0) c.dets.assign(scratch.begin(), scratch.begin() + found);
                    done.push_back(std::move(c));
                }
                if (q < 0) failed_ = true;
                lane.reqs.clear();
                lane.busy = false;
                progress = true;
                const std::lock_guard<std::mutex> lock(mutex_);
                for (Completed& c : done) completed_.push_back(std::move(c));
                outstanding_ -= static_cast<int64_t>(done.size());
            }
            for (Lane& lane : lanes_) {
                if (lane.busy) continue;
                const int cap = lane.engine->frame_count();
                {
                    const std::lock_guard<std::mutex> lock(mutex_);
                    std::deque<Request>& q = pending_[lane.cls];
                    while (!q.empty() && static_cast<int>(lane.reqs.size()) < cap) {
                        lane.reqs.push_back(q.front());
                        q.pop_front();
                    }
                }
                if (lane.reqs.empty()) continue;
                grays.clear();
                strides.clear();
                for (const Request& r : lane.reqs) {
                    grays.push_back(r.gray);
                    strides.push_back(r.stride);
                }
                std::string error;
                if (!lane.engine->launch_async(grays.data(), strides.data(), static_cast<int>(lane.reqs.size()), &error)) {
                    // Nothing was queued: the requests complete at once with
                    // no results and the lane stays free.
                    failed_ = true;
                    const auto t_fail = Clock::now();
                    const std::lock_guard<std::mutex> lock(mutex_);
                    for (const Request& r : lane.reqs) {
                        Completed c{};
                        c.result.request = r.id;
                        c.result.class_index = r.cls;
                        c.result.count = -1;
                        c.result.batch_frames = static_cast<int32_t>(lane.reqs.size());
                        c.result.latency_ms = std::chrono::duration<double, std::milli>(t_fail - r.t_arrival).count();
                        completed_.push_back(std::move(c));
                    }
                    outstanding_ -= static_cast<int64_t>(lane.reqs.size());
                    lane.reqs.clear();
                    progress = true;
                    continue;
                }
                lane.busy = true;
                ++launches_;
                launched_frames_ += static_cast<int64_t>(lane.reqs.size());
                progress = true;
            }
            if (!progress) {
                for (int i = 0; i < 32; ++i) cpu_relax();
            }
        }
        // Drain lanes still in flight so the engines can be destroyed safely.
        for (Lane& lane : lanes_) {
            if (!lane.busy) continue;
            while (lane.engine->query_async() == 0) cpu_relax();
        }
    }

    std::vector<bd_server_class> classes_;
    std::vector<Lane> lanes_;
    std::vector<std::deque<Request>> pending_;
    std::deque<Completed> completed_;
    std::mutex mutex_;
    std::thread thread_;
    std::atomic<bool> running_{false};
    std::atomic<bool> keep_dets_{true};
    std::atomic<bool> failed_{false};
    int64_t next_id_ = 0;
    int64_t outstanding_ = 0;
    std::atomic<int64_t> dropped_{0};
    std::atomic<int64_t> launches_{0};
    std::atomic<int64_t> launched_frames_{0};
    int max_det_ = 64;
};

class ResidentServer : public ServerBase {
public:
    ResidentServer() = default;
    ~ResidentServer() override { release(); }

    bool init(const Params& params, const std::vector<bd_server_class>& classes, double wait_us, std::string* error) {
        classes_ = classes;
        max_det_ = max_detections_for(params);
        int dev = 0, khz = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
            khz = 100000;  // 100 MHz, the usual device wall clock
        const uint64_t wait_ticks = static_cast<uint64_t>(wait_us * 1e-3 * static_cast<double>(khz));
        if (!host_alloc(reinterpret_cast<void**>(&stop_), sizeof(uint32_t), error)) return false;
        *stop_ = 0;
        rings_.resize(classes.size());
        for (size_t c = 0; c < classes.size(); ++c) {
            Ring& r = rings_[c];
            r.depth = classes[c].queue_depth > 0 ? classes[c].queue_depth : 256;
            if (!host_alloc(reinterpret_cast<void**>(&r.slots), sizeof(RequestSlot) * r.depth, error) ||
                !host_alloc(reinterpret_cast<void**>(&r.done), sizeof(CompletionSlot) * r.depth, error) ||
                !host_alloc(reinterpret_cast<void**>(&r.out), sizeof(bd_detection) * r.depth * max_det_, error) ||
                !dev_alloc(reinterpret_cast<void**>(&r.head), sizeof(uint32_t), error))
                return false;
            std::memset(r.slots, 0, sizeof(RequestSlot) * r.depth);
            std::memset(r.done, 0, sizeof(CompletionSlot) * r.depth);
            r.ids.assign(r.depth, -1);
            r.arrival.assign(r.depth, Clock::time_point{});
            r.recorded.assign(r.depth, 0);
        }
        for (size_t c = 0; c < classes.size(); ++c) {
            const bd_server_class& k = classes[c];
            for (int l = 0; l < k.lanes; ++l) {
                RLane lane;
                lane.cls = static_cast<int>(c);
                lane.engine.reset(Engine::create(params, std::vector<int>(k.lane_frames, k.width),
                                                 std::vector<int>(k.lane_frames, k.height), error));
                if (!lane.engine) return false;
                if (!dev_alloc(reinterpret_cast<void**>(&lane.tickets), sizeof(uint32_t) * (k.lane_frames + 1), error) ||
                    !dev_alloc(reinterpret_cast<void**>(&lane.out),
                               sizeof(bd_detection) * static_cast<size_t>(k.lane_frames) * max_det_, error))
                    return false;
                Ring& r = rings_[c];
                ResidentArgs a{};
                a.slots = r.slots;
                a.done = r.done;
                a.out = r.out;
                a.stop = stop_;
                a.head = r.head;
                a.lane_tickets = lane.tickets;
                a.lane_out = lane.out;
                a.depth = r.depth;
                a.max_det = max_det_;
                a.wait_ticks = wait_ticks;
                a.max_polls = 1u << 20;
                if (!lane.engine->prepare_resident(a, error)) return false;
                lanes_.push_back(std::move(lane));
            }
        }
        running_ = true;
        thread_ = std::thread([this] { loop(); });
        return true;
    }

    int64_t submit(int cls, const uint8_t* gray, int32_t stride, Clock::time_point t_arrival) override {
        if (cls < 0 || cls >= static_cast<int>(classes_.size())) return BD_ERROR_INVALID_ARGUMENT;
        const std::lock_guard<std::mutex> lock(submit_mutex_);
        Ring& r = rings_[cls];
        const uint32_t t = r.tail;
        if (t - r.harvested.load(std::memory_order_acquire) >= static_cast<uint32_t>(r.depth)) {
            ++dropped_;
            return BD_ERROR_OUT_OF_MEMORY;
        }
        const uint32_t slot = t % static_cast<uint32_t>(r.depth);
        const int64_t id = next_id_++;
        r.ids[slot] = id;
        r.arrival[slot] = t_arrival;
        r.recorded[slot] = 0;
        r.slots[slot].gray = gray;
        r.slots[slot].stride = stride;
        __atomic_store_n(&r.slots[slot].seq, t + 1u, __ATOMIC_RELEASE);  // publishes the request
        __atomic_store_n(&r.tail, t + 1u, __ATOMIC_RELEASE);
        outstanding_.fetch_add(1);
        return id;
    }

    int poll(bd_server_result* results, int max_results, bd_detection* dets, int dets_per_result) override {
        const std::lock_guard<std::mutex> lock(done_mutex_);
        int n = 0;
        while (n < max_results && !completed_.empty()) {
            Completed& c = completed_.front();
            bd_server_result r = c.result;
            r.written = 0;
            if (dets && dets_per_result > 0) {
                r.written = std::min(static_cast<int>(c.dets.size()), dets_per_result);
                std::memcpy(dets + static_cast<size_t>(n) * dets_per_result, c.dets.data(),
                            static_cast<size_t>(r.written) * sizeof(bd_detection));
            }
            results[n++] = r;
            completed_.pop_front();
        }
        return n;
    }

    bool wait_idle(int timeout_ms) override {
        const auto deadline = Clock::now() + std::chrono::milliseconds(timeout_ms);
        while (Clock::now() < deadline) {
            if (outstanding_.load() == 0) return true;
            std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
        return false;
    }

    void set_keep_detections(bool keep) override { keep_dets_ = keep; }
    std::vector<bd_server_result> take_all() override {
        const std::lock_guard<std::mutex> lock(done_mutex_);
        std::vector<bd_server_result> out;
        out.reserve(completed_.size());
        for (const Completed& c : completed_) out.push_back(c.result);
        completed_.clear();
        return out;
    }
    int64_t dropped() const override { return dropped_; }
    // Launches that admitted requests, counted from the completions: a
    // launch of b requests gives each of them batch = b.
    int64_t launches() const override {
        const std::lock_guard<std::mutex> lock(done_mutex_);
        return static_cast<int64_t>(launch_weight_ + 0.5);
    }
    int64_t launched_frames() const override {
        const std::lock_guard<std::mutex> lock(done_mutex_);
        return completed_total_;
    }
    const std::vector<bd_server_class>& classes() const override { return classes_; }
    bool failed() const override { return failed_; }

private:
    struct Ring {
        RequestSlot* slots = nullptr;
        CompletionSlot* done = nullptr;
        bd_detection* out = nullptr;
        uint32_t* head = nullptr;  // device memory
        int depth = 0;
        uint32_t tail = 0;                       // next ticket (submit_mutex_)
        std::atomic<uint32_t> harvested{0};      // tickets below are recorded and reusable
        std::vector<int64_t> ids;
        std::vector<Clock::time_point> arrival;
        std::vector<char> recorded;
        Ring() = default;
        Ring(Ring&& o) noexcept
            : slots(o.slots), done(o.done), out(o.out), head(o.head), depth(o.depth), tail(o.tail),
              harvested(o.harvested.load()), ids(std::move(o.ids)), arrival(std::move(o.arrival)),
              recorded(std::move(o.recorded)) {}
    };
    struct RLane {
        std::unique_ptr<Engine> engine;
        int cls = 0;
        bool in_flight = false;
        uint32_t* tickets = nullptr;   // device memory
        bd_detection* out = nullptr;   // device memory
    };

    static bool host_alloc(void** p, size_t bytes, std::string* error) {
        const hipError_t st = hipHostMalloc(p, bytes, hipHostMallocPortable);
        if (st != hipSuccess) {
            *p = nullptr;
            if (error) *error = std::string("hipHostMalloc: ") + hipGetErrorString(st);
            return false;
        }
        return true;
    }
    static bool dev_alloc(void** p, size_t bytes, std::string* error) {
        hipError_t st = hipMalloc(p, bytes);
        if (st == hipSuccess) st = hipMemset(*p, 0, bytes);
        if (st != hipSuccess) {
            if (error) *error = std::string("hipMalloc: ") + hipGetErrorString(st);
            return false;
        }
        return true;
    }

    // Records every completion the device has published, in any order, and
    // advances each ring's harvested mark over its recorded prefix.
    bool harvest() {
        bool any = false;
        double weight = 0.0;
        std::vector<Completed> done;
        for (size_t c = 0; c < rings_.size(); ++c) {
            Ring& r = rings_[c];
            const uint32_t tail = __atomic_load_n(&r.tail, __ATOMIC_ACQUIRE);
            uint32_t h = r.harvested.load(std::memory_order_relaxed);
            for (uint32_t t = h; t != tail; ++t) {
                const uint32_t slot = t % static_cast<uint32_t>(r.depth);
                if (r.recorded[slot]) continue;
                if (__atomic_load_n(&r.done[slot].seq, __ATOMIC_ACQUIRE) != t + 1u) continue;
                const auto t_done = Clock::now();
                Completed out{};
                out.result.request = r.ids[slot];
                out.result.class_index = static_cast<int32_t>(c);
                out.result.count = r.done[slot].count;
                out.result.batch_frames = r.done[slot].batch;
                out.result.latency_ms = std::chrono::duration<double, std::milli>(t_done - r.arrival[slot]).count();
                if (keep_dets_ && out.result.count > 0) {
                    const bd_detection* src = r.out + static_cast<size_t>(slot) * max_det_;
                    out.dets.assign(src, src + out.result.count);
                }
                if (out.result.batch_frames > 0) weight += 1.0 / out.result.batch_frames;
                r.recorded[slot] = 1;
                done.push_back(std::move(out));
                any = true;
            }
            while (h != tail && r.recorded[h % static_cast<uint32_t>(r.depth)]) ++h;
            r.harvested.store(h, std::memory_order_release);
        }
        if (!done.empty()) {
            const std::lock_guard<std::mutex> lock(done_mutex_);
            for (Completed& c : done) completed_.push_back(std::move(c));
            completed_total_ += static_cast<int64_t>(done.size());
            launch_weight_ += weight;
            outstanding_.fetch_sub(static_cast<int64_t>(done.size()));
        }
        return any;
    }

    void loop() {
        while (running_) {
            bool progress = false;
            for (RLane& lane : lanes_) {
                if (lane.in_flight) {
                    const int q = lane.engine->query_async();
                    if (q == 0) continue;
                    if (q < 0) failed_ = true;
                    lane.in_flight = false;
                }
                std::string error;
                if (!lane.engine->launch_resident(&error)) {
                    failed_ = true;
                    continue;
                }
                lane.in_flight = true;
                progress = true;
            }
            progress |= harvest();
            if (!progress) {
                for (int i = 0; i < 32; ++i) cpu_relax();
            }
        }
        // Every admission wait ends at the stop flag; drain the lanes.
        __atomic_store_n(stop_, 1u, __ATOMIC_RELEASE);
        for (RLane& lane : lanes_) {
            while (lane.in_flight && lane.engine->query_async() == 0) cpu_relax();
            lane.in_flight = false;
        }
        harvest();
    }

    void release() {
        if (running_.exchange(false) && thread_.joinable()) thread_.join();
        else if (stop_) __atomic_store_n(stop_, 1u, __ATOMIC_RELEASE);
        for (RLane& lane : lanes_) {
            lane.engine.reset();
            if (lane.tickets) (void)hipFree(lane.tickets);
            if (lane.out) (void)hipFree(lane.out);
        }
        lanes_.clear();
        for (Ring& r : rings_) {
            if (r.slots) (void)hipHostFree(r.slots);
            if (r.done) (void)hipHostFree(r.done);
            if (r.out) (void)hipHostFree(r.out);
            if (r.head) (void)hipFree(r.head);
        }
        rings_.clear();
        if (stop_) (void)hipHostFree(stop_);
        stop_ = nullptr;
    }

    std::vector<bd_server_class> classes_;
    std::vector<Ring> rings_;
    std::vector<RLane> lanes_;
    uint32_t* stop_ = nullptr;  // pinned host memory
    std::deque<Completed> completed_;
    std::mutex submit_mutex_;
    mutable std::mutex done_mutex_;
    std::thread thread_;
    std::atomic<bool> running_{false};
    std::atomic<bool> keep_dets_{true};
    std::atomic<bool> failed_{false};
    int64_t next_id_ = 0;
    std::atomic<int64_t> outstanding_{0};
    std::atomic<int64_t> dropped_{0};
    double launch_weight_ = 0.0;
    int64_t completed_total_ = 0;
    int max_det_ = 64;
};

}  // namespace bd

struct bd_server {
    std::unique_ptr<bd::ServerBase> server;
};

namespace {
thread_local std::string g_server_error;

double percentile(std::vector<double>& v, double q) {
    if (v.empty()) return 0.0;
    const size_t k = std::min(v.size() - 1, static_cast<size_t>(std::ceil(q * v.size())) - (q > 0 ? 1 : 0));
    std::nth_element(v.begin(), v.begin() + static_cast<long>(k), v.end());
    return v[k];
}

uint64_t splitmix(uint64_t& x) {
    x += 0x9e3779b97f4a7c15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
}  // namespace

extern "C" {

const char* bd_server_last_error(void) { return g_server_error.c_str(); }

}  // extern "C"

namespace {

// Checks the arguments and builds the server of either kind.
bd_server* create_server(int32_t family, int32_t nclasses, const bd_server_class* classes, const bd_config* config,
                         bool resident, double wait_us, int32_t* status_out) {
    if (status_out) *status_out = BD_OK;
    if ((family != 0 && family != 1) || nclasses < 1 || !classes) {
        g_server_error = "invalid family or class list";
        if (status_out) *status_out = BD_ERROR_INVALID_ARGUMENT;
        return nullptr;
    }
    int lanes = 0;
    for (int i = 0; i < nclasses; ++i) {
        const bd_server_class& c = classes[i];
        if (c.width < 16 || c.width > BD_MAX_WIDTH || c.height < 16 || c.height > BD_MAX_HEIGHT ||
            c.lanes < 1 || c.lane_frames < 1 || c.queue_depth < 0) {
            g_server_error = "invalid server class";
            if (status_out) *status_out = BD_ERROR_INVALID_ARGUMENT;
            return nullptr;
        }
        lanes += c.lanes;
    }
    if (resident && (lanes > 16 || !(wait_us > 0.0 && wait_us <= 1e6))) {
        g_server_error = "a resident server takes at most 16 lanes and an admission wait in (0, 1e6] us";
        if (status_out) *status_out = BD_ERROR_INVALID_ARGUMENT;
        return nullptr;
    }
    if (config && config->struct_size != 0 && config->struct_size < sizeof(bd_config)) {
        g_server_error = "bd_config.struct_size is too small for this library version";
        if (status_out) *status_out = BD_ERROR_INVALID_ARGUMENT;
        return nullptr;
    }
    bd_config defaults;
    bd_config_init(&defaults);
    const bd_config* c = config ? config : &defaults;
    bd::Params p;
    p.family = family;
    p.max_hamming = c->max_hamming;
    p.min_cluster_pixels = c->min_cluster_pixels;
    p.min_tag_pixels = c->min_tag_pixels;
    p.max_detections = c->max_detections;
    p.max_candidates = c->max_candidates;
    p.refine_edges = c->refine_edges != 0;
    p.use_graph = c->use_cuda_graph != 0;
    p.min_decision_margin = c->min_decision_margin;
    p.decode_sharpening = c->decode_sharpening;
    bd_server* s = new (std::nothrow) bd_server();
    if (!s) {
        g_server_error = "out of memory";
        if (status_out) *status_out = BD_ERROR_OUT_OF_MEMORY;
        return nullptr;
    }
    std::string error;
    const std::vector<bd_server_class> list(classes, classes + nclasses);
    bool ok;
    if (resident) {
        auto srv = std::make_unique<bd::ResidentServer>();
        ok = srv->init(p, list, wait_us, &error);
        s->server = std::move(srv);
    } else {
        auto srv = std::make_unique<bd::HostServer>();
        ok = srv->init(p, list, &error);
        s->server = std::move(srv);
    }
    if (!ok) {
        delete s;
        g_server_error = error.empty() ? "failed to create HIP server" : error;
        if (status_out) *status_out = BD_ERROR_CUDA;
        return nullptr;
    }
    return s;
}

}  // namespace

extern "C" {

bd_server* bd_server_create(int32_t family, int32_t nclasses, const bd_server_class* classes,
                            const bd_config* config, int32_t* status_out) {
    return create_server(family, nclasses, classes, config, false, 0.0, status_out);
}

bd_server* bd_server_create_resident(int32_t family, int32_t nclasses, const bd_server_class* classes,
                                     const bd_config* config, double admit_wait_us, int32_t* status_out) {
    return create_server(family, nclasses, classes, config, true, admit_wait_us, status_out);
}

void bd_server_destroy(bd_server* server) { delete server; }

int64_t bd_server_submit(bd_server* server, int32_t class_index, const uint8_t* gray_device,
                         int32_t stride) {
    if (!server || !gray_device) return BD_ERROR_INVALID_ARGUMENT;
    const auto& classes = server->server->classes();
    if (class_index < 0 || class_index >= static_cast<int>(classes.size()) || stride < classes[class_index].width ||
        !bd::frame_span_fits(classes[class_index].width, classes[class_index].height, stride))
        return BD_ERROR_INVALID_ARGUMENT;
    return server->server->submit(class_index, gray_device, stride, bd::Clock::now());
}

int32_t bd_server_poll(bd_server* server, bd_server_result* results, int32_t max_results,
                       bd_detection* dets, int32_t dets_per_result) {
    if (!server || !results || max_results < 0) return BD_ERROR_INVALID_ARGUMENT;
    if (server->server->failed()) return BD_ERROR_CUDA;
    return server->server->poll(results, max_results, dets, dets_per_result);
}

int32_t bd_server_wait_idle(bd_server* server, int32_t timeout_ms) {
    if (!server) return BD_ERROR_INVALID_ARGUMENT;
    if (!server->server->wait_idle(timeout_ms)) return BD_ERROR_UNKNOWN;
    return server->server->failed() ? BD_ERROR_CUDA : BD_OK;
}

int32_t bd_server_simulate(bd_server* server, const bd_stream_spec* streams, int32_t nstreams,
                           double seconds, uint64_t seed, bd_serve_report* report) {
    if (!server || !streams || nstreams < 1 || seconds <= 0.0 || !report) return BD_ERROR_INVALID_ARGUMENT;
    bd::ServerBase& srv = *server->server;
    const auto& classes = srv.classes();
    struct Arrival {
        double t;
        int stream;
        int frame;
    };
    std::vector<Arrival> arrivals;
    for (int s = 0; s < nstreams; ++s) {
        const bd_stream_spec& st = streams[s];
        if (st.class_index < 0 || st.class_index >= static_cast<int>(classes.size()) || st.fps <= 0.0 ||
            st.frame_count < 1 || !st.frames || st.stride < classes[st.class_index].width ||
            !bd::frame_span_fits(classes[st.class_index].width, classes[st.class_index].height, st.stride))
            return BD_ERROR_INVALID_ARGUMENT;
        uint64_t rng = seed ^ (0x632be59bd9b4e019ull * static_cast<uint64_t>(s + 1));
        double t = 0.0;
        int k = 0;
        for (;;) {
            const double u = (static_cast<double>(splitmix(rng) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
            t += -std::log(u) / st.fps;
            if (t >= seconds) break;
            arrivals.push_back({t, s, k++ % st.frame_count});
        }
    }
    std::sort(arrivals.begin(), arrivals.end(), [](const Arrival& a, const Arrival& b) { return a.t < b.t; });
    srv.set_keep_detections(false);
    (void)srv.take_all();
    const int64_t dropped0 = srv.dropped();
    const int64_t launches0 = srv.launches();
    const int64_t frames0 = srv.launched_frames();
    std::vector<double> lat;
    lat.reserve(arrivals.size());
    double pixels = 0.0;
    const auto start = bd::Clock::now();
    int64_t offered = 0;
    for (const Arrival& a : arrivals) {
        const auto target = start + std::chrono::duration_cast<bd::Clock::duration>(std::chrono::duration<double>(a.t));
        for (;;) {
            const auto now = bd::Clock::now();
            if (now >= target) break;
            if (target - now > std::chrono::microseconds(300)) {
                std::this_thread::sleep_for(std::chrono::microseconds(100));
            } else {
                bd::cpu_relax();
            }
        }
        const bd_stream_spec& st = streams[a.stream];
        ++offered;
        srv.submit(st.class_index, st.frames[a.frame], st.stride, target);
        // Results are harvested as they come so the completion list stays small.
        if ((offered & 63) == 0) {
            for (const bd_server_result& r : srv.take_all()) {
                lat.push_back(r.latency_ms);
                pixels += static_cast<double>(classes[r.class_index].width) * classes[r.class_index].height;
            }
        }
    }
    const bool idle = srv.wait_idle(10000);
    const auto end = bd::Clock::now();
    for (const bd_server_result& r : srv.take_all()) {
        lat.push_back(r.latency_ms);
        pixels += static_cast<double>(classes[r.class_index].width) * classes[r.class_index].height;
    }
    srv.set_keep_detections(true);
    const double elapsed = std::chrono::duration<double>(end - start).count();
    std::memset(report, 0, sizeof(*report));
    report->offered = offered;
    report->completed = static_cast<int64_t>(lat.size());
    report->dropped = srv.dropped() - dropped0;
    report->seconds = elapsed;
    report->fps = static_cast<double>(lat.size()) / elapsed;
    report->mpix_s = pixels / elapsed / 1e6;
    report->launches = srv.launches() - launches0;
    report->mean_batch = report->launches > 0
                             ? static_cast<double>(srv.launched_frames() - frames0) / static_cast<double>(report->launches)
                             : 0.0;
    if (!lat.empty()) {
        double sum = 0.0, mx = 0.0;
        for (double v : lat) {
            sum += v;
            mx = std::max(mx, v);
        }
        report->mean_ms = sum / static_cast<double>(lat.size());
        report->max_ms = mx;
        report->p50_ms = percentile(lat, 0.50);
        report->p99_ms = percentile(lat, 0.99);
        report->p999_ms = percentile(lat, 0.999);
    }
    if (srv.failed()) return BD_ERROR_CUDA;
    return idle ? BD_OK : BD_ERROR_UNKNOWN;
}

}  // extern "C"
