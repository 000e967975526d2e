// Serving: continuous batching of device-resident frames over several HIP
// lanes (no counterpart in the reference, whose calls are synchronous and
// serialised per handle, gpu_detector.cu:1839-1841).
//
// A server owns, per frame-size class, `lanes` batch engines of
// `lane_frames` positions each: the preallocated device arena.  Requests are
// admitted into a per-class queue of bounded depth (beyond it they are
// dropped).  One scheduler thread busy-polls: a finished lane hands its
// results to the completion list and becomes free; a free lane takes up to
// `lane_frames` queued requests of its class and launches them at once, so
// requests that arrive while other lanes are on the GPU start as soon as a
// lane frees up and their work overlaps on the device (one stream per lane).
// Every request is processed by exactly the same kernels as a single-frame
// call, so its result does not depend on the batch it rides in.
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

class Server {
public:
    Server() = default;
    ~Server() { stop(); }

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

    int64_t submit(int cls, const uint8_t* gray, int32_t stride, Clock::time_point t_arrival) {
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

    int poll(bd_server_result* results, int max_results, bd_detection* dets, int dets_per_result) {
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

    bool wait_idle(int timeout_ms) {
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

    void set_keep_detections(bool keep) { keep_dets_ = keep; }
    std::vector<bd_server_result> take_all() {
        const std::lock_guard<std::mutex> lock(mutex_);
        std::vector<bd_server_result> out;
        out.reserve(completed_.size());
        for (const Completed& c : completed_) out.push_back(c.result);
        completed_.clear();
        return out;
    }
    int64_t dropped() const { return dropped_; }
    int64_t launches() const { return launches_; }
    int64_t launched_frames() const { return launched_frames_; }
    const std::vector<bd_server_class>& classes() const { return classes_; }
    bool failed() const { return failed_; }

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
                    failed_ = true;
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

}  // namespace bd

struct bd_server {
    bd::Server server;
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

bd_server* bd_server_create(int32_t family, int32_t nclasses, const bd_server_class* classes,
                            const bd_config* config, int32_t* status_out) {
    if (status_out) *status_out = BD_OK;
    if ((family != 0 && family != 1) || nclasses < 1 || !classes) {
        g_server_error = "invalid family or class list";
        if (status_out) *status_out = BD_ERROR_INVALID_ARGUMENT;
        return nullptr;
    }
    for (int i = 0; i < nclasses; ++i) {
        const bd_server_class& c = classes[i];
        if (c.width < 16 || c.width > BD_MAX_WIDTH || c.height < 16 || c.height > BD_MAX_HEIGHT ||
            c.lanes < 1 || c.lane_frames < 1 || c.queue_depth < 0) {
            g_server_error = "invalid server class";
            if (status_out) *status_out = BD_ERROR_INVALID_ARGUMENT;
            return nullptr;
        }
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
    if (!s->server.init(p, std::vector<bd_server_class>(classes, classes + nclasses), &error)) {
        delete s;
        g_server_error = error.empty() ? "failed to create HIP server" : error;
        if (status_out) *status_out = BD_ERROR_CUDA;
        return nullptr;
    }
    return s;
}

void bd_server_destroy(bd_server* server) { delete server; }

int64_t bd_server_submit(bd_server* server, int32_t class_index, const uint8_t* gray_device,
                         int32_t stride) {
    if (!server || !gray_device) return BD_ERROR_INVALID_ARGUMENT;
    const auto& classes = server->server.classes();
    if (class_index < 0 || class_index >= static_cast<int>(classes.size()) || stride < classes[class_index].width ||
        !bd::frame_span_fits(classes[class_index].width, classes[class_index].height, stride))
        return BD_ERROR_INVALID_ARGUMENT;
    return server->server.submit(class_index, gray_device, stride, bd::Clock::now());
}

int32_t bd_server_poll(bd_server* server, bd_server_result* results, int32_t max_results,
                       bd_detection* dets, int32_t dets_per_result) {
    if (!server || !results || max_results < 0) return BD_ERROR_INVALID_ARGUMENT;
    if (server->server.failed()) return BD_ERROR_CUDA;
    return server->server.poll(results, max_results, dets, dets_per_result);
}

int32_t bd_server_wait_idle(bd_server* server, int32_t timeout_ms) {
    if (!server) return BD_ERROR_INVALID_ARGUMENT;
    if (!server->server.wait_idle(timeout_ms)) return BD_ERROR_UNKNOWN;
    return server->server.failed() ? BD_ERROR_CUDA : BD_OK;
}

int32_t bd_server_simulate(bd_server* server, const bd_stream_spec* streams, int32_t nstreams,
                           double seconds, uint64_t seed, bd_serve_report* report) {
    if (!server || !streams || nstreams < 1 || seconds <= 0.0 || !report) return BD_ERROR_INVALID_ARGUMENT;
    bd::Server& srv = server->server;
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
