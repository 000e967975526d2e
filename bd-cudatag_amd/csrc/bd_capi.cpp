// C ABI and C++ API on top of the HIP engine.  The C ABI keeps the
// reference's observable behaviour (detector/src/detector.cpp:1-225): the
// same validation order, status codes, error strings, thread-local last
// error and config defaults.  The C++ Detector (detector.cpp:227-380 for the
// behaviour: exceptions, clamps, messages) owns its engine directly; both
// APIs share one argument check and call path per frame.
#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "bd_engine.h"
#include "bdcudatag/bdcudatag.h"
#include "bdcudatag/detector.hpp"

namespace {

thread_local std::string g_last_error;

void set_error(const std::string& message) { g_last_error = message; }

bool valid_family(int family) { return family == 0 || family == 1; }

bool valid_dimensions(int width, int height) {
    return width >= 16 && width <= BD_MAX_WIDTH && height >= 16 && height <= BD_MAX_HEIGHT;
}

// The argument checks of detector creation, in the reference's order
// (detector.cpp:112-129); nullptr when the family and size are acceptable.
const char* creation_problem(int family, int width, int height) {
    if (!valid_family(family)) return "unknown tag family";
    if (!valid_dimensions(width, height)) return "image dimensions out of range (16..8192 x 16..16384)";
    return nullptr;
}

// detector.cpp:39-57 (config_to_gpu)
bd::Params to_params(int family, const bd_config* config) {
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
    return p;
}

// One frame through a single-frame engine: the reference's per-call checks
// (detector.cpp:164-218, in its order), then the detection.  Returns the
// number of tags found (records written: min(found, capacity)), or a
// negative bd_status with *why set.
int32_t run_single(bd::Engine& engine, const uint8_t* gray, int32_t stride, bd_detection* out,
                   int32_t capacity, bool device, std::string* why) {
    const int width = engine.width(0), height = engine.height(0);
    const char* bad = nullptr;
    if (!gray) {
        bad = device ? "null device image pointer" : "null image pointer";
    } else if (stride < width) {
        bad = "stride is smaller than the image width";
    } else if (device && !bd::frame_span_fits(width, height, stride)) {
        bad = "device frame spans 4 GiB or more";
    } else if (capacity < 0 || (capacity > 0 && !out)) {
        bad = "invalid output buffer";
    }
    if (bad) {
        *why = bad;
        return BD_ERROR_INVALID_ARGUMENT;
    }
    const bd_frame frame{gray, width, height, stride, 0};
    int32_t found = 0;
    if (engine.detect(&frame, device, out, capacity, &found, why) < 0) {
        if (why->empty()) *why = "HIP detection failed";
        return BD_ERROR_CUDA;
    }
    return found;
}

}  // namespace

struct bd_batch {
    bd::Engine* engine = nullptr;
};

struct bd_detector {
    bd_batch batch;
    int family = 0;
    int width = 0;
    int height = 0;
};

extern "C" {

uint32_t bd_api_version(void) { return BD_API_VERSION; }

const char* bd_version_string(void) { return "bdcudatag 0.1.0 (hip gfx950)"; }

int32_t bd_family_count(void) { return 2; }

const char* bd_family_name(int32_t family) {
    if (!valid_family(family)) return nullptr;
    return family == 0 ? "tag36h11" : "tag16h5";
}

int32_t bd_family_bits_per_side(int32_t family) {
    if (!valid_family(family)) return -1;
    return family == 0 ? 6 : 4;
}

int32_t bd_family_min_hamming(int32_t family) {
    if (!valid_family(family)) return -1;
    return family == 0 ? 11 : 5;
}

int32_t bd_family_default_max_hamming(int32_t family) {
    if (!valid_family(family)) return -1;
    return family == 0 ? 2 : 1;
}

void bd_config_init(bd_config* config) {
    if (!config) return;
    std::memset(config, 0, sizeof(*config));
    config->struct_size = sizeof(bd_config);
    config->max_hamming = -1;
    config->min_cluster_pixels = 24;
    config->min_tag_pixels = 8;
    config->max_detections = 64;
    config->max_candidates = 0;
    config->refine_edges = 1;
    config->use_cuda_graph = 1;
    config->min_decision_margin = 3.0f;
    config->decode_sharpening = 0.25f;
}

// detector.cpp:112-156
bd_detector* bd_detector_create(int32_t family, int32_t width, int32_t height,
                                const bd_config* config, int32_t* status_out) {
    if (status_out) *status_out = BD_OK;
    if (const char* problem = creation_problem(family, width, height)) {
        set_error(problem);
        if (status_out) *status_out = BD_ERROR_INVALID_ARGUMENT;
        return nullptr;
    }
    if (config && config->struct_size != 0 && config->struct_size < sizeof(bd_config)) {
        set_error("bd_config.struct_size is too small for this library version");
        if (status_out) *status_out = BD_ERROR_INVALID_ARGUMENT;
        return nullptr;
    }
    std::string error;
    bd::Engine* engine = bd::Engine::create(to_params(family, config), {width}, {height}, &error);
    if (!engine) {
        set_error(error.empty() ? "failed to create HIP detector" : error);
        if (status_out) *status_out = BD_ERROR_CUDA;
        return nullptr;
    }
    bd_detector* detector = new (std::nothrow) bd_detector();
    if (!detector) {
        delete engine;
        set_error("out of memory allocating detector handle");
        if (status_out) *status_out = BD_ERROR_OUT_OF_MEMORY;
        return nullptr;
    }
    detector->batch.engine = engine;
    detector->family = family;
    detector->width = width;
    detector->height = height;
    return detector;
}

void bd_detector_destroy(bd_detector* detector) {
    if (!detector) return;
    delete detector->batch.engine;
    delete detector;
}

static int32_t detect_one(bd_detector* detector, const uint8_t* gray, int32_t stride,
                          bd_detection* out, int32_t out_capacity, bool device) {
    if (!detector) {
        set_error("null detector");
        return BD_ERROR_INVALID_ARGUMENT;
    }
    std::string why;
    const int32_t found = run_single(*detector->batch.engine, gray, stride, out, out_capacity, device, &why);
    if (found < 0) set_error(why);
    return found;
}

// detector.cpp:164-190
int32_t bd_detector_detect_host(bd_detector* detector, const uint8_t* gray, int32_t stride,
                                bd_detection* out, int32_t out_capacity) {
    return detect_one(detector, gray, stride, out, out_capacity, false);
}

// detector.cpp:192-218
int32_t bd_detector_detect_device(bd_detector* detector, const uint8_t* gray_device,
                                  int32_t stride, bd_detection* out, int32_t out_capacity) {
    return detect_one(detector, gray_device, stride, out, out_capacity, true);
}

void bd_detector_get_stats(bd_detector* detector, bd_stats* out) {
    if (!detector || !out) return;
    detector->batch.engine->stats(0, out);
}

const char* bd_last_error(void) { return g_last_error.c_str(); }

// ---------------------------------------------------------------------------
// Version 3: batches
// ---------------------------------------------------------------------------

bd_batch* bd_batch_create(int32_t family, int32_t count, const int32_t* widths,
                          const int32_t* heights, const bd_config* config, int32_t* status_out) {
    if (status_out) *status_out = BD_OK;
    if (!valid_family(family)) {
        set_error("unknown tag family");
        if (status_out) *status_out = BD_ERROR_INVALID_ARGUMENT;
        return nullptr;
    }
    if (count < 1 || !widths || !heights) {
        set_error("a batch needs at least one frame size");
        if (status_out) *status_out = BD_ERROR_INVALID_ARGUMENT;
        return nullptr;
    }
    for (int i = 0; i < count; ++i) {
        if (!valid_dimensions(widths[i], heights[i])) {
            set_error("image dimensions out of range (16..8192 x 16..16384)");
            if (status_out) *status_out = BD_ERROR_INVALID_ARGUMENT;
            return nullptr;
        }
    }
    if (config && config->struct_size != 0 && config->struct_size < sizeof(bd_config)) {
        set_error("bd_config.struct_size is too small for this library version");
        if (status_out) *status_out = BD_ERROR_INVALID_ARGUMENT;
        return nullptr;
    }
    std::string error;
    bd::Engine* engine = bd::Engine::create(to_params(family, config),
                                            std::vector<int>(widths, widths + count),
                                            std::vector<int>(heights, heights + count), &error);
    if (!engine) {
        set_error(error.empty() ? "failed to create HIP batch detector" : error);
        if (status_out) *status_out = BD_ERROR_CUDA;
        return nullptr;
    }
    bd_batch* batch = new (std::nothrow) bd_batch();
    if (!batch) {
        delete engine;
        set_error("out of memory allocating batch handle");
        if (status_out) *status_out = BD_ERROR_OUT_OF_MEMORY;
        return nullptr;
    }
    batch->engine = engine;
    return batch;
}

void bd_batch_destroy(bd_batch* batch) {
    if (!batch) return;
    delete batch->engine;
    delete batch;
}

static int32_t batch_detect(bd_batch* batch, const bd_frame* frames, int32_t count,
                            bd_detection* out, int32_t out_stride, int32_t* counts, bool device) {
    if (!batch || !batch->engine) {
        set_error("null batch");
        return BD_ERROR_INVALID_ARGUMENT;
    }
    bd::Engine& e = *batch->engine;
    if (!frames || count != e.frame_count()) {
        set_error("frame count does not match the batch");
        return BD_ERROR_INVALID_ARGUMENT;
    }
    for (int i = 0; i < count; ++i) {
        if (!frames[i].gray) {
            set_error(device ? "null device image pointer" : "null image pointer");
            return BD_ERROR_INVALID_ARGUMENT;
        }
        if (frames[i].width != e.width(i) || frames[i].height != e.height(i)) {
            set_error("frame size does not match the batch");
            return BD_ERROR_INVALID_ARGUMENT;
        }
        if (frames[i].stride < frames[i].width) {
            set_error("stride is smaller than the image width");
            return BD_ERROR_INVALID_ARGUMENT;
        }
        if (device && !bd::frame_span_fits(frames[i].width, frames[i].height, frames[i].stride)) {
            set_error("device frame spans 4 GiB or more");
            return BD_ERROR_INVALID_ARGUMENT;
        }
    }
    if (out_stride < 0 || (out_stride > 0 && !out)) {
        set_error("invalid output buffer");
        return BD_ERROR_INVALID_ARGUMENT;
    }
    std::string error;
    const int written = e.detect(frames, device, out, out_stride, counts, &error);
    if (written < 0) {
        set_error(error.empty() ? "HIP detection failed" : error);
        return BD_ERROR_CUDA;
    }
    return written;
}

int32_t bd_batch_detect_host(bd_batch* batch, const bd_frame* frames, int32_t count,
                             bd_detection* out, int32_t out_stride, int32_t* counts) {
    return batch_detect(batch, frames, count, out, out_stride, counts, false);
}

int32_t bd_batch_detect_device(bd_batch* batch, const bd_frame* frames, int32_t count,
                               bd_detection* out, int32_t out_stride, int32_t* counts) {
    return batch_detect(batch, frames, count, out, out_stride, counts, true);
}

void bd_batch_get_stats(bd_batch* batch, int32_t index, bd_stats* out) {
    if (!batch || !batch->engine || !out || index < 0 || index >= batch->engine->frame_count()) return;
    batch->engine->stats(index, out);
}

int64_t bd_batch_debug_copy(bd_batch* batch, int32_t index, int32_t stage, void* dst,
                            int64_t capacity) {
    if (!batch || !batch->engine) {
        set_error("null batch");
        return BD_ERROR_INVALID_ARGUMENT;
    }
    std::string error;
    const int64_t n = batch->engine->debug_copy(index, stage, dst, capacity, &error);
    if (n == -1) {
        set_error(error);
        return BD_ERROR_INVALID_ARGUMENT;
    }
    if (n < 0) {
        set_error(error);
        return BD_ERROR_CUDA;
    }
    return n;
}

int32_t bd_batch_frame_info(bd_batch* batch, int32_t index, int32_t out[8]) {
    if (!batch || !batch->engine || !out || !batch->engine->frame_info(index, out)) {
        set_error("invalid batch or frame index");
        return BD_ERROR_INVALID_ARGUMENT;
    }
    return BD_OK;
}

bd_batch* bd_detector_batch(bd_detector* detector) { return detector ? &detector->batch : nullptr; }

int32_t bd_batch_time_stages(bd_batch* batch, int32_t reps, float* ms) {
    if (!batch || !batch->engine) {
        set_error("null batch");
        return BD_ERROR_INVALID_ARGUMENT;
    }
    std::string error;
    if (!batch->engine->time_stages(reps, ms, &error)) {
        set_error(error);
        return BD_ERROR_CUDA;
    }
    return BD_OK;
}

int32_t bd_batch_replay(bd_batch* batch, int32_t reps, int32_t device_path, float* ms) {
    if (!batch || !batch->engine) {
        set_error("null batch");
        return BD_ERROR_INVALID_ARGUMENT;
    }
    std::string error;
    if (!batch->engine->replay(reps, device_path != 0, ms, &error)) {
        set_error(error);
        return BD_ERROR_CUDA;
    }
    return BD_OK;
}

int32_t bd_batch_get_tune(bd_batch* batch, bd_tune* out) {
    if (!batch || !batch->engine || !out) {
        set_error("null batch or output");
        return BD_ERROR_INVALID_ARGUMENT;
    }
    const bd::Tune t = batch->engine->tune();
    *out = bd_tune{t.fit_grid, t.decode_grid, t.head_ranks, t.streams, t.decode_waves, t.chunk_px, t.head_area,
                   t.band_rows};
    return BD_OK;
}

int32_t bd_batch_set_capture(bd_batch* batch, int32_t on) {
    if (!batch || !batch->engine) {
        set_error("null batch");
        return BD_ERROR_INVALID_ARGUMENT;
    }
    std::string error;
    if (!batch->engine->set_capture(on != 0, &error)) {
        set_error(error);
        return BD_ERROR_CUDA;
    }
    return BD_OK;
}

int32_t bd_batch_set_tune(bd_batch* batch, const bd_tune* tune) {
    if (!batch || !batch->engine || !tune) {
        set_error("null batch or parameters");
        return BD_ERROR_INVALID_ARGUMENT;
    }
    bd::Tune t;
    t.fit_grid = tune->fit_grid;
    t.decode_grid = tune->decode_grid;
    t.head_ranks = tune->head_ranks;
    t.streams = tune->streams == 0 ? bd::Tune{}.streams : tune->streams;
    t.decode_waves = tune->decode_waves == 0 ? bd::Tune{}.decode_waves : tune->decode_waves;
    t.chunk_px = tune->chunk_px == 0 ? bd::Tune{}.chunk_px : tune->chunk_px;
    t.head_area = tune->head_area == 0 ? bd::Tune{}.head_area : tune->head_area;
    t.band_rows = tune->band_rows == 0 ? bd::Tune{}.band_rows : tune->band_rows;
    std::string error;
    if (!batch->engine->set_tune(t, &error)) {
        set_error(error);
        return error == "launch parameters out of range" ? BD_ERROR_INVALID_ARGUMENT : BD_ERROR_CUDA;
    }
    return BD_OK;
}

int32_t bd_batch_autotune(bd_batch* batch, int32_t reps, bd_tune* best, float* best_ms) {
    if (!batch || !batch->engine || reps < 1) {
        set_error("null batch or invalid repetition count");
        return BD_ERROR_INVALID_ARGUMENT;
    }
    bd::Tune t;
    float ms = 0.0f;
    std::string error;
    if (!batch->engine->autotune(reps, &t, &ms, &error)) {
        set_error(error);
        return BD_ERROR_CUDA;
    }
    if (best)
        *best = bd_tune{t.fit_grid, t.decode_grid, t.head_ranks, t.streams, t.decode_waves, t.chunk_px, t.head_area,
                        t.band_rows};
    if (best_ms) *best_ms = ms;
    return BD_OK;
}

int32_t bd_detect_device_noalloc(int32_t family, int32_t width, int32_t height, const bd_config* config,
                                 const uint8_t* gray_device, int32_t stride, bd_detection* out_device,
                                 int32_t out_capacity, int32_t* count_device, void* stream) {
    const char* bad = creation_problem(family, width, height);
    if (!bad && config && config->struct_size != 0 && config->struct_size < sizeof(bd_config))
        bad = "bd_config.struct_size is too small for this library version";
    if (!bad && !gray_device) bad = "null device image pointer";
    if (!bad && stride < width) bad = "stride is smaller than the image width";
    if (!bad && !bd::frame_span_fits(width, height, stride)) bad = "device frame spans 4 GiB or more";
    if (!bad && (out_capacity < 0 || (out_capacity > 0 && !out_device))) bad = "invalid output buffer";
    if (!bad && !count_device) bad = "null device count pointer";
    if (bad) {
        set_error(bad);
        return BD_ERROR_INVALID_ARGUMENT;
    }
    std::string why;
    const int rc = bd::detect_noalloc(to_params(family, config), width, height, gray_device, stride, out_device,
                                      out_capacity, count_device, static_cast<hipStream_t>(stream), &why);
    if (rc == 0) return BD_OK;
    set_error(why);
    return rc == -1 ? BD_ERROR_INVALID_ARGUMENT : BD_ERROR_CUDA;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// C++ API
// ---------------------------------------------------------------------------

namespace bdcudatag {
namespace {

Detection to_cpp(const bd_detection& r) {
    Detection d;
    d.id = r.id;
    d.hamming = r.hamming;
    d.decision_margin = r.decision_margin;
    for (size_t k = 0; k < d.corners.size(); ++k) d.corners[k] = {r.corners[2 * k], r.corners[2 * k + 1]};
    std::copy(r.homography, r.homography + 9, d.homography.begin());
    return d;
}

}  // namespace

// The C++ detector owns its engine outright, with no C handle in between.
// `records` is the engine's output for the current call, reused across
// calls; `lock` serialises the calls that share it (the engine serialises
// its own state too), so one Detector may be shared between threads.
struct Detector::Impl {
    std::unique_ptr<bd::Engine> engine;
    std::vector<bd_detection> records;
    std::mutex lock;

    static Impl& live(Impl* impl) {
        if (!impl) throw std::runtime_error("detector has been moved from");
        return *impl;
    }

    // Host or device input: the same checks, errors and record conversion.
    // Writes at most min(capacity, max_detections) detections.
    int run(const uint8_t* gray, int stride, Detection* out, int capacity, bool device) {
        if (capacity < 0 || (capacity > 0 && !out)) throw std::invalid_argument("invalid detection output buffer");
        capacity = std::min(capacity, static_cast<int>(records.size()));
        const std::lock_guard<std::mutex> hold(lock);
        std::string why;
        const int32_t found = run_single(*engine, gray, stride, records.data(), capacity, device, &why);
        if (found < 0) {
            set_error(why);
            throw std::runtime_error(why);
        }
        const int written = std::min(static_cast<int>(found), capacity);
        std::transform(records.begin(), records.begin() + written, out, to_cpp);
        return written;
    }
};

Detector::Detector(Family family, int width, int height, const DetectorConfig& config)
    : family_(family), width_(width), height_(height) {
    const int fam = static_cast<int>(family);
    if (const char* problem = creation_problem(fam, width, height)) {
        set_error(problem);
        throw std::runtime_error(problem);
    }
    bd::Params p;
    p.family = fam;
    p.max_hamming = config.max_hamming;
    p.min_cluster_pixels = config.min_cluster_pixels;
    p.min_tag_pixels = config.min_tag_pixels;
    p.max_detections = config.max_detections;
    p.max_candidates = config.max_candidates;
    p.refine_edges = config.refine_edges;
    p.use_graph = config.use_cuda_graph;
    p.min_decision_margin = config.min_decision_margin;
    p.decode_sharpening = config.decode_sharpening;
    std::string why;
    std::unique_ptr<bd::Engine> engine(bd::Engine::create(p, {width}, {height}, &why));
    if (!engine) {
        if (why.empty()) why = "failed to create bdcudatag detector";
        set_error(why);
        throw std::runtime_error(why);
    }
    std::unique_ptr<Impl> impl(new Impl());
    impl->records.resize(static_cast<size_t>(engine->max_detections()));
    impl->engine = std::move(engine);
    impl_ = impl.release();
}

Detector::~Detector() { delete impl_; }

Detector::Detector(Detector&& other) noexcept
    : impl_(std::exchange(other.impl_, nullptr)), family_(other.family_), width_(other.width_),
      height_(other.height_) {}

Detector& Detector::operator=(Detector&& other) noexcept {
    if (this != &other) {
        delete std::exchange(impl_, std::exchange(other.impl_, nullptr));
        family_ = other.family_;
        width_ = other.width_;
        height_ = other.height_;
    }
    return *this;
}

int Detector::detect(const uint8_t* gray, int stride, Detection* out, int out_capacity) {
    return Impl::live(impl_).run(gray, stride, out, out_capacity, false);
}

int Detector::detectDevice(const uint8_t* gray_device, int stride, Detection* out, int out_capacity) {
    return Impl::live(impl_).run(gray_device, stride, out, out_capacity, true);
}

std::vector<Detection> Detector::detect(const uint8_t* gray, int stride) {
    Impl& impl = Impl::live(impl_);
    std::vector<Detection> found(impl.records.size());
    found.resize(static_cast<size_t>(impl.run(gray, stride, found.data(), static_cast<int>(found.size()), false)));
    return found;
}

Stats Detector::stats() const {
    Stats s;
    if (!impl_) return s;
    bd_stats n{};
    impl_->engine->stats(0, &n);
    s.frames = n.frames;
    s.last_gpu_ms = n.last_gpu_ms;
    s.average_gpu_ms = n.average_gpu_ms;
    s.last_total_ms = n.last_total_ms;
    s.runs = n.runs;
    s.candidates = n.candidates;
    s.detections = n.detections;
    s.quads = n.quads;
    s.roi_rows = n.roi_rows;
    return s;
}

}  // namespace bdcudatag
