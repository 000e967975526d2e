// Host engine: owns the device arena, the stream, the events and the
// captured hipGraphs for one list of frame sizes.  A single-frame detector
// is an engine with one frame.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "bd_internal.h"
#include "bdcudatag/bdcudatag.h"

namespace bd {

struct Params {
    int family = 0;
    int max_hamming = -1;
    int min_cluster_pixels = 24;
    int min_tag_pixels = 8;
    int max_detections = 64;
    int max_candidates = 0;
    bool refine_edges = true;
    bool use_graph = true;
    float min_decision_margin = 3.0f;
    float decode_sharpening = 0.25f;
};

// One frame's device buffers, as byte offsets from the base of the memory
// that holds them (an engine's arena or the no-allocation scratch array).
struct FrameLayout {
    size_t tmin, tmax, tmean, tthr, trlo, trhi, roi, hist, bits, rowc, runs, label, area, bbox, keys, roots,
        nroots, items, rowi, rext, rxfull, cands, quads, hquads, dec;
};

// Candidates kept per frame: the configured count clamped to [16, 1024], or
// clamp(next_pow2(w * h / 4096) * 3 / 2, 96, 1024) when it is 0 (reference
// gpu_detector.cu:1918-1923).
int max_candidates_for(const Params& p, int width, int height);
// Places one frame's buffers at *cursor (each 256-byte aligned), advancing it.
FrameLayout layout_frame(int width, int height, int max_candidates, int row_ext_log2, size_t* cursor);
// The descriptor of a frame whose buffers start at `base` per `o`: sizes,
// clamped parameters (gpu_detector.cu:1906-1943) and buffer pointers.  The
// input pointer, stride, scalar block and output records are the caller's.
FrameDesc bind_frame(const Params& p, int width, int height, int max_candidates, int max_detections,
                     int row_ext_log2, uint8_t* base, const FrameLayout& o, const uint64_t* codes);
// Resolved family constants.
int max_detections_for(const Params& p);
int code_count_for(int family);
const uint64_t* host_codes_for(int family);

// One frame through the pipeline with no allocation (bd_noalloc.hip): the
// working memory is a static per-device array, the work goes on `stream`
// and the records and count land in the caller's device buffers.  Returns 0
// when queued, -1 when the frame needs more working memory than the array
// holds, -2 on a HIP error (*error says which).
int detect_noalloc(const Params& p, int width, int height, const uint8_t* gray, int stride, bd_detection* out,
                   int capacity, int32_t* count, hipStream_t stream, std::string* error);

struct FrameStats {
    uint64_t frames = 0;
    float last_gpu_ms = 0.0f;
    float average_gpu_ms = 0.0f;
    float last_total_ms = 0.0f;
    int runs = 0, candidates = 0, detections = 0, quads = 0, roi_rows = 0;
};

class Engine {
public:
    static Engine* create(const Params& params, const std::vector<int>& widths,
                          const std::vector<int>& heights, std::string* error);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    int frame_count() const { return static_cast<int>(w_.size()); }
    int width(int i) const { return w_[i]; }
    int height(int i) const { return h_[i]; }
    int max_detections() const { return max_det_; }

    // Runs every frame; per-frame results go to out + i * out_stride.
    // `device` selects whether frames[i].gray are device pointers.  Returns the
    // total written or -1 (error filled).
    int detect(const bd_frame* frames, bool device, bd_detection* out, int out_stride,
               int32_t* counts, std::string* error);

    void stats(int index, bd_stats* out) const;
    int64_t debug_copy(int index, int stage, void* dst, int64_t capacity, std::string* error);
    bool frame_info(int index, int32_t out[8]) const;

    // Per-kernel device time, averaged over `reps` runs of the pipeline on
    // the frames already resident in the engine: HIP events on the engine
    // stream between consecutive kernels.  ms[0..9] per kernel, ms[10] whole.
    bool time_stages(int reps, float* ms, std::string* error);
    // Replays the host-path or device-path graph `reps` times back to back
    // and returns the elapsed device time.
    bool replay(int reps, bool device_path, float* ms, std::string* error);
    hipStream_t stream() const { return stream_; }

    // Launch parameters (results never depend on them).  set_tune waits for
    // the stream and re-captures the graphs.
    Tune tune() const;
    bool set_tune(const Tune& t, std::string* error);
    // Times graph replays of the last call's frames under a small set of
    // launch parameters (one parameter at a time, keeping the best so far)
    // and leaves the fastest set in place.  *best_ms is per pipeline run.
    bool autotune(int reps, Tune* best, float* best_ms, std::string* error);
    // Stage capture: with `on`, later runs also write the runs and labels of
    // the rows inside each band (no kernel reads them; debug_copy shows
    // them).  Waits for the streams and re-captures the graphs.
    bool set_capture(bool on, std::string* error);

    // Asynchronous device-path launch of positions [0, n) (the rest are
    // inactive for this launch).  The caller must not launch again before
    // query_async() reports completion.
    bool launch_async(const uint8_t* const* gray, const int32_t* stride, int n, std::string* error);
    // 1 when the last async launch has finished, 0 while running, -1 on error.
    int query_async();
    // Detections of position `index` of the finished async launch.
    int collect_async(int index, bd_detection* out, int capacity) const;

    // Device-resident serving: every launch admits requests from the ring of
    // `args` on the device (k_admit), runs the pipeline over the admitted
    // positions and publishes their results (k_publish), captured once as a
    // graph on the engine's stream.  The engine is the serving lane's alone
    // from then on: its device-path descriptors belong to the admissions.
    bool prepare_resident(const ResidentArgs& args, std::string* error);
    // Queues one resident launch; query_async() reports its completion.
    bool launch_resident(std::string* error);

private:
    Engine() = default;
    bool init(const Params& params, const std::vector<int>& widths,
              const std::vector<int>& heights, std::string* error);
    bool two_lanes() const { return lanes_ == 2 && plan_host_.tune.streams == 2; }
    // Copies in, kernels, copies out for lane 0 or 1 (on its own stream) or,
    // with lane = -1, for the whole batch on stream_.
    hipError_t enqueue_lane(bool device_path, int lane);
    bool capture(bool device_path, int lane, hipGraphExec_t* exec);
    // (Re)captures the graphs the current launch parameters use.
    void capture_all();
    // One pipeline run on the staged frames, by graph where one was
    // captured; with `join`, everything later queued on stream_ runs after
    // it (without, stream2_'s lane ends on its own: the caller waits for
    // both streams).
    hipError_t launch(bool device_path, bool join = true);
    // Copies the device-path descriptors to the device when a frame pointer,
    // stride or active flag changed; *uploaded says whether it queued a copy.
    hipError_t upload_descriptors(bool* uploaded);
    // Waits until every frame's k_finalize has published its pass number
    // (pinned host memory, polled), or reports why it never will.
    bool wait_passes(bool split, std::string* error);
    bool set_tune_locked(const Tune& t, std::string* error);
    bool replay_locked(int reps, bool device_path, float* ms, std::string* error);
    void release();

    Params p_;
    std::vector<int> w_, h_;
    int max_det_ = 0;
    bool last_device_ = false;  // path of the last detect call
    std::vector<FrameDesc> desc_;      // host copy, gray -> engine input buffers
    std::vector<size_t> input_off_;    // per frame offset in the input region
    size_t input_bytes_ = 0;

    void* arena_ = nullptr;             // all per-frame device buffers
    uint8_t* d_input_ = nullptr;        // frames for the host path
    FrameDesc* d_desc_host_ = nullptr;  // descriptors of the host path
    FrameDesc* d_desc_dev_ = nullptr;   // descriptors of the device path
    int32_t* d_prefix_ = nullptr;
    uint64_t* d_codes_ = nullptr;
    uint32_t* d_scal_ = nullptr;        // nframes * S_COUNT
    uint32_t* d_work_ = nullptr;        // batch candidate list (header, head and tail regions)

    uint8_t* h_input_ = nullptr;        // pinned staging
    uint32_t* h_scal_ = nullptr;        // S_COUNT per frame; k_finalize writes the counts
    bd_detection* h_out_ = nullptr;     // max_det records per frame, written by k_finalize
    std::vector<FrameDesc> h_desc_dev_;  // host copy of the device-path descriptors
    // Frame pointer, stride and active flag of each device-path descriptor as
    // last bound; a launch rebinds the descriptors only when one changed.
    std::vector<const uint8_t*> up_gray_;
    std::vector<int32_t> up_stride_, up_active_;
    bool up_valid_ = false;

    Plan plan_host_{};
    Plan plan_dev_{};
    // Two-lane execution: the frames split into two contiguous runs of about
    // equal pixel count, each with its own plan (prefix tables and candidate
    // list) and its own graphs (copies in, kernels, copies out of its own
    // frames), launched on stream_ and stream2_ and joined on stream_.  The
    // lanes share no buffer.  The whole-batch plans above serve one-lane
    // launches and the per-kernel timing.
    int lanes_ = 1;
    int lane_first_[2] = {0, 0}, lane_count_[2] = {0, 0};
    Plan lane_host_[2]{};
    Plan lane_dev_[2]{};
    hipGraphExec_t lane_graph_host_[2] = {}, lane_graph_dev_[2] = {};
    hipStream_t stream2_ = nullptr;
    hipEvent_t ev_fork_ = nullptr, ev_join_ = nullptr;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev_start_ = nullptr, ev_stop_ = nullptr;
    hipEvent_t ev_stop2_ = nullptr;  // end of stream2_'s lane in a synchronous call
    hipEvent_t stage_events_[kStageCount + 1] = {};
    hipGraphExec_t graph_host_ = nullptr, graph_dev_ = nullptr;
    ResidentArgs resident_{};
    bool resident_ready_ = false;
    hipGraphExec_t graph_resident_ = nullptr;
    hipError_t enqueue_resident();

    std::vector<FrameStats> stats_;
    std::vector<uint32_t> seq_before_;   // per frame: pass number before the current call
    double clock_khz_ = 0.0;             // device wall clock (s_memrealtime) rate
    int async_frames_ = 0;
    mutable std::mutex mutex_;
};

}  // namespace bd
