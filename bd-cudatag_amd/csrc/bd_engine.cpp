// Host engine: arena layout, graph capture and the detect call.
//
// Mirrors the reference's GpuDetector::Impl (detector/src/gpu_detector.cu:
// 1784-2254): construction-time allocation of every buffer, a non-blocking
// stream, events around each frame, and a captured graph replayed per call
// (H2D copy -> kernels -> D2H of counts and detections).  Differences: one
// engine serves a list of frames (the batch API), the device-pointer path is
// a graph too (a small k_bind launch before it rebinds the frame descriptors
// when a pointer changed), and the homography is computed on the device
// inside the graph.
#include "bd_engine.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <sstream>

#include "tag_codes.inc"

namespace bd {
namespace {

std::string hip_error(const char* what, hipError_t e) {
    std::ostringstream ss;
    ss << what << ": " << hipGetErrorString(e);
    return ss.str();
}

int next_power_of_two(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace

int max_candidates_for(const Params& p, int W, int H) {
    return p.max_candidates > 0 ? std::min(std::max(p.max_candidates, 16), 1024)
                                : std::min(std::max(next_power_of_two((W * H) / 4096) * 3 / 2, 96), 1024);
}

int max_detections_for(const Params& p) { return std::min(std::max(p.max_detections, 1), 256); }

int code_count_for(int family) { return family == 0 ? 587 : 30; }

const uint64_t* host_codes_for(int family) { return family == 0 ? bd_codes_tag36h11 : bd_codes_tag16h5; }

FrameLayout layout_frame(int W, int H, int maxc, int row_ext_log2, size_t* cursor) {
    auto take = [cursor](size_t bytes) {
        const size_t at = *cursor;
        *cursor = align_up(*cursor + std::max<size_t>(bytes, 1), 256);
        return at;
    };
    const int tw = (W + kTile - 1) / kTile, th = (H + kTile - 1) / kTile;
    const int w64 = (W + 63) / 64;
    const size_t slots = static_cast<size_t>(H) * (W / 2 + 2);
    const size_t tiles = static_cast<size_t>(tw) * th;
    FrameLayout o;
    o.tmin = take(tiles);
    o.tmax = take(tiles);
    o.tmean = take(tiles);
    o.tthr = take(tiles);
    o.trlo = take(th * 4);
    o.trhi = take(th * 4);
    o.roi = take(th * 16);
    o.hist = take(320 * 4);
    o.bits = take(static_cast<size_t>(H) * w64 * 8);
    o.rowc = take(static_cast<size_t>(H) * 4);
    o.runs = take(slots * 4);
    o.label = take(slots * 4);
    o.area = take(slots * 4);
    o.bbox = take(slots * 16);
    o.keys = take(slots * 8);
    o.roots = take(slots * 4);
    o.nroots = take(static_cast<size_t>(th) * 4);
    o.items = take(slots * 8);
    o.rowi = take(static_cast<size_t>(H) * 4);
    o.rext = take((static_cast<size_t>(H) << row_ext_log2) * 8);
    o.rxfull = take(static_cast<size_t>(H) * 4);
    o.cands = take(static_cast<size_t>(maxc) * 4);
    o.quads = take(static_cast<size_t>(maxc) * 32);
    o.hquads = take(static_cast<size_t>(maxc) * 32);
    o.dec = take(static_cast<size_t>(maxc) * sizeof(bd_detection));
    return o;
}

FrameDesc bind_frame(const Params& p, int W, int H, int maxc, int max_det, int row_ext_log2, uint8_t* base,
                     const FrameLayout& o, const uint64_t* codes) {
    FrameDesc d{};
    d.active = 1;
    d.width = W;
    d.height = H;
    d.stride = W;
    d.tile_w = (W + kTile - 1) / kTile;
    d.tile_h = (H + kTile - 1) / kTile;
    d.words64 = (W + 63) / 64;
    d.row_cap = W / 2 + 2;
    d.row_ext_log2 = row_ext_log2;
    d.max_candidates = maxc;
    d.max_detections = max_det;
    d.dim = (p.family == 0 ? 6 : 4) + 2;
    d.code_count = code_count_for(p.family);
    d.max_hamming = p.max_hamming >= 0 ? std::min(p.max_hamming, 3) : (p.family == 0 ? 2 : 1);
    d.min_cluster = std::max(p.min_cluster_pixels, 1);
    d.min_tag = std::max(p.min_tag_pixels, 4);
    d.refine = p.refine_edges ? 1 : 0;
    d.min_margin = p.min_decision_margin;
    d.sharpening = p.decode_sharpening;
    d.codes = codes;
    d.tile_min = base + o.tmin;
    d.tile_max = base + o.tmax;
    d.tile_mean = base + o.tmean;
    d.tile_thresh = base + o.tthr;
    d.trow_lo = reinterpret_cast<int32_t*>(base + o.trlo);
    d.trow_hi = reinterpret_cast<int32_t*>(base + o.trhi);
    d.row_roi = reinterpret_cast<int32_t*>(base + o.roi);
    d.hist = reinterpret_cast<uint32_t*>(base + o.hist);
    d.bits = reinterpret_cast<uint64_t*>(base + o.bits);
    d.row_count = reinterpret_cast<uint32_t*>(base + o.rowc);
    d.runs = reinterpret_cast<uint32_t*>(base + o.runs);
    d.label = reinterpret_cast<uint32_t*>(base + o.label);
    d.area = reinterpret_cast<uint32_t*>(base + o.area);
    d.bbox = reinterpret_cast<uint32_t*>(base + o.bbox);
    d.cand_keys = reinterpret_cast<uint64_t*>(base + o.keys);
    d.roots = reinterpret_cast<uint32_t*>(base + o.roots);
    d.band_nroots = reinterpret_cast<uint32_t*>(base + o.nroots);
    d.items = reinterpret_cast<uint64_t*>(base + o.items);
    d.row_items = reinterpret_cast<uint32_t*>(base + o.rowi);
    d.row_ext = reinterpret_cast<uint64_t*>(base + o.rext);
    d.row_ext_full = reinterpret_cast<uint32_t*>(base + o.rxfull);
    d.candidates = reinterpret_cast<uint32_t*>(base + o.cands);
    d.quads = reinterpret_cast<float*>(base + o.quads);
    d.hull_quads = reinterpret_cast<float*>(base + o.hquads);
    d.dec = reinterpret_cast<bd_detection*>(base + o.dec);
    return d;
}

Engine* Engine::create(const Params& params, const std::vector<int>& widths,
                       const std::vector<int>& heights, std::string* error) {
    Engine* e = new (std::nothrow) Engine();
    if (!e) {
        if (error) *error = "out of memory allocating detector";
        return nullptr;
    }
    if (!e->init(params, widths, heights, error)) {
        delete e;
        return nullptr;
    }
    return e;
}

Engine::~Engine() { release(); }

void Engine::release() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (stream2_) (void)hipStreamSynchronize(stream2_);
    if (graph_host_) (void)hipGraphExecDestroy(graph_host_);
    if (graph_dev_) (void)hipGraphExecDestroy(graph_dev_);
    if (graph_resident_) (void)hipGraphExecDestroy(graph_resident_);
    graph_resident_ = nullptr;
    for (int l = 0; l < 2; ++l) {
        if (lane_graph_host_[l]) (void)hipGraphExecDestroy(lane_graph_host_[l]);
        if (lane_graph_dev_[l]) (void)hipGraphExecDestroy(lane_graph_dev_[l]);
        lane_graph_host_[l] = lane_graph_dev_[l] = nullptr;
    }
    if (ev_start_) (void)hipEventDestroy(ev_start_);
    if (ev_stop_) (void)hipEventDestroy(ev_stop_);
    for (hipEvent_t& e : stage_events_) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    if (ev_fork_) (void)hipEventDestroy(ev_fork_);
    if (ev_join_) (void)hipEventDestroy(ev_join_);
    if (ev_stop2_) (void)hipEventDestroy(ev_stop2_);
    if (stream2_) (void)hipStreamDestroy(stream2_);
    if (stream_) (void)hipStreamDestroy(stream_);
    if (arena_) (void)hipFree(arena_);
    if (h_input_) (void)hipHostFree(h_input_);
    if (h_scal_) (void)hipHostFree(h_scal_);
    if (h_out_) (void)hipHostFree(h_out_);
    graph_host_ = graph_dev_ = nullptr;
    ev_start_ = ev_stop_ = nullptr;
    ev_fork_ = ev_join_ = ev_stop2_ = nullptr;
    stream2_ = nullptr;
    stream_ = nullptr;
    arena_ = nullptr;
    h_input_ = nullptr;
    h_scal_ = nullptr;
    h_out_ = nullptr;
    h_desc_dev_.clear();
}

bool Engine::init(const Params& params, const std::vector<int>& widths,
                  const std::vector<int>& heights, std::string* error) {
    p_ = params;
    w_ = widths;
    h_ = heights;
    const int nf = static_cast<int>(w_.size());
    // Clamps of GpuDetector::Impl::initialize (gpu_detector.cu:1906-1943).
    max_det_ = max_detections_for(p_);
    const uint64_t* codes = host_codes_for(p_.family);
    const int code_count = code_count_for(p_.family);

    // Arena layout: every frame's buffers, then the shared scalar and output
    // blocks (contiguous so one copy returns them for all frames).
    std::vector<FrameLayout> off(nf);
    std::vector<int> maxc(nf);
    size_t cur = 0;
    auto take = [&cur](size_t bytes) {
        const size_t at = cur;
        cur = align_up(cur + std::max<size_t>(bytes, 1), 256);
        return at;
    };
    input_off_.assign(nf, 0);
    input_bytes_ = 0;
    for (int f = 0; f < nf; ++f) {
        const int W = w_[f], H = h_[f];
        maxc[f] = max_candidates_for(p_, W, H);
        off[f] = layout_frame(W, H, maxc[f], row_ext_log2_for_width(W), &cur);
        input_off_[f] = input_bytes_;
        input_bytes_ = align_up(input_bytes_ + static_cast<size_t>(W) * H, 256);
    }
    size_t work_cap = 0;
    for (int f = 0; f < nf; ++f) work_cap += static_cast<size_t>(maxc[f]);
    const size_t work_words = kWorkItems + static_cast<size_t>(kHeadSlots) * nf + work_cap;
    const size_t o_work = take(work_words * 4);
    // Lanes: frames [0, split) and [split, nf), split where the first lane
    // holds kLaneLeadPx (at most 1/16 of the batch) more pixels than the
    // second, or as near as the frame sizes allow: the second lane's graph
    // starts about 20 us later (the host submits the first graph meanwhile),
    // and a synchronous call waits for the later of the two.  Measured on 64
    // frames of 1280x800 (three engines per split, calls interleaved): 34/30
    // frames 0.394 ms per call, 33/31 0.396, 35/29 0.398, 32/32 0.415.
    int split = 0;
    if (nf >= 2) {
        int64_t total_px = 0, run = 0;
        uint64_t best = ~0ull;
        for (int f = 0; f < nf; ++f) total_px += static_cast<int64_t>(w_[f]) * h_[f];
        for (int f = 1; f < nf; ++f) {
            run += static_cast<int64_t>(w_[f - 1]) * h_[f - 1];
            const int64_t diff = run - (total_px - run) - std::min(kLaneLeadPx, total_px / 16);
            const uint64_t mag = static_cast<uint64_t>(diff < 0 ? -diff : diff);
            if (mag < best) {
                best = mag;
                split = f;
            }
        }
    }
    lanes_ = split > 0 ? 2 : 1;
    lane_first_[0] = 0;
    lane_first_[1] = split;
    lane_count_[0] = split;
    lane_count_[1] = nf - split;
    const int* lane_first = lane_first_;
    const int* lane_count = lane_count_;
    size_t lane_cap[2] = {0, 0}, o_lane_work[2] = {0, 0}, o_lane_prefix[2] = {0, 0};
    for (int l = 0; l < 2 && lanes_ == 2; ++l) {
        for (int f = lane_first[l]; f < lane_first[l] + lane_count[l]; ++f) lane_cap[l] += static_cast<size_t>(maxc[f]);
        o_lane_work[l] = take((kWorkItems + static_cast<size_t>(kHeadSlots) * lane_count[l] + lane_cap[l]) * 4);
        o_lane_prefix[l] = take(static_cast<size_t>(SH_COUNT) * (lane_count[l] + 1) * 4);
    }
    const size_t o_scal = take(static_cast<size_t>(nf) * S_COUNT * 4);
    const size_t o_desc_h = take(static_cast<size_t>(nf) * sizeof(FrameDesc));
    const size_t o_desc_d = take(static_cast<size_t>(nf) * sizeof(FrameDesc));
    const size_t o_prefix = take(static_cast<size_t>(SH_COUNT) * (nf + 1) * 4);
    const size_t o_codes = take(static_cast<size_t>(code_count) * 8);
    const size_t o_input = take(input_bytes_);
    const size_t total = cur;

    hipError_t st = hipMalloc(&arena_, total);
    if (st != hipSuccess) {
        if (error) *error = hip_error("hipMalloc", st);
        arena_ = nullptr;
        return false;
    }
    uint8_t* base = static_cast<uint8_t*>(arena_);
    d_scal_ = reinterpret_cast<uint32_t*>(base + o_scal);
    d_work_ = reinterpret_cast<uint32_t*>(base + o_work);
    d_desc_host_ = reinterpret_cast<FrameDesc*>(base + o_desc_h);
    d_desc_dev_ = reinterpret_cast<FrameDesc*>(base + o_desc_d);
    d_prefix_ = reinterpret_cast<int32_t*>(base + o_prefix);
    d_codes_ = reinterpret_cast<uint64_t*>(base + o_codes);
    d_input_ = base + o_input;

    st = hipHostMalloc(reinterpret_cast<void**>(&h_input_), std::max<size_t>(input_bytes_, 1), hipHostMallocPortable);
    if (st == hipSuccess) st = hipHostMalloc(reinterpret_cast<void**>(&h_scal_), static_cast<size_t>(nf) * S_COUNT * 4, hipHostMallocPortable);
    if (st == hipSuccess) st = hipHostMalloc(reinterpret_cast<void**>(&h_out_), static_cast<size_t>(nf) * max_det_ * sizeof(bd_detection), hipHostMallocPortable);
    // k_finalize writes each frame's detection records and counts straight
    // into the pinned host arrays (only the records found cross PCIe; copy
    // nodes cost a blit kernel each, ~4 us at the end of every lane).
    bd_detection* out_dev = nullptr;
    uint32_t* scal_dev = nullptr;
    if (st == hipSuccess) st = hipHostGetDevicePointer(reinterpret_cast<void**>(&out_dev), h_out_, 0);
    if (st == hipSuccess) st = hipHostGetDevicePointer(reinterpret_cast<void**>(&scal_dev), h_scal_, 0);
    if (st != hipSuccess) {
        if (error) *error = hip_error("hipHostMalloc", st);
        return false;
    }
    // Pass numbers start at 0 on both sides (k_finalize publishes 1 first).
    std::memset(h_scal_, 0, static_cast<size_t>(nf) * S_COUNT * 4);

    desc_.assign(nf, FrameDesc{});
    std::vector<int32_t> prefix(static_cast<size_t>(SH_COUNT) * (nf + 1), 0);
    for (int f = 0; f < nf; ++f) {
        FrameDesc& d = desc_[f];
        d = bind_frame(p_, w_[f], h_[f], maxc[f], max_det_, row_ext_log2_for_width(w_[f]), base, off[f], d_codes_);
        d.gray = d_input_ + input_off_[f];
        d.scal = d_scal_ + static_cast<size_t>(f) * S_COUNT;
        d.out = out_dev + static_cast<size_t>(f) * max_det_;
        d.host_counts = scal_dev + static_cast<size_t>(f) * S_COUNT + S_COUNTS;
        prefix[SH_TILE_ROW * (nf + 1) + f + 1] = prefix[SH_TILE_ROW * (nf + 1) + f] + d.tile_h;
        prefix[SH_FRAME * (nf + 1) + f + 1] = f + 1;
        prefix[SH_TSTAT * (nf + 1) + f + 1] = prefix[SH_TSTAT * (nf + 1) + f] + tile_stat_blocks(d.tile_w, d.tile_h);
        for (int c = 0; c < kWidthClasses; ++c) {
            const bool mine = d.row_ext_log2 == kRowExtLog2Min + c;
            int32_t* pc = prefix.data() + static_cast<size_t>(SH_CLASS + c) * (nf + 1);
            int32_t* pp = prefix.data() + static_cast<size_t>(SH_CLASS_PAIR + c) * (nf + 1);
            pc[f + 1] = pc[f] + (mine ? d.tile_h : 0);
            pp[f + 1] = pp[f] + (mine ? (d.tile_h + 1) / 2 : 0);
        }
    }
    h_desc_dev_ = desc_;
    // Per-lane prefix tables: the whole-batch table restricted to the lane's
    // frames and rebased to zero.
    std::vector<int32_t> lane_prefix[2];
    for (int l = 0; l < 2 && lanes_ == 2; ++l) {
        const int n = lane_count[l], first = lane_first[l];
        lane_prefix[l].assign(static_cast<size_t>(SH_COUNT) * (n + 1), 0);
        for (int k = 0; k < SH_COUNT; ++k) {
            const int32_t* src = prefix.data() + static_cast<size_t>(k) * (nf + 1);
            for (int i = 0; i <= n; ++i) lane_prefix[l][static_cast<size_t>(k) * (n + 1) + i] = src[first + i] - src[first];
        }
    }

    st = hipMemcpy(d_desc_host_, desc_.data(), desc_.size() * sizeof(FrameDesc), hipMemcpyHostToDevice);
    if (st == hipSuccess) st = hipMemcpy(d_desc_dev_, desc_.data(), desc_.size() * sizeof(FrameDesc), hipMemcpyHostToDevice);
    if (st == hipSuccess) st = hipMemcpy(d_prefix_, prefix.data(), prefix.size() * 4, hipMemcpyHostToDevice);
    for (int l = 0; l < 2 && lanes_ == 2 && st == hipSuccess; ++l) {
        st = hipMemcpy(base + o_lane_prefix[l], lane_prefix[l].data(), lane_prefix[l].size() * 4, hipMemcpyHostToDevice);
        if (st == hipSuccess)
            st = hipMemset(base + o_lane_work[l], 0,
                           (kWorkItems + static_cast<size_t>(kHeadSlots) * lane_count[l] + lane_cap[l]) * 4);
    }
    if (st == hipSuccess) st = hipMemcpy(d_codes_, codes, static_cast<size_t>(code_count) * 8, hipMemcpyHostToDevice);
    // Accumulators start cleared; afterwards each consumer kernel clears them.
    for (int f = 0; f < nf && st == hipSuccess; ++f) {
        // Each tile row's active extent starts empty: (tile_w, -1).
        const std::vector<int32_t> lo(static_cast<size_t>(desc_[f].tile_h), desc_[f].tile_w);
        const std::vector<int32_t> hi(static_cast<size_t>(desc_[f].tile_h), -1);
        st = hipMemcpy(desc_[f].trow_lo, lo.data(), lo.size() * 4, hipMemcpyHostToDevice);
        if (st == hipSuccess) st = hipMemcpy(desc_[f].trow_hi, hi.data(), hi.size() * 4, hipMemcpyHostToDevice);
        if (st == hipSuccess) st = hipMemset(desc_[f].hist, 0, 320 * 4);
        if (st == hipSuccess) st = hipMemset(desc_[f].dec, 0xff, static_cast<size_t>(maxc[f]) * sizeof(bd_detection));
    }
    if (st == hipSuccess) {
        std::vector<uint32_t> scal(static_cast<size_t>(nf) * S_COUNT, 0);
        for (int f = 0; f < nf; ++f) scal[static_cast<size_t>(f) * S_COUNT + S_TOP] = 0xffffffffu;
        st = hipMemcpy(d_scal_, scal.data(), scal.size() * 4, hipMemcpyHostToDevice);
    }
    if (st == hipSuccess) st = hipMemset(d_input_, 0, std::max<size_t>(input_bytes_, 1));
    if (st == hipSuccess) st = hipMemset(d_work_, 0, work_words * 4);
    if (st != hipSuccess) {
        if (error) *error = hip_error("detector setup copy", st);
        return false;
    }

    plan_host_.tune = default_tune();
    plan_host_.wide_blocks = binarize_wide_blocks();
    plan_host_.frames = d_desc_host_;
    plan_host_.prefix = d_prefix_;
    plan_host_.nframes = nf;
    plan_host_.work = d_work_;
    plan_host_.work_cap = static_cast<int32_t>(work_cap);
    // Each width class's LDS band capacity, from its widest frame in the plan.
    const auto set_band_runs = [this](Plan* plan, int first, int n) {
        for (int c = 0; c < kWidthClasses; ++c) {
            int widest = 0;
            for (int f = first; f < first + n; ++f)
                if (row_ext_log2_for_width(w_[f]) == kRowExtLog2Min + c) widest = std::max(widest, w_[f]);
            plan->class_band_runs[c] = widest > 0 ? band_runs_for_width(widest) : 0;
            plan->class_merge_stage[c] = widest > 0 ? merge_stage_for_width(widest) : 0;
        }
    };
    set_band_runs(&plan_host_, 0, nf);
    plan_set_blocks(&plan_host_, prefix.data(), nf);
    plan_dev_ = plan_host_;
    plan_dev_.frames = d_desc_dev_;
    for (int l = 0; l < 2 && lanes_ == 2; ++l) {
        const int n = lane_count[l];
        Plan& p = lane_host_[l];
        p = plan_host_;
        p.frames = d_desc_host_ + lane_first[l];
        p.prefix = reinterpret_cast<int32_t*>(base + o_lane_prefix[l]);
        p.work = reinterpret_cast<uint32_t*>(base + o_lane_work[l]);
        p.nframes = n;
        p.work_cap = static_cast<int32_t>(lane_cap[l]);
        set_band_runs(&p, lane_first[l], n);
        plan_set_blocks(&p, lane_prefix[l].data(), n);
        lane_dev_[l] = p;
        lane_dev_[l].frames = d_desc_dev_ + lane_first[l];
    }

    // Every stream at the default priority.
    st = hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking);
    if (st == hipSuccess && lanes_ == 2) st = hipStreamCreateWithFlags(&stream2_, hipStreamNonBlocking);
    if (st == hipSuccess && lanes_ == 2) st = hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming);
    if (st == hipSuccess && lanes_ == 2) st = hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming);
    if (st == hipSuccess && lanes_ == 2) st = hipEventCreate(&ev_stop2_);
    // Device wall-clock rate (kHz) for the pass stamps.
    int dev = 0, khz = 0;
    if (st == hipSuccess && hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess)
        clock_khz_ = static_cast<double>(khz);
    if (st == hipSuccess) st = hipEventCreate(&ev_start_);
    if (st == hipSuccess) st = hipEventCreate(&ev_stop_);
    if (st == hipSuccess) st = hipDeviceSynchronize();
    if (st != hipSuccess) {
        if (error) *error = hip_error("stream/event setup", st);
        return false;
    }
    stats_.assign(nf, FrameStats{});
    seq_before_.assign(nf, 0u);

    capture_all();
    return true;
}

hipError_t Engine::enqueue_lane(bool device_path, int lane) {
    const int nf = frame_count();
    const int first = lane < 0 ? 0 : lane_first_[lane];
    const int n = lane < 0 ? nf : lane_count_[lane];
    const Plan& plan = lane < 0 ? (device_path ? plan_dev_ : plan_host_) : (device_path ? lane_dev_[lane] : lane_host_[lane]);
    hipStream_t s = lane == 1 ? stream2_ : stream_;
    hipError_t st;
    if (device_path) {
        st = hipSuccess;  // descriptors: upload_descriptors() before the launch
    } else {
        const size_t begin = input_off_[first];
        const size_t end = first + n < nf ? input_off_[first + n] : input_bytes_;
        st = hipMemcpyAsync(d_input_ + begin, h_input_ + begin, end - begin, hipMemcpyHostToDevice, s);
    }
    if (st == hipSuccess) st = launch_pipeline(plan, s);
    return st;
}

// The device path's descriptors change only in frame pointer, stride and
// active flag; they are rebound (by k_bind launches carrying them as kernel
// arguments, before the lanes fork) only when one of these differs from the
// last binding, so repeated calls on the same buffers launch the graphs
// alone.
hipError_t Engine::upload_descriptors(bool* uploaded) {
    *uploaded = false;
    const int nf = frame_count();
    bool changed = !up_valid_;
    for (int f = 0; f < nf && !changed; ++f) {
        const FrameDesc& d = h_desc_dev_[f];
        changed = d.gray != up_gray_[f] || d.stride != up_stride_[f] || d.active != up_active_[f];
    }
    if (!changed) return hipSuccess;
    for (int first = 0; first < nf; first += kBindFrames) {
        BindArgs args = {};
        args.first = first;
        args.count = std::min(kBindFrames, nf - first);
        for (int k = 0; k < args.count; ++k) {
            const FrameDesc& d = h_desc_dev_[first + k];
            args.gray[k] = d.gray;
            args.stride[k] = d.stride;
            args.active[k] = d.active;
        }
        const hipError_t st = launch_bind(d_desc_dev_, args, stream_);
        if (st != hipSuccess) return st;
    }
    up_gray_.resize(nf);
    up_stride_.resize(nf);
    up_active_.resize(nf);
    for (int f = 0; f < nf; ++f) {
        up_gray_[f] = h_desc_dev_[f].gray;
        up_stride_[f] = h_desc_dev_[f].stride;
        up_active_[f] = h_desc_dev_[f].active;
    }
    up_valid_ = true;
    *uploaded = true;
    return hipSuccess;
}

hipError_t Engine::launch(bool device_path, bool join) {
    bool uploaded = false;
    if (device_path) {
        const hipError_t st = upload_descriptors(&uploaded);
        if (st != hipSuccess) return st;
    }
    if (!two_lanes()) {
        hipGraphExec_t g = device_path ? graph_dev_ : graph_host_;
        return g ? hipGraphLaunch(g, stream_) : enqueue_lane(device_path, -1);
    }
    // The second lane starts after the descriptor upload when this launch
    // queued one on stream_ (a cross-stream wait costs ~35 us of host time,
    // so it is skipped otherwise: every earlier launch of this engine has
    // completed on both streams, or is one of back-to-back replays of the
    // same frames, whose lanes share no buffer); stream_ waits for the
    // second lane before anything queued later.
    hipError_t st = hipSuccess;
    if (uploaded) {
        st = hipEventRecord(ev_fork_, stream_);
        if (st == hipSuccess) st = hipStreamWaitEvent(stream2_, ev_fork_, 0);
    }
    for (int l = 0; l < 2 && st == hipSuccess; ++l) {
        hipGraphExec_t g = device_path ? lane_graph_dev_[l] : lane_graph_host_[l];
        st = g ? hipGraphLaunch(g, l == 1 ? stream2_ : stream_) : enqueue_lane(device_path, l);
    }
    if (!join) return st;
    if (st == hipSuccess) st = hipEventRecord(ev_join_, stream2_);
    if (st == hipSuccess) st = hipStreamWaitEvent(stream_, ev_join_, 0);
    return st;
}

bool Engine::capture(bool device_path, int lane, hipGraphExec_t* exec) {
    *exec = nullptr;
    hipStream_t s = lane == 1 ? stream2_ : stream_;
    hipError_t st = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (st != hipSuccess) return false;
    const hipError_t queued = enqueue_lane(device_path, lane);
    hipGraph_t graph = nullptr;
    st = hipStreamEndCapture(s, &graph);
    if (queued != hipSuccess || st != hipSuccess || !graph) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        return false;
    }
    st = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (st != hipSuccess) {
        *exec = nullptr;
        return false;
    }
    return true;
}

void Engine::capture_all() {
    if (graph_host_) (void)hipGraphExecDestroy(graph_host_);
    if (graph_dev_) (void)hipGraphExecDestroy(graph_dev_);
    graph_host_ = graph_dev_ = nullptr;
    for (int l = 0; l < 2; ++l) {
        if (lane_graph_host_[l]) (void)hipGraphExecDestroy(lane_graph_host_[l]);
        if (lane_graph_dev_[l]) (void)hipGraphExecDestroy(lane_graph_dev_[l]);
        lane_graph_host_[l] = lane_graph_dev_[l] = nullptr;
    }
    // Graph capture is an optimisation: a graph that fails to capture leaves
    // direct launches in its place, as the reference does
    // (gpu_detector.cu:2051-2056).
    if (!p_.use_graph) return;
    if (two_lanes()) {
        for (int l = 0; l < 2; ++l) {
            (void)capture(false, l, &lane_graph_host_[l]);
            (void)capture(true, l, &lane_graph_dev_[l]);
        }
    } else {
        (void)capture(false, -1, &graph_host_);
        (void)capture(true, -1, &graph_dev_);
    }
}

int Engine::detect(const bd_frame* frames, bool device, bd_detection* out, int out_stride,
                   int32_t* counts, std::string* error) {
    const std::lock_guard<std::mutex> lock(mutex_);
    const int nf = frame_count();
    const auto start = std::chrono::steady_clock::now();
    last_device_ = device;
    if (device) {
        for (int f = 0; f < nf; ++f) {
            h_desc_dev_[f].gray = frames[f].gray;
            h_desc_dev_[f].stride = frames[f].stride;
            h_desc_dev_[f].active = 1;
        }
    } else {
        // Repack into pinned staging (gpu_detector.cu:2169-2176).
        for (int f = 0; f < nf; ++f) {
            const int W = w_[f], H = h_[f], stride = frames[f].stride;
            uint8_t* dst = h_input_ + input_off_[f];
            if (stride == W) {
                std::memcpy(dst, frames[f].gray, static_cast<size_t>(W) * H);
            } else {
                for (int y = 0; y < H; ++y)
                    std::memcpy(dst + static_cast<size_t>(y) * W, frames[f].gray + static_cast<size_t>(y) * stride, W);
            }
        }
    }
    for (int f = 0; f < nf; ++f)
        seq_before_[f] = __atomic_load_n(h_scal_ + static_cast<size_t>(f) * S_COUNT + S_COUNTS + kHostSeq,
                                         __ATOMIC_ACQUIRE);
    // Two lanes end on their own streams (a cross-stream join cost ~20 us
    // per call).  The host returns once every frame's k_finalize has
    // published its pass number into pinned memory (an event's completion
    // reaches a waiting host ~10 us after the last kernel); nothing is
    // queued on either stream before both lanes' kernels have ended.  The
    // pass is timed by the device clock stamps k_tile_stats and k_finalize
    // leave beside the counts, so the call records no event.
    const bool split_wait = two_lanes();
    const hipError_t st = launch(device, !split_wait);
    if (st != hipSuccess) {
        if (error) *error = hip_error("pipeline launch", st);
        return -1;
    }
    if (!wait_passes(split_wait, error)) return -1;
    const auto end = std::chrono::steady_clock::now();
    const float total_ms = std::chrono::duration<float, std::milli>(end - start).count();
    uint64_t t_first = ~0ull, t_last = 0;
    for (int f = 0; f < nf; ++f) {
        const uint32_t* c = h_scal_ + static_cast<size_t>(f) * S_COUNT + S_COUNTS;
        t_first = std::min(t_first, *reinterpret_cast<const uint64_t*>(c + kHostStart));
        t_last = std::max(t_last, *reinterpret_cast<const uint64_t*>(c + kHostEnd));
    }
    const float gpu_ms = t_last > t_first && clock_khz_ > 0
                             ? static_cast<float>(static_cast<double>(t_last - t_first) / clock_khz_)
                             : 0.0f;

    // finish() of gpu_detector.cu:2100-2127, per frame.
    int written_total = 0;
    for (int f = 0; f < nf; ++f) {
        const uint32_t* c = h_scal_ + static_cast<size_t>(f) * S_COUNT + S_COUNTS;
        const int found = std::min(static_cast<int>(c[1]), max_det_);
        const int written = std::min(found, std::max(out_stride, 0));
        if (written > 0 && out) {
            std::memcpy(out + static_cast<size_t>(f) * out_stride, h_out_ + static_cast<size_t>(f) * max_det_,
                        static_cast<size_t>(written) * sizeof(bd_detection));
        }
        if (counts) counts[f] = found;
        written_total += written;
        FrameStats& s = stats_[f];
        s.candidates = std::min(static_cast<int>(c[0]), desc_[f].max_candidates);
        s.runs = static_cast<int>(c[2]);
        s.detections = found;
        s.quads = static_cast<int>(c[3]);
        s.roi_rows = std::min(static_cast<int>(c[4]) * kTile, h_[f]);
        s.last_gpu_ms = gpu_ms;
        s.last_total_ms = total_ms;
        s.average_gpu_ms = s.frames == 0 ? gpu_ms : 0.9f * s.average_gpu_ms + 0.1f * gpu_ms;
        ++s.frames;
    }
    return written_total;
}

bool Engine::wait_passes(bool split, std::string* error) {
    const int nf = frame_count();
    const auto seq_of = [this](int f) {
        return __atomic_load_n(h_scal_ + static_cast<size_t>(f) * S_COUNT + S_COUNTS + kHostSeq, __ATOMIC_ACQUIRE);
    };
    int f = 0;
    for (uint32_t spin = 1;; ++spin) {
        while (f < nf && seq_of(f) != seq_before_[f]) ++f;
        if (f == nf) return true;
        // Every so often the streams are asked too: a pipeline that failed
        // never publishes, and its error surfaces here.
        if ((spin & 1023u) == 0) {
            hipError_t q = hipStreamQuery(stream_);
            if (q == hipSuccess && split) q = hipStreamQuery(stream2_);
            if (q == hipSuccess) {
                while (f < nf && seq_of(f) != seq_before_[f]) ++f;
                if (f == nf) return true;
                if (error) *error = "pipeline ended without publishing frame " + std::to_string(f);
                return false;
            }
            if (q != hipErrorNotReady) {
                if (error) *error = hip_error("hipStreamQuery", q);
                return false;
            }
        }
        __builtin_ia32_pause();
    }
}

void Engine::stats(int index, bd_stats* out) const {
    const std::lock_guard<std::mutex> lock(mutex_);
    const FrameStats& s = stats_[index];
    out->frames = s.frames;
    out->last_gpu_ms = s.last_gpu_ms;
    out->average_gpu_ms = s.average_gpu_ms;
    out->last_total_ms = s.last_total_ms;
    out->runs = s.runs;
    out->candidates = s.candidates;
    out->detections = s.detections;
    out->quads = s.quads;
    out->roi_rows = s.roi_rows;
}

bool Engine::frame_info(int index, int32_t out[8]) const {
    if (index < 0 || index >= frame_count()) return false;
    const FrameDesc& d = desc_[index];
    out[0] = d.width;
    out[1] = d.height;
    out[2] = d.tile_w;
    out[3] = d.tile_h;
    out[4] = d.words64;
    out[5] = d.row_cap;
    out[6] = d.max_candidates;
    out[7] = d.max_detections;
    return true;
}

int64_t Engine::debug_copy(int index, int stage, void* dst, int64_t capacity, std::string* error) {
    const std::lock_guard<std::mutex> lock(mutex_);
    if (index < 0 || index >= frame_count()) {
        if (error) *error = "frame index out of range";
        return -1;
    }
    const FrameDesc& d = desc_[index];
    const size_t tiles = static_cast<size_t>(d.tile_w) * d.tile_h;
    const size_t slots = static_cast<size_t>(d.height) * d.row_cap;
    const void* src = nullptr;
    size_t bytes = 0;
    switch (stage) {
        case BD_STAGE_TILE_MIN: src = d.tile_min; bytes = tiles; break;
        case BD_STAGE_TILE_MAX: src = d.tile_max; bytes = tiles; break;
        case BD_STAGE_TILE_MEAN: src = d.tile_mean; bytes = tiles; break;
        case BD_STAGE_TILE_THRESH: src = d.tile_thresh; bytes = tiles; break;
        case BD_STAGE_HIST: src = d.hist; bytes = 320 * 4; break;
        case BD_STAGE_SCALARS: src = d.scal; bytes = S_COUNT * 4; break;
        case BD_STAGE_ROW_ROI: src = d.row_roi; bytes = static_cast<size_t>(d.tile_h) * 16; break;
        case BD_STAGE_BITS: src = d.bits; bytes = static_cast<size_t>(d.height) * d.words64 * 8; break;
        case BD_STAGE_ROW_COUNT: src = d.row_count; bytes = static_cast<size_t>(d.height) * 4; break;
        case BD_STAGE_RUNS: src = d.runs; bytes = slots * 4; break;
        case BD_STAGE_LABELS: src = d.label; bytes = slots * 4; break;
        case BD_STAGE_AREA: src = d.area; bytes = slots * 4; break;
        case BD_STAGE_BBOX: src = d.bbox; bytes = slots * 16; break;
        case BD_STAGE_CANDIDATES: src = d.candidates; bytes = static_cast<size_t>(d.max_candidates) * 4; break;
        case BD_STAGE_QUADS: src = d.quads; bytes = static_cast<size_t>(d.max_candidates) * 32; break;
        case BD_STAGE_HULL_QUADS: src = d.hull_quads; bytes = static_cast<size_t>(d.max_candidates) * 32; break;
        case BD_STAGE_DECODE: src = d.dec; bytes = static_cast<size_t>(d.max_candidates) * sizeof(bd_detection); break;
        default:
            if (error) *error = "unknown stage";
            return -1;
    }
    if (dst && capacity > 0) {
        const size_t n = std::min(bytes, static_cast<size_t>(capacity));
        const hipError_t st = hipMemcpy(dst, src, n, hipMemcpyDeviceToHost);
        if (st != hipSuccess) {
            if (error) *error = hip_error("hipMemcpy(debug)", st);
            return -2;
        }
    }
    return static_cast<int64_t>(bytes);
}

bool Engine::launch_async(const uint8_t* const* gray, const int32_t* stride, int n,
                          std::string* error) {
    const std::lock_guard<std::mutex> lock(mutex_);
    const int nf = frame_count();
    if (n < 1 || n > nf) {
        if (error) *error = "invalid frame count";
        return false;
    }
    for (int f = 0; f < nf; ++f) {
        h_desc_dev_[f].active = f < n ? 1 : 0;
        if (f < n) {
            h_desc_dev_[f].gray = gray[f];
            h_desc_dev_[f].stride = stride[f];
        }
    }
    hipError_t st = launch(true);
    if (st == hipSuccess) st = hipEventRecord(ev_stop_, stream_);
    if (st != hipSuccess) {
        if (error) *error = hip_error("async launch", st);
        return false;
    }
    async_frames_ = n;
    return true;
}

int Engine::query_async() {
    const hipError_t st = hipEventQuery(ev_stop_);
    if (st == hipSuccess) return 1;
    if (st == hipErrorNotReady) return 0;
    return -1;
}

int Engine::collect_async(int index, bd_detection* out, int capacity) const {
    const uint32_t* c = h_scal_ + static_cast<size_t>(index) * S_COUNT + S_COUNTS;
    const int found = std::min(static_cast<int>(c[1]), max_det_);
    const int written = std::min(found, std::max(capacity, 0));
    if (written > 0 && out) {
        std::memcpy(out, h_out_ + static_cast<size_t>(index) * max_det_,
                    static_cast<size_t>(written) * sizeof(bd_detection));
    }
    return found;
}

hipError_t Engine::enqueue_resident() {
    hipError_t st = launch_admit(plan_dev_, resident_, stream_);
    if (st == hipSuccess) st = launch_pipeline(plan_dev_, stream_);
    if (st == hipSuccess) st = launch_publish(plan_dev_, resident_, stream_);
    return st;
}

bool Engine::prepare_resident(const ResidentArgs& args, std::string* error) {
    const std::lock_guard<std::mutex> lock(mutex_);
    resident_ = args;
    resident_ready_ = true;
    if (graph_resident_) (void)hipGraphExecDestroy(graph_resident_);
    graph_resident_ = nullptr;
    if (!p_.use_graph) return true;
    hipError_t st = hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal);
    if (st != hipSuccess) {
        if (error) *error = hip_error("hipStreamBeginCapture", st);
        return false;
    }
    const hipError_t queued = enqueue_resident();
    hipGraph_t graph = nullptr;
    st = hipStreamEndCapture(stream_, &graph);
    if (queued == hipSuccess && st == hipSuccess && graph) st = hipGraphInstantiate(&graph_resident_, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (queued != hipSuccess || st != hipSuccess) {
        graph_resident_ = nullptr;
        if (error) *error = hip_error("resident graph capture", queued != hipSuccess ? queued : st);
        return false;
    }
    return true;
}

bool Engine::launch_resident(std::string* error) {
    const std::lock_guard<std::mutex> lock(mutex_);
    if (!resident_ready_) {
        if (error) *error = "resident launch not prepared";
        return false;
    }
    hipError_t st = graph_resident_ ? hipGraphLaunch(graph_resident_, stream_) : enqueue_resident();
    if (st == hipSuccess) st = hipEventRecord(ev_stop_, stream_);
    if (st != hipSuccess) {
        if (error) *error = hip_error("resident launch", st);
        return false;
    }
    return true;
}

bool Engine::time_stages(int reps, float* ms, std::string* error) {
    const std::lock_guard<std::mutex> lock(mutex_);
    if (reps < 1 || !ms) {
        if (error) *error = "invalid repetition count";
        return false;
    }
    if (!stage_events_[0]) {
        for (int i = 0; i <= kStageCount; ++i) {
            const hipError_t st = hipEventCreate(&stage_events_[i]);
            if (st != hipSuccess) {
                if (error) *error = hip_error("hipEventCreate", st);
                return false;
            }
        }
    }
    for (int i = 0; i <= kStageCount; ++i) ms[i] = 0.0f;
    for (int r = 0; r < reps; ++r) {
        for (int stage = 0; stage < kStageCount; ++stage) {
            hipError_t st = hipEventRecord(stage_events_[stage], stream_);
            if (st == hipSuccess) st = launch_stage(plan_host_, stage, stream_);
            if (st != hipSuccess) {
                if (error) *error = hip_error("kernel launch", st);
                return false;
            }
        }
        hipError_t st = hipEventRecord(stage_events_[kStageCount], stream_);
        if (st == hipSuccess) st = hipEventSynchronize(stage_events_[kStageCount]);
        if (st != hipSuccess) {
            if (error) *error = hip_error("hipEventSynchronize", st);
            return false;
        }
        for (int stage = 0; stage < kStageCount; ++stage) {
            float t = 0.0f;
            (void)hipEventElapsedTime(&t, stage_events_[stage], stage_events_[stage + 1]);
            ms[stage] += t;
        }
        float t = 0.0f;
        (void)hipEventElapsedTime(&t, stage_events_[0], stage_events_[kStageCount]);
        ms[kStageCount] += t;
    }
    for (int i = 0; i <= kStageCount; ++i) ms[i] /= static_cast<float>(reps);
    return true;
}

bool Engine::replay(int reps, bool device_path, float* ms, std::string* error) {
    const std::lock_guard<std::mutex> lock(mutex_);
    return replay_locked(reps, device_path, ms, error);
}

bool Engine::replay_locked(int reps, bool device_path, float* ms, std::string* error) {
    if (reps < 1) {
        if (error) *error = "invalid repetition count";
        return false;
    }
    hipError_t st = hipEventRecord(ev_start_, stream_);
    for (int r = 0; r < reps && st == hipSuccess; ++r) st = launch(device_path);
    if (st == hipSuccess) st = hipEventRecord(ev_stop_, stream_);
    if (st == hipSuccess) st = hipEventSynchronize(ev_stop_);
    if (st != hipSuccess) {
        if (error) *error = hip_error("hipEventSynchronize", st);
        return false;
    }
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, ev_start_, ev_stop_);
    if (ms) *ms = t;
    return true;
}

Tune Engine::tune() const {
    Tune t = plan_host_.tune;
    if (lanes_ < 2) t.streams = 1;
    return t;
}

bool Engine::set_tune(const Tune& t, std::string* error) {
    const std::lock_guard<std::mutex> lock(mutex_);
    return set_tune_locked(t, error);
}

bool Engine::set_tune_locked(const Tune& t, std::string* error) {
    if (t.fit_grid < 64 || t.fit_grid > 16384 || t.decode_grid < 64 || t.decode_grid > 16384 ||
        t.head_ranks < 0 || t.head_ranks > kHeadRanksMax || t.streams < 1 || t.streams > 2 ||
        (t.decode_waves != 4 && t.decode_waves != 8 && t.decode_waves != 16) ||
        (t.chunk_px != 64 && t.chunk_px != 128) || t.head_area < 1 || (t.band_rows != 8 && t.band_rows != 16)) {
        if (error) *error = "launch parameters out of range";
        return false;
    }
    hipError_t st = hipStreamSynchronize(stream_);
    if (st == hipSuccess && stream2_) st = hipStreamSynchronize(stream2_);
    if (st != hipSuccess) {
        if (error) *error = hip_error("hipStreamSynchronize", st);
        return false;
    }
    plan_host_.tune = t;
    plan_dev_.tune = t;
    for (int l = 0; l < 2; ++l) {
        lane_host_[l].tune = t;
        lane_dev_[l].tune = t;
    }
    capture_all();
    return true;
}

bool Engine::set_capture(bool on, std::string* error) {
    const std::lock_guard<std::mutex> lock(mutex_);
    if (resident_ready_) {
        if (error) *error = "stage capture is not available on a resident serving lane";
        return false;
    }
    hipError_t st = hipStreamSynchronize(stream_);
    if (st == hipSuccess && stream2_) st = hipStreamSynchronize(stream2_);
    if (st != hipSuccess) {
        if (error) *error = hip_error("hipStreamSynchronize", st);
        return false;
    }
    const int32_t v = on ? 1 : 0;
    plan_host_.capture = plan_dev_.capture = v;
    for (int l = 0; l < 2; ++l) lane_host_[l].capture = lane_dev_[l].capture = v;
    capture_all();
    return true;
}

bool Engine::autotune(int reps, Tune* best, float* best_ms, std::string* error) {
    const std::lock_guard<std::mutex> lock(mutex_);
    if (reps < 1) {
        if (error) *error = "invalid repetition count";
        return false;
    }
    const bool dev = last_device_;
    const auto measure = [&](const Tune& t, float* ms) {
        float warm = 0.0f;
        return set_tune_locked(t, error) && replay_locked(2, dev, &warm, error) &&
               replay_locked(reps, dev, ms, error);
    };
    Tune cur = plan_host_.tune;
    float cur_ms = 0.0f;
    if (!measure(cur, &cur_ms)) return false;
    const auto consider = [&](Tune t) {
        float ms = 0.0f;
        if (!measure(t, &ms)) return false;
        if (ms < cur_ms) {
            cur = t;
            cur_ms = ms;
        }
        return true;
    };
    static const int32_t kGrids[] = {256, 512, 1024, 1280, 1536, 1792, 2048, 4096};
    for (const int32_t g : kGrids) {
        Tune t = cur;
        t.fit_grid = g;
        if (!consider(t)) return false;
    }
    for (const int32_t g : kGrids) {
        Tune t = cur;
        t.decode_grid = g;
        if (!consider(t)) return false;
    }
    for (const int32_t w : {4, 8, 16}) {
        if (w == cur.decode_waves) continue;
        Tune t = cur;
        t.decode_waves = w;
        if (!consider(t)) return false;
    }
    {
        Tune t = cur;
        t.chunk_px = cur.chunk_px == 64 ? 128 : 64;
        if (!consider(t)) return false;
    }
    {
        Tune t = cur;
        t.band_rows = cur.band_rows == 8 ? 16 : 8;
        if (!consider(t)) return false;
    }
    for (int32_t h = 0; h <= kHeadRanksMax; ++h) {
        Tune t = cur;
        t.head_ranks = h;
        if (!consider(t)) return false;
    }
    for (const int32_t a : {500, 1000, 2000, 4000, INT32_MAX}) {
        if (a == cur.head_area) continue;
        Tune t = cur;
        t.head_area = a;
        if (!consider(t)) return false;
    }
    if (lanes_ == 2) {
        Tune t = cur;
        t.streams = 3 - cur.streams;
        if (!consider(t)) return false;
    }
    if (!set_tune_locked(cur, error)) return false;
    if (best) *best = cur;
    if (best_ms) *best_ms = cur_ms / static_cast<float>(reps);
    return true;
}

}  // namespace bd
