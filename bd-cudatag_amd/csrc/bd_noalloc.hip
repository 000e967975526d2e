// Single-frame detection that allocates nothing (bd_detect_device_noalloc).
//
// For callers that cannot call hipMalloc: code under HIP graph capture,
// serving loops that must not allocate, harnesses that forbid it.  The frame
// and the output records are the caller's device memory and the work is
// queued on the caller's stream; everything else the pipeline needs (the
// frame's stage buffers, its descriptor, the candidate work list, the
// scalars and the codeword table) lives in `scratch`, a file-scope device
// array of kNoAllocScratchBytes that every GPU the library runs on holds until the
// process ends.  The kernels are the detector's own, so results are those of
// bd_detector_detect_device.
//
// Each call clears the accumulators it needs (a fill kernel, no memset, so
// the call can be captured into a graph) and overwrites the rest.  Calls that
// overlap on one GPU share its array and race; calls on different GPUs do
// not.
#include <hip/hip_runtime.h>

#include <mutex>
#include <sstream>
#include <string>

#include "bd_engine.h"
#define BD_CODES_SPACE __device__
#include "tag_codes.inc"

namespace bd {

// 64 MiB: the stage buffers of one frame up to 1920x1200 (63.4 MiB) plus
// the descriptor, work list, scalars, codewords and output records.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;

namespace {

// In host code this name is a host-side stand-in, not a device address:
// scratch_base() finds each device's copy with hipGetSymbolAddress.
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
uint8_t* g_scratch_base[kMaxDevices] = {};

// Device address of the current device's scratch array, looked up once per
// device with that device current.
uint8_t* scratch_base(std::string* error) {
    int dev = -1;
    hipError_t st = hipGetDevice(&dev);
    if (st != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        if (error) *error = st != hipSuccess ? std::string("hipGetDevice: ") + hipGetErrorString(st)
                                             : std::string("device index beyond the supported 64");
        return nullptr;
    }
    const std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_base[dev]) {
        void* p = nullptr;
        st = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (st != hipSuccess || !p) {
            if (error) *error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(st);
            return nullptr;
        }
        g_scratch_base[dev] = static_cast<uint8_t*>(p);
    }
    return g_scratch_base[dev];
}

struct SetupArgs {
    FrameDesc desc;           // the frame's descriptor, pointers into scratch
    FrameDesc* desc_dev;      // where the kernels read it
    int32_t* prefix_dev;      // SH_COUNT x 2 prefix table of the one-frame plan
    int32_t prefix[SH_COUNT * 2];
    uint32_t* work;           // the work list's header
    uint64_t* codes;          // codeword table in scratch
    int32_t family;
};

// Writes the descriptor, prefix table and codewords, and clears what the
// pipeline accumulates into: the histograms, the scalar block (top row at
// its sentinel), the work-list header and the per-candidate decode ids.
__global__ __launch_bounds__(256) void k_noalloc_setup(SetupArgs a) {
    const int t = threadIdx.x;
    if (t == 0) *a.desc_dev = a.desc;
    if (t < SH_COUNT * 2) a.prefix_dev[t] = a.prefix[t];
    if (t < kWorkItems) a.work[t] = 0u;
    if (t < S_COUNT) a.desc.scal[t] = t == S_TOP ? 0xffffffffu : 0u;
    for (int i = t; i < 320; i += blockDim.x) a.desc.hist[i] = 0u;
    for (int i = t; i < a.desc.tile_h; i += blockDim.x) {
        a.desc.trow_lo[i] = a.desc.tile_w;
        a.desc.trow_hi[i] = -1;
    }
    for (int i = t; i < a.desc.max_candidates; i += blockDim.x) a.desc.dec[i].id = -1;
    const uint64_t* src = a.family == 0 ? bd_codes_tag36h11 : bd_codes_tag16h5;
    for (int i = t; i < a.desc.code_count; i += blockDim.x) a.codes[i] = src[i];
}

// The frame's detection count, min(found, max_detections), and its first
// min(count, capacity) records in the caller's buffers.
__global__ __launch_bounds__(256) void k_noalloc_publish(const FrameDesc* d, bd_detection* out, int32_t capacity,
                                                         int32_t* count) {
    const int found = min(static_cast<int>(d->scal[S_COUNTS + 1]), d->max_detections);
    const int n = min(found, capacity);
    for (int i = threadIdx.x; i < n; i += blockDim.x) out[i] = d->out[i];
    if (threadIdx.x == 0) *count = found;
}

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

}  // namespace

int detect_noalloc(const Params& p, int width, int height, const uint8_t* gray, int stride, bd_detection* out,
                   int capacity, int32_t* count, hipStream_t stream, std::string* error) {
    const int maxc = max_candidates_for(p, width, height);
    const int max_det = max_detections_for(p);
    const int row_ext_log2 = row_ext_log2_for_width(width);
    size_t cur = 0;
    const FrameLayout o = layout_frame(width, height, maxc, row_ext_log2, &cur);
    auto take = [&cur](size_t bytes) {
        const size_t at = cur;
        cur = align256(cur + bytes);
        return at;
    };
    const size_t o_desc = take(sizeof(FrameDesc));
    const size_t o_prefix = take(SH_COUNT * 2 * sizeof(int32_t));
    const size_t o_work = take((kWorkItems + kHeadSlots + static_cast<size_t>(maxc)) * sizeof(uint32_t));
    const size_t o_scal = take(S_COUNT * sizeof(uint32_t));
    const size_t o_codes = take(static_cast<size_t>(code_count_for(p.family)) * sizeof(uint64_t));
    const size_t o_out = take(static_cast<size_t>(max_det) * sizeof(bd_detection));
    if (cur > kNoAllocScratchBytes) {
        if (error) {
            std::ostringstream ss;
            ss << "a " << width << "x" << height << " frame needs " << cur
               << " bytes of working memory; the no-allocation entry point holds " << kNoAllocScratchBytes
               << " (frames up to 1920x1200)";
            *error = ss.str();
        }
        return -1;
    }
    uint8_t* base = scratch_base(error);
    if (!base) return -2;

    SetupArgs a{};
    a.desc = bind_frame(p, width, height, maxc, max_det, row_ext_log2, base, o,
                        reinterpret_cast<const uint64_t*>(base + o_codes));
    a.desc.gray = gray;
    a.desc.stride = stride;
    a.desc.scal = reinterpret_cast<uint32_t*>(base + o_scal);
    a.desc.out = reinterpret_cast<bd_detection*>(base + o_out);
    a.desc_dev = reinterpret_cast<FrameDesc*>(base + o_desc);
    a.prefix_dev = reinterpret_cast<int32_t*>(base + o_prefix);
    a.work = reinterpret_cast<uint32_t*>(base + o_work);
    a.codes = reinterpret_cast<uint64_t*>(base + o_codes);
    a.family = p.family;
    // One frame: every prefix table is {0, blocks of the frame}; only the
    // frame's own width class has blocks.
    int32_t blocks[SH_COUNT] = {a.desc.tile_h, 1, tile_stat_blocks(a.desc.tile_w, a.desc.tile_h)};
    const int cls = row_ext_log2 - kRowExtLog2Min;
    for (int c = 0; c < kWidthClasses; ++c) {
        blocks[SH_CLASS + c] = c == cls ? a.desc.tile_h : 0;
        blocks[SH_CLASS_PAIR + c] = c == cls ? (a.desc.tile_h + 1) / 2 : 0;
    }
    for (int k = 0; k < SH_COUNT; ++k) {
        a.prefix[2 * k] = 0;
        a.prefix[2 * k + 1] = blocks[k];
    }

    Plan plan{};
    plan.frames = a.desc_dev;
    plan.prefix = a.prefix_dev;
    plan.work = a.work;
    plan.nframes = 1;
    plan.work_cap = maxc;
    plan.class_band_runs[cls] = band_runs_for_width(width);
    plan.class_merge_stage[cls] = merge_stage_for_width(width);
    plan.tune = default_tune();
    plan.wide_blocks = binarize_wide_blocks();
    plan.tune.streams = 1;
    for (int k = 0; k < SH_COUNT; ++k) plan.blocks[k] = plan.per_frame[k] = blocks[k];

    hipLaunchKernelGGL(k_noalloc_setup, dim3(1), dim3(256), 0, stream, a);
    hipError_t st = hipGetLastError();
    if (st == hipSuccess) st = launch_pipeline(plan, stream);
    if (st == hipSuccess) {
        hipLaunchKernelGGL(k_noalloc_publish, dim3(1), dim3(256), 0, stream, a.desc_dev, out, capacity, count);
        st = hipGetLastError();
    }
    if (st != hipSuccess) {
        if (error) *error = std::string("kernel launch: ") + hipGetErrorString(st);
        return -2;
    }
    return 0;
}

}  // namespace bd
