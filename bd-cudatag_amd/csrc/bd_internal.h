// Internal declarations shared by the HIP kernels and the host engine.
//
// One launch processes a list of frames ("a batch"; the single-frame
// detector is a batch of one).  Every frame has a FrameDesc in device memory
// pointing at its own slice of the detector's arena; every kernel's grid
// spans all frames and a block finds its frame through a per-kernel prefix
// table (Plan).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "bdcudatag/bdcudatag.h"

namespace bd {

constexpr int kTile = 8;             // reference gpu_detector.cu:38
constexpr int kMinTileContrast = 12; // reference gpu_detector.cu:39
constexpr int kFitDirections = 32;   // reference gpu_detector.cu:46
constexpr int kMaxBoundaryRows = 1024;
constexpr int kMaxHullVertices = 64;
constexpr int kFitIterations = 2;
constexpr int kFitProbeHalf = 24;
constexpr int kMaxProbes = 2 * kFitProbeHalf + 1;
constexpr int kMaxSamples = 16;
constexpr int kMaxCodes = 587;       // largest codebook (tag36h11)

// Device frames are addressed by 32-bit byte offsets from their base in the
// sampling kernels; a frame must span less than 4 GiB.
inline bool frame_span_fits(int width, int height, int stride) {
    return static_cast<uint64_t>(stride) * static_cast<uint64_t>(height - 1) + static_cast<uint64_t>(width) <=
           0xffffffffull;
}

// Candidate keys sorted inside one workgroup's LDS before falling back to a
// radix select over global memory (at most max_candidates <= 1024 keys are
// kept, so the select always leaves them within the cap).  16 KB keeps
// k_select small enough to start on CUs that the other lane's persistent
// k_fit workgroups hold.
constexpr int kSortCap = 2048;
// Runs of one 8-row band joined in LDS by k_binarize_runs (dynamic LDS of
// Plan::class_band_runs runs, 8 bytes each); a band with more runs is joined
// row by row in global memory by k_merge instead.  The capacity follows the
// widest frame of the width class, band_runs = clamp(1.6 * width rounded up
// to 512, kBandRunsMin, kBandRunsMax): 2048 runs (16 KB, eight workgroups
// per CU) at 1280 px, where noisy scenes reach 983 runs per band on average,
// 1826 at the 99th percentile and 1989 at most; 3072 at 1920 px (1605, 2746,
// 2876); 6656 runs at 4096 px (3500 on average, about 1 % of bands above).
// The maximum keeps the workgroup within 64 KB of LDS.  The band's pixel
// masks and end counts (96 bytes per 64 pixels of width) and the pooled
// threshold staging (6 bytes per tile column) fit in any capacity of at
// least 1.6 * width.
constexpr int kBandRunsMin = 1024;
// Largest k_merge row staging (runs per row; 32 KB of LDS with the labels of
// both rows): every row of frames up to 4096 px wide.
constexpr int kMergeStageMax = 2050;
inline int merge_stage_for_width(int width) {
    const int want = width / 2 + 2;
    return want < kMergeStageMax ? want : kMergeStageMax;
}
constexpr int kBandRunsMax = 7680;
inline int band_runs_for_width(int width) {
    const int want = (width * 8 / 5 + 511) / 512 * 512;
    return want < kBandRunsMin ? kBandRunsMin : want > kBandRunsMax ? kBandRunsMax : want;
}
// Component slots per row of the row-extreme tables (log2), per frame from
// its width: the least power of two of at least width / 20, 64 to 256.  Rows
// of noisy eight-tag scenes hold 8.7 components on average and 44 at most
// at 1280 px (64 slots), 28 and 48 at 1920 px (128), 68 and 100 at 4096 px
// (256); a row with more components than slots is scanned whole by k_fit.
// The three sizes are the plan's width classes: k_stats is compiled once
// per size, and k_binarize_runs and k_stats are launched once per class
// present, each launch over that class's frames only, with the class's own
// LDS band capacity (from its widest frame).
constexpr int kRowExtLog2Min = 6, kRowExtLog2Max = 8;
// Pixels the first of two lanes holds beyond the second (about 20 us of the
// device's batch throughput, the head start the first lane's graph has).
constexpr int64_t kLaneLeadPx = 4 << 20;
constexpr uint32_t kRowTableFull = 0xffffffffu;
inline int row_ext_log2_for_width(int width) {
    int lg = kRowExtLog2Min;
    while (lg < kRowExtLog2Max && (20 << lg) < width) ++lg;
    return lg;
}

// k_tile_stats: 16x8-pixel blocks per workgroup (one per thread), numbered
// across each frame's tile rows.
#ifndef BD_TILE_STAT_BLOCKS
#define BD_TILE_STAT_BLOCKS 256
#endif
constexpr int kTileStatBlocks = BD_TILE_STAT_BLOCKS;
inline int tile_stat_blocks(int tile_w, int tile_h) {
    const int pairs = (tile_w + 1) / 2;
    return (tile_h * pairs + kTileStatBlocks - 1) / kTileStatBlocks;
}

// Per-frame scalar block (u32 each).  Accumulators are cleared by the kernel
// that consumes them, so a replayed graph needs no memset nodes.
enum Scalar : int {
    S_TOP = 0,        // min active tile row (atomicMin), 0xffffffff = none
    S_BOTTOM1 = 1,    // max active tile row + 1 (atomicMax), 0 = none
    S_REF0 = 2,       // lower-quartile tile span
    S_REF1 = 3,       // Otsu split of the tile means
    S_COUNTS = 4,     // counts[5]: candidates, detections, runs, quads, roi tile rows
    S_CAND_ACC = 9,   // candidates appended by k_collect
    S_NCAND = 10,     // min(candidates, max_candidates)
    S_MAXRUNS = 11,   // maximal black runs in the frame
    S_DECODED = 12,   // candidates that decoded (before de-duplication)
    S_SEQ = 13,       // pipeline passes completed (k_finalize), never cleared
    S_COUNT = 16
};

struct FrameDesc {
    const uint8_t* gray;
    int32_t active;  // 0: this position carries no frame in the current launch
    int32_t width, height, stride;
    int32_t tile_w, tile_h, words64, row_cap;
    int32_t row_ext_log2;  // log2 of the component slots per row of row_ext
    int32_t max_candidates, max_detections;
    int32_t dim, code_count, max_hamming;
    int32_t min_cluster, min_tag, refine;
    float min_margin, sharpening;
    const uint64_t* codes;
    uint8_t* tile_min;
    uint8_t* tile_max;
    uint8_t* tile_mean;
    uint8_t* tile_thresh;
    int32_t* trow_lo;     // per tile row: leftmost active tile or tile_w
    int32_t* trow_hi;     // per tile row: rightmost active tile or -1
    int32_t* row_roi;     // 4 * tile_h: active, start, end, band joined locally
    uint32_t* hist;       // 256 span bins + 64 mean bins
    uint32_t* scal;       // S_COUNT scalars
    uint64_t* bits;       // height * words64
    uint32_t* row_count;  // height
    uint32_t* runs;       // height * row_cap, x0 | x1 << 16
    uint32_t* label;      // height * row_cap
    uint32_t* area;       // height * row_cap
    uint32_t* bbox;       // height * row_cap * 4
    uint64_t* cand_keys;  // height * row_cap
    uint32_t* roots;      // height * row_cap: band b's component roots from b * 8 * row_cap
    uint32_t* band_nroots;  // tile_h: roots listed per band
    // Row items (k_binarize_runs -> k_stats, and k_fit's scan of a row whose
    // component table overflowed) of the rows of bands joined in LDS:
    // stretches of consecutive runs of one row with one band root, folded.
    // height * row_cap, row y's row_items[y] items from y * row_cap:
    // x0 | x1 << 13 (leftmost start, rightmost end) | (pixels - 1) << 26 |
    // band-root flag << 39 | band root's band row << 40 | its index in that
    // row << 43 (the flag marks the item holding the band root's own run).
    uint64_t* items;
    uint32_t* row_items;  // height
    uint64_t* row_ext;    // height << row_ext_log2: per row, the row's components
                          // root << 32 | x0 << 16 | x1 (leftmost start, rightmost
                          // end), packed from the row's first slot
    uint32_t* row_ext_full;  // height: the row's entries, or kRowTableFull when
                             // its table overflowed (the row is scanned instead)
    uint32_t* candidates; // max_candidates
    float* quads;         // max_candidates * 8
    float* hull_quads;    // max_candidates * 8
    bd_detection* dec;    // max_candidates, per-candidate decode result
    bd_detection* out;    // max_detections, final detections (pinned host memory)
    // Pinned host memory, or null: the five counts from S_COUNTS, then at
    // kHostSeq the frame's completed-pass number, stored last with system
    // scope release (the host polls it instead of waiting on an event); the
    // device wall clock (s_memrealtime) at the pass's start (k_tile_stats)
    // and end (k_finalize) as u64 words at kHostStart and kHostEnd.
    uint32_t* host_counts;
};
constexpr int kHostSeq = 5;
constexpr int kHostEnd = 6;    // u64 at host_counts[6..7] (8-byte aligned)
constexpr int kHostStart = 8;  // u64 at host_counts[8..9]

// Prefix tables over frames: prefix[k][f] = first block of frame f for
// kernel shape k; prefix[k][nframes] = total blocks.
// SH_CLASS + c: tile rows of the frames of width class c (row-table log2
// kRowExtLog2Min + c), zero blocks for every other frame; SH_CLASS_PAIR + c:
// pairs of tile rows (16-row execution tiling of k_binarize_runs).
constexpr int kWidthClasses = kRowExtLog2Max - kRowExtLog2Min + 1;
enum Shape : int {
    SH_TILE_ROW = 0,
    SH_FRAME = 1,
    SH_TSTAT = 2,
    SH_CLASS = 3,
    SH_CLASS_PAIR = SH_CLASS + kWidthClasses,
    SH_COUNT = SH_CLASS_PAIR + kWidthClasses
};

// Batch candidate list: a small header, then one entry per selected
// candidate of any frame (frame << 10 | candidate index), appended by
// k_select and consumed by the persistent k_fit and k_decode grids.  The
// first Tune::head_ranks candidates of every frame (its largest), and every
// further candidate of at least Tune::head_area pixels (up to kHeadSlots per
// frame), sit in a head region of kHeadSlots * nframes entries that is
// handed out first; the rest follow in a tail region in frame order.
enum WorkSlot : int { kWorkCount = 0, kWorkFitCursor = 1, kWorkHeadCount = 3, kWorkItems = 4 };  // slot 2 unused
constexpr int kHeadRanksMax = 4;
constexpr int kHeadSlots = 32;

// Launch parameters that change speed only, never results (autotuned per
// batch shape): persistent grid sizes of k_fit and k_decode, the number of
// each frame's largest candidates scheduled first and the area above which
// a candidate is scheduled first whatever its rank, whether a batch of two
// or more frames runs as two concurrent lanes (streams = 2) or as one, the
// decode waves per workgroup, the pixels per binarization item and the
// pixel rows per binarization workgroup.
struct Tune {
    int32_t fit_grid = 2048;     // 8 resident 256-thread workgroups per CU (default_tune: the occupancy)
    int32_t decode_grid = 1024;  // 4 resident 512-thread workgroups per CU
    int32_t head_ranks = 2;
    int32_t streams = 2;
    int32_t decode_waves = 8;    // 4, 8 or 16 waves per decode workgroup
    int32_t chunk_px = 64;       // pixels packed per k_binarize_runs item: 64 or 128
    int32_t head_area = 1000;    // pixels; larger candidates join the head region
    int32_t band_rows = 8;       // pixel rows per k_binarize_runs workgroup: 8 or 16
};

struct Plan {
    const FrameDesc* frames;
    const int32_t* prefix;  // SH_COUNT * (nframes + 1)
    uint32_t* work;         // kWorkItems + kHeadSlots * nframes + work_cap entries
    int32_t nframes;
    int32_t work_cap;       // sum of max_candidates over frames
    // LDS band capacity of k_binarize_runs (runs) per width class, from the
    // widest frame of the class in this plan (0: no frame of the class).
    int32_t class_band_runs[kWidthClasses];
    // Runs per row that k_merge stages in LDS per width class: the widest
    // frame's maximal runs per row (width / 2 + 2), at most kMergeStageMax.
    int32_t class_merge_stage[kWidthClasses];
    Tune tune;
    int32_t blocks[SH_COUNT];
    // Blocks of every frame for a shape whose frames all have the same
    // count (0: the counts differ), so that kernels find a block's frame by
    // one division instead of the prefix table.
    int32_t per_frame[SH_COUNT];
    // Nonzero: k_binarize_runs also writes the runs and labels of the rows
    // inside each band, which no later kernel reads (stage capture for
    // stage-by-stage comparisons).
    int32_t capture;
    // A k_binarize_runs launch of at most this many workgroups (one round of
    // 1024-thread workgroups, two per CU) runs them with 1024 threads, which
    // shortens each band's chain; larger launches keep 256 threads per
    // workgroup and more workgroups in flight.
    int32_t wide_blocks;
};

// Fills blocks and per_frame from the shapes' prefix tables (SH_COUNT
// tables of n + 1 entries).
inline void plan_set_blocks(Plan* p, const int32_t* prefix, int n) {
    for (int k = 0; k < SH_COUNT; ++k) {
        const int32_t* q = prefix + static_cast<size_t>(k) * (n + 1);
        p->blocks[k] = q[n];
        int32_t per = n > 0 ? q[1] - q[0] : 0;
        for (int f = 1; f < n && per > 0; ++f)
            if (q[f + 1] - q[f] != per) per = 0;
        p->per_frame[k] = per;
    }
}

// Device-resident serving (bd_server_create_resident): requests of one frame
// size class arrive in a ring in pinned host memory; a lane's launch admits
// them on the device (k_admit) and publishes their results (k_publish).
// A request slot is published by writing seq = ticket + 1 after gray and
// stride; a completion slot by writing seq = ticket + 1 after count.
struct RequestSlot {
    const uint8_t* gray;
    int32_t stride;
    uint32_t seq;
};
struct CompletionSlot {
    int32_t count;  // detections (min(found, max_detections))
    int32_t batch;  // requests admitted by the same launch
    uint32_t seq;
    uint32_t reserved;
};
struct ResidentArgs {
    RequestSlot* slots;        // ring of `depth` request slots (pinned host memory)
    CompletionSlot* done;      // ring of `depth` completion slots (pinned host memory)
    bd_detection* out;         // depth x max_det records (pinned host memory)
    const uint32_t* stop;      // pinned host memory: nonzero ends every wait
    uint32_t* head;            // device memory: the class's next ticket to admit
    uint32_t* lane_tickets;    // device memory: tickets of this lane's positions, count last
    bd_detection* lane_out;    // device memory: max_det records per position (k_finalize's output)
    int32_t depth, max_det;
    uint64_t wait_ticks;       // device wall-clock ticks an admission waits for a request
    uint32_t max_polls;        // and at most this many polls of the ring
};

// Kernels in pipeline order: tile_stats, binarize_runs, merge, stats,
// collect, select, fit, decode, finalize.
constexpr int kStageCount = 9;

// Default launch parameters: persistent grids of one full round of the
// workgroups the device's CUs hold for the compiled k_fit and k_decode.
Tune default_tune();
// Plan::wide_blocks for the current device: twice its CU count.
int32_t binarize_wide_blocks();

// Enqueues one kernel (0..kStageCount-1) for every frame of the plan.
hipError_t launch_stage(const Plan& plan, int stage, hipStream_t stream);
// Enqueues the whole pipeline for every frame of the plan on `stream`.
hipError_t launch_pipeline(const Plan& plan, hipStream_t stream);
// Device-resident serving: admission before the pipeline and publication
// after it, for the plan's frame positions (one wave each).
hipError_t launch_admit(const Plan& plan, const ResidentArgs& args, hipStream_t stream);
hipError_t launch_publish(const Plan& plan, const ResidentArgs& args, hipStream_t stream);

// Frame bindings of positions first .. first + count - 1 (count <=
// kBindFrames) of a device-path call, carried as kernel arguments: a new
// frame pointer reaches the descriptors by one small kernel instead of a
// host-to-device copy of the descriptors (a rebinding call measured 1 us
// faster).
constexpr int kBindFrames = 64;
struct BindArgs {
    const uint8_t* gray[kBindFrames];
    int32_t stride[kBindFrames];
    int32_t active[kBindFrames];
    int32_t first, count;
};
hipError_t launch_bind(FrameDesc* frames, const BindArgs& args, hipStream_t stream);

}  // namespace bd
