// HIP kernels of the detection pipeline for gfx950 (CDNA4, wave64).
//
// Nine kernels, each with a grid that spans every frame of the plan:
//   k_tile_stats      8x8 tile min/max/mean, histograms, active extents   (a)
//   k_binarize_runs   quartile span + Otsu split and the band's ROI row
//                     (a), pooled threshold, 64-px bit packing, black runs
//                     per row, band-local union-find in LDS               (b)
//   k_merge           8-connected union across band boundaries            (b)
//   k_stats           band roots resolved to component roots, root labels,
//                     component area and bounding box, per-row component
//                     tables (leftmost start, rightmost end)               (b)
//   k_collect         tag-shaped component filter                         (c)
//   k_select          largest-first candidate order (LDS bitonic sort)    (c)
//   k_fit             row extremes (table lookups), hull probe, quad,
//                     edge refinement                                     (d)
//   k_decode          4 rotations x {refined, hull} decode                (e)
//   k_finalize        de-duplication, stable id sort, FP64 homography     (e)
//
// Integer stages are bit-exact with the reference algorithm.  Floating-point
// stages are written as the same sequence of IEEE +, -, *, /, sqrt as the
// oracle and are compiled with -ffp-contract=off, so they are bit-identical
// to it as well.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdint>

#include "bd_internal.h"

namespace bd {
namespace {

constexpr int kWave = 64;

// cos/sin of fl32(fl32(6.28318530718f * d) / 32), correctly rounded to FP32:
// the probe directions of the reference hull search (gpu_detector.cu:1139-1142).
__constant__ float c_dir_cos[32] = {
    0x1p+0f,          0x1.f6297cp-1f,   0x1.d906bcp-1f,   0x1.a9b662p-1f,
    0x1.6a09e6p-1f,   0x1.1c73b2p-1f,   0x1.87de2ap-2f,   0x1.8f8b78p-3f,
    -0x1.777a5cp-25f, -0x1.8f8b84p-3f,  -0x1.87de30p-2f,  -0x1.1c73b8p-1f,
    -0x1.6a09e6p-1f,  -0x1.a9b664p-1f,  -0x1.d906c0p-1f,  -0x1.f6297ep-1f,
    -0x1p+0f,         -0x1.f6297cp-1f,  -0x1.d906bcp-1f,  -0x1.a9b660p-1f,
    -0x1.6a09e2p-1f,  -0x1.1c73acp-1f,  -0x1.87de16p-2f,  -0x1.8f8b8cp-3f,
    0x1.99bc5cp-27f,  0x1.8f8b90p-3f,   0x1.87de36p-2f,   0x1.1c73bap-1f,
    0x1.6a09eep-1f,   0x1.a9b662p-1f,   0x1.d906bep-1f,   0x1.f6297ep-1f,
};
__constant__ float c_dir_sin[32] = {
    0x0p+0f,          0x1.8f8b84p-3f,   0x1.87de2cp-2f,   0x1.1c73b4p-1f,
    0x1.6a09e6p-1f,   0x1.a9b664p-1f,   0x1.d906bcp-1f,   0x1.f6297ep-1f,
    0x1p+0f,          0x1.f6297cp-1f,   0x1.d906bcp-1f,   0x1.a9b660p-1f,
    0x1.6a09e6p-1f,   0x1.1c73b2p-1f,   0x1.87de20p-2f,   0x1.8f8b82p-3f,
    -0x1.777a5cp-24f, -0x1.8f8b9ap-3f,  -0x1.87de2ap-2f,  -0x1.1c73b6p-1f,
    -0x1.6a09eap-1f,  -0x1.a9b668p-1f,  -0x1.d906c2p-1f,  -0x1.f6297cp-1f,
    -0x1p+0f,         -0x1.f6297cp-1f,  -0x1.d906bap-1f,  -0x1.a9b65ep-1f,
    -0x1.6a09dep-1f,  -0x1.1c73b6p-1f,  -0x1.87de2ap-2f,  -0x1.8f8b78p-3f,
};

// Frame bytes in the global address space, addressed by 32-bit offsets (the
// host rejects device frames spanning 4 GiB or more), so that loads take
// the frame base from scalar registers and are issued as global loads.
typedef const uint8_t __attribute__((address_space(1))) gbyte;
typedef unsigned int vec4u __attribute__((ext_vector_type(4)));
typedef const vec4u __attribute__((address_space(1))) gvec16;

__device__ inline const gbyte* frame_bytes(const uint8_t* gray) { return (const gbyte*)gray; }

typedef const uint64_t __attribute__((address_space(1))) gword64;

// 16 frame bytes at byte offset o (16-byte aligned frame rows).
__device__ inline uint4 load16(const gbyte* image, size_t o) {
    const vec4u v = *reinterpret_cast<const gvec16*>(image + o);
    return make_uint4(v.x, v.y, v.z, v.w);
}

// Frame of a block for one grid shape.  Frames with equal block counts for
// the shape (a uniform batch) divide; otherwise the prefix table (one entry
// per frame; every table starts at 0 and ends at the shape's block count) is
// searched.  The tables do not change while a kernel runs, so they are read
// as constant memory, scalar loads through the scalar cache: the frame at
// the proportional guess and its successor's entry in one 8-byte load
// first, a binary search on the side the block lies if the guess misses.
typedef const int32_t __attribute__((address_space(4))) kint32;
typedef const uint64_t __attribute__((address_space(4))) kuint64;
__device__ inline int frame_of(const Plan& plan, int shape, int block, int* local) {
    const int n = plan.nframes;
    const int per = plan.per_frame[shape];
    if (n == 1 || per > 0) {
        const int f = n == 1 ? 0 : block / per;
        *local = block - f * per;
        return f;
    }
    const int32_t* prefix = plan.prefix + shape * (n + 1);
    const uint32_t total = static_cast<uint32_t>(plan.blocks[shape]);
    const int g = min(n - 1, static_cast<int>(static_cast<uint32_t>(block) * static_cast<uint32_t>(n) / max(total, 1u)));
    // One 8-byte load for both entries (4-byte alignment suffices for
    // scalar loads).
    const uint64_t ab = *(const kuint64*)(prefix + g);
    const int a = static_cast<int>(static_cast<uint32_t>(ab)), b = static_cast<int>(ab >> 32);
    if (a <= block && block < b) {
        *local = block - a;
        return g;
    }
    const kint32* p = (const kint32*)prefix;
    int lo = block < a ? 0 : g + 1, hi = block < a ? g - 1 : n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (p[mid] <= block) lo = mid; else hi = mid - 1;
    }
    *local = block - p[lo];
    return lo;
}

// Frame descriptors as the pipeline kernels read them: constant memory too
// (only the host and the admission kernels, in earlier launches, write
// them), so the uniform descriptor fields come through scalar loads.
typedef const FrameDesc __attribute__((address_space(4))) KFrame;

// The descriptor's first eight words (frame pointer, active flag, width,
// height, stride, tile columns and rows) in one scalar load, so that the
// active check and the frame geometry share one round trip.
struct FrameHead {
    const uint8_t* gray;
    int active, width, height, stride, tile_w, tile_h;
};
__device__ inline FrameHead frame_head(const KFrame* d) {
    static_assert(offsetof(FrameDesc, gray) == 0 && offsetof(FrameDesc, active) == 8 &&
                      offsetof(FrameDesc, width) == 12 && offsetof(FrameDesc, height) == 16 &&
                      offsetof(FrameDesc, stride) == 20 && offsetof(FrameDesc, tile_w) == 24 &&
                      offsetof(FrameDesc, tile_h) == 28,
                  "descriptor head layout");
    typedef int i8v __attribute__((ext_vector_type(8)));
    const i8v v = *(const __attribute__((address_space(4))) i8v*)d;
    FrameHead h;
    h.gray = reinterpret_cast<const uint8_t*>((static_cast<uint64_t>(static_cast<uint32_t>(v[1])) << 32) |
                                              static_cast<uint32_t>(v[0]));
    h.active = v[2];
    h.width = v[3];
    h.height = v[4];
    h.stride = v[5];
    h.tile_w = v[6];
    h.tile_h = v[7];
    return h;
}

// A uniform read of data that no workgroup of the running kernel writes
// (earlier launches wrote it): a constant-memory load, so a scalar load
// through the scalar cache.
template <class T>
__device__ inline T kld(const T* p) {
    return *(const T __attribute__((address_space(4)))*)p;
}

// Inclusive prefix sum over the wave by DPP (the whole wave must be
// active): row_shr 1/2/4/8 scan each 16-lane row, row_bcast:15 carries row
// 0 into row 1 and row 2 into row 3, row_bcast:31 carries lanes 0-31 into
// rows 2 and 3.
template <int kCtrl, int kRowMask>
__device__ inline uint32_t dpp_add(uint32_t x) {
    return x + static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), kCtrl, kRowMask, 0xf, false));
}

__device__ inline uint32_t wave_incl_scan(uint32_t v) {
    v = dpp_add<0x111, 0xf>(v);
    v = dpp_add<0x112, 0xf>(v);
    v = dpp_add<0x114, 0xf>(v);
    v = dpp_add<0x118, 0xf>(v);
    v = dpp_add<0x142, 0xa>(v);
    v = dpp_add<0x143, 0xc>(v);
    return v;
}

__device__ inline uint32_t wave_excl_scan(uint32_t v, int lane, uint32_t* total) {
    (void)lane;
    const uint32_t x = wave_incl_scan(v);
    *total = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x), kWave - 1));
    return x - v;
}

__device__ inline uint32_t wave_sum(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(wave_incl_scan(v)), kWave - 1));
}

// Cross-lane moves by DPP, no LDS traffic.  dpp_from_lower gives lane i the
// value of lane i - 1 (lane 0 gets `edge`), dpp_from_upper the value of lane
// i + 1 (lane 63 gets `edge`).  A source lane that is inactive in EXEC also
// yields `edge`, so call these with the whole wave active.
__device__ inline uint32_t dpp_from_lower(uint32_t edge, uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(edge), static_cast<int>(v),
                                                             0x138 /* wave_shr:1 */, 0xf, 0xf, false));
}
__device__ inline uint32_t dpp_from_upper(uint32_t edge, uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(edge), static_cast<int>(v),
                                                             0x130 /* wave_shl:1 */, 0xf, 0xf, false));
}
__device__ inline uint32_t wave_lane(uint32_t v, int lane) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), lane));
}
// Minimum of v over lanes 0-31, returned to every lane: row_shr 1/2/4/8
// leaves each 16-lane row's minimum in its lane 15, row_bcast:15 folds row
// 0 into row 1, and lane 31 is read back.
__device__ inline uint32_t min_lanes_0_31(uint32_t v) {
    const int all = static_cast<int>(0xffffffffu);
    v = min(v, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(all, static_cast<int>(v), 0x111, 0xf, 0xf, false)));
    v = min(v, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(all, static_cast<int>(v), 0x112, 0xf, 0xf, false)));
    v = min(v, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(all, static_cast<int>(v), 0x114, 0xf, 0xf, false)));
    v = min(v, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(all, static_cast<int>(v), 0x118, 0xf, 0xf, false)));
    v = min(v, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(all, static_cast<int>(v), 0x142, 0xa, 0xf, false)));
    return wave_lane(v, 31);
}

// Minimum / maximum of v over the whole wave (every lane active), by the
// same DPP row shifts and broadcasts as wave_incl_scan; lane 63 ends up
// holding the result, which is returned to every lane.
template <int kCtrl, int kRowMask>
__device__ inline uint32_t dpp_min(uint32_t x) {
    return min(x, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(0xffffffffu), static_cast<int>(x),
                                                                    kCtrl, kRowMask, 0xf, false)));
}
template <int kCtrl, int kRowMask>
__device__ inline uint32_t dpp_max(uint32_t x) {
    return max(x, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), kCtrl, kRowMask, 0xf, false)));
}
__device__ inline uint32_t wave_min(uint32_t v) {
    v = dpp_min<0x111, 0xf>(v);
    v = dpp_min<0x112, 0xf>(v);
    v = dpp_min<0x114, 0xf>(v);
    v = dpp_min<0x118, 0xf>(v);
    v = dpp_min<0x142, 0xa>(v);
    v = dpp_min<0x143, 0xc>(v);
    return wave_lane(v, kWave - 1);
}
__device__ inline uint32_t wave_max(uint32_t v) {
    v = dpp_max<0x111, 0xf>(v);
    v = dpp_max<0x112, 0xf>(v);
    v = dpp_max<0x114, 0xf>(v);
    v = dpp_max<0x118, 0xf>(v);
    v = dpp_max<0x142, 0xa>(v);
    v = dpp_max<0x143, 0xc>(v);
    return wave_lane(v, kWave - 1);
}

// ---------------------------------------------------------------------------
// (a) tile statistics
// ---------------------------------------------------------------------------

__device__ inline void acc_word(uint32_t w, uint32_t& mn, uint32_t& mx, uint32_t& sum) {
    const uint32_t b0 = w & 0xffu, b1 = (w >> 8) & 0xffu, b2 = (w >> 16) & 0xffu, b3 = w >> 24;
    mn = min(mn, min(min(b0, b1), min(b2, b3)));
    mx = max(mx, max(max(b0, b1), max(b2, b3)));
    sum = __builtin_amdgcn_sad_u8(w, 0u, sum);
}

// Each thread owns a 16x8 pixel block (two tiles of one tile row) read as
// eight 16-byte rows issued together, so a wave reads 1 KiB contiguous per
// pixel row.  A frame's blocks are numbered tile row by tile row and each
// workgroup takes 256 consecutive ones: every thread has a block but the
// frame's last workgroup's tail, and the 64-frame 1280x800 batch is 2048
// workgroups, one round of resident ones.  A tile row may straddle two
// workgroups, so its active extent is combined by global atomics (k_finalize
// of the previous pass, or creation, resets it).  reference
// gpu_detector.cu:104-166 (k_tile_stats, k_tile_histograms) plus the per-row
// active extents that k_build_row_roi (:282-357) needs.
__global__ __launch_bounds__(kTileStatBlocks) void k_tile_stats(Plan plan) {
    // The batch's candidate list starts empty (block 0 exists in every
    // launch, whichever frames are active).
    if (blockIdx.x == 0 && threadIdx.x < kWorkItems) plan.work[threadIdx.x] = 0;
    int blk;
    const KFrame* d = (const KFrame*)plan.frames + frame_of(plan, SH_TSTAT, blockIdx.x, &blk);
    const FrameHead fh = frame_head(d);
    if (!fh.active) return;
    const int W = fh.width, H = fh.height, stride = fh.stride, tw = fh.tile_w, th = fh.tile_h;
    const uint8_t* gray = fh.gray;
    const int t = threadIdx.x;
    // The counts later kernels accumulate start at zero (the histograms and
    // the active extents this kernel accumulates were cleared by the
    // previous pass's k_finalize, or at creation).
    if (blk == 0 && t < 4) d->scal[S_COUNTS + t] = 0;
    // The pass's start on the device clock, for the host's timing.
    if (blk == 0 && t == 5 && d->host_counts)
        *reinterpret_cast<uint64_t*>(d->host_counts + kHostStart) = __builtin_amdgcn_s_memrealtime();
    if (blk == 0 && t == 4) {
        d->scal[S_CAND_ACC] = 0;
        d->scal[S_NCAND] = 0;
        d->scal[S_MAXRUNS] = 0;
        d->scal[S_DECODED] = 0;
    }
    const int pairs = (tw + 1) >> 1;
    const int g0 = blk * kTileStatBlocks;
    const int g1 = min(g0 + kTileStatBlocks, th * pairs);
    const int ty0 = g0 / pairs;             // first tile row of the workgroup
    const int nrows = (g1 - 1) / pairs - ty0 + 1;
    // This thread's block: its eight 16-byte rows are issued first, so
    // that their latency overlaps the LDS set-up below.
    const int g = g0 + t;
    const bool have = g < g1;
    const int ty = have ? g / pairs : ty0;
    const int p = g - ty * pairs;
    const int y0 = ty * kTile;
    const int rows = min(y0 + kTile, H) - y0;
    const int x0 = p * 16;
    const bool vec = ((reinterpret_cast<uintptr_t>(gray) | static_cast<uintptr_t>(stride)) & 15u) == 0;
    const bool whole = have && vec && x0 + 16 <= W && rows == kTile;
    uint4 v[kTile];
    if (whole) {
        const gbyte* src = frame_bytes(gray) + static_cast<size_t>(y0) * stride + x0;
#pragma unroll
        for (int r = 0; r < kTile; ++r) v[r] = load16(src, static_cast<size_t>(r) * stride);
    }
    // Histograms, one copy per wave: neighbouring tiles often share a bin,
    // and atomics to one LDS address serialise.
    constexpr int kWaves = kTileStatBlocks / kWave;
    __shared__ uint32_t hs[kWaves][256 + 64];
    __shared__ int lo_s[kTileStatBlocks], hi_s[kTileStatBlocks];
    // The workgroup's tiles (contiguous in the tile arrays) are staged here
    // and written with coalesced word stores.
    constexpr int kStage = 2 * kTileStatBlocks;
    __shared__ __align__(4) uint8_t s_min[kStage], s_max[kStage], s_mean[kStage];
    __shared__ uint32_t s_top, s_bottom1;
    for (int i = t; i < kWaves * 320; i += blockDim.x) (&hs[0][0])[i] = 0;
    if (t < nrows) {
        lo_s[t] = tw;
        hi_s[t] = -1;
    }
    if (t == 0) {
        s_top = 0xffffffffu;
        s_bottom1 = 0;
    }
    const int first_tile = ty0 * tw + 2 * (g0 - ty0 * pairs);
    __syncthreads();

    uint32_t* my_hs = hs[t / kWave];
    // The workgroup's first and last active tile rows (a wave reduction,
    // one LDS atomic per wave): 1 + row of this thread's active tiles.
    uint32_t act_top = 0xffffffffu, act_bottom1 = 0;
    if (have) {
        uint32_t mnA = 255, mxA = 0, sumA = 0, mnB = 255, mxB = 0, sumB = 0;
        const int nA = min(8, W - x0);
        const int nB = min(8, W - x0 - 8);
        if (whole) {
#pragma unroll
            for (int r = 0; r < kTile; ++r) {
                acc_word(v[r].x, mnA, mxA, sumA);
                acc_word(v[r].y, mnA, mxA, sumA);
                acc_word(v[r].z, mnB, mxB, sumB);
                acc_word(v[r].w, mnB, mxB, sumB);
            }
        } else if (vec && x0 + 16 <= W) {
            const int y1 = y0 + rows;
            for (int y = y0; y < y1; ++y) {
                const uint4 w = load16(frame_bytes(gray), static_cast<size_t>(y) * stride + x0);
                acc_word(w.x, mnA, mxA, sumA);
                acc_word(w.y, mnA, mxA, sumA);
                acc_word(w.z, mnB, mxB, sumB);
                acc_word(w.w, mnB, mxB, sumB);
            }
        } else {
            const int y1 = y0 + rows;
            for (int y = y0; y < y1; ++y) {
                const uint8_t* row = gray + static_cast<size_t>(y) * stride + x0;
                for (int x = 0; x < nA; ++x) {
                    const uint32_t px = row[x];
                    mnA = min(mnA, px);
                    mxA = max(mxA, px);
                    sumA += px;
                }
                for (int x = 0; x < nB; ++x) {
                    const uint32_t px = row[8 + x];
                    mnB = min(mnB, px);
                    mxB = max(mxB, px);
                    sumB += px;
                }
            }
        }
        const int ta = 2 * p;
        const uint32_t cA = static_cast<uint32_t>(nA * rows);
        const uint32_t meanA = (sumA + cA / 2) / cA;
        const int ia = ty * tw + ta - first_tile;  // staging index (< kStage)
        s_min[ia] = static_cast<uint8_t>(mnA);
        s_max[ia] = static_cast<uint8_t>(mxA);
        s_mean[ia] = static_cast<uint8_t>(meanA);
        atomicAdd(&my_hs[(mxA - mnA) & 0xffu], 1u);
        atomicAdd(&my_hs[256 + (meanA >> 2)], 1u);
        int lo = tw, hi = -1;
        if (static_cast<int>(mxA - mnA) >= kMinTileContrast) lo = hi = ta;
        if (nB > 0) {
            const uint32_t cB = static_cast<uint32_t>(nB * rows);
            const uint32_t meanB = (sumB + cB / 2) / cB;
            s_min[ia + 1] = static_cast<uint8_t>(mnB);
            s_max[ia + 1] = static_cast<uint8_t>(mxB);
            s_mean[ia + 1] = static_cast<uint8_t>(meanB);
            atomicAdd(&my_hs[(mxB - mnB) & 0xffu], 1u);
            atomicAdd(&my_hs[256 + (meanB >> 2)], 1u);
            if (static_cast<int>(mxB - mnB) >= kMinTileContrast) {
                lo = min(lo, ta + 1);
                hi = ta + 1;
            }
        }
        if (hi >= 0) {
            atomicMin(&lo_s[ty - ty0], lo);
            atomicMax(&hi_s[ty - ty0], hi);
            act_top = static_cast<uint32_t>(ty);
            act_bottom1 = static_cast<uint32_t>(ty) + 1u;
        }
    }
    act_top = wave_min(act_top);
    act_bottom1 = wave_max(act_bottom1);
    if (t % kWave == 0 && act_bottom1 != 0) {
        atomicMin(&s_top, act_top);
        atomicMax(&s_bottom1, act_bottom1);
    }
    __syncthreads();
    // One pair of global atomics per workgroup for the frame's first and
    // last active tile rows.
    if (t == 0 && s_bottom1 != 0) {
        atomicMin(&d->scal[S_TOP], s_top);
        atomicMax(&d->scal[S_BOTTOM1], s_bottom1);
    }
    for (int i = t; i < 320; i += blockDim.x) {
        uint32_t n = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) n += hs[w][i];
        if (n) atomicAdd(&d->hist[i], n);
    }
    if (t < nrows && hi_s[t] >= 0) {
        const int row = ty0 + t;
        atomicMin(&d->trow_lo[row], lo_s[t]);
        atomicMax(&d->trow_hi[row], hi_s[t]);
    }
    // The staged tiles: bytes [first, first + count) of each tile array, as
    // aligned words with the unaligned ends byte by byte.
    const int last_ty = (g1 - 1) / pairs;
    const int last_p = g1 - 1 - last_ty * pairs;
    const int first = first_tile;
    const int count = last_ty * tw + min(2 * last_p + 2, tw) - first;
    const int a0 = min((first + 3) & ~3, first + count), a1 = max(a0, (first + count) & ~3);
    for (int i = t; i < (a1 - a0) / 4; i += blockDim.x) {
        const int o = a0 - first + 4 * i;
        const auto word = [&](const uint8_t* src) {
            return static_cast<uint32_t>(src[o]) | (static_cast<uint32_t>(src[o + 1]) << 8) |
                   (static_cast<uint32_t>(src[o + 2]) << 16) | (static_cast<uint32_t>(src[o + 3]) << 24);
        };
        reinterpret_cast<uint32_t*>(d->tile_min + a0)[i] = word(s_min);
        reinterpret_cast<uint32_t*>(d->tile_max + a0)[i] = word(s_max);
        reinterpret_cast<uint32_t*>(d->tile_mean + a0)[i] = word(s_mean);
    }
    const int head = a0 - first, tail = first + count - a1;
    if (t < head + tail) {
        const int o = t < head ? t : a1 - first + (t - head);
        d->tile_min[first + o] = s_min[o];
        d->tile_max[first + o] = s_max[o];
        d->tile_mean[first + o] = s_mean[o];
    }
}

// The frame's references and active tile rows, for one workgroup of
// k_binarize_runs (every band workgroup derives them from the 1.3 KB of
// histograms k_tile_stats left, instead of a one-workgroup kernel between
// the two): the noise-floor quartile and the Otsu split (reference
// gpu_detector.cu:174-230) and the first and last active tile rows
// (:282-300).  Called by one whole wave (lanes 0-63).
struct FrameRef {
    int ref0;    // lower-quartile tile span (k_tile_reference)
    int ref1;    // Otsu split of the tile means, scaled to a pixel level
    int top;     // first active tile row (-1 when none)
    int bottom;  // last active tile row (-1 when none)
};

// The inputs, loaded first so that their latency overlaps other loads:
// four span bins and one mean bin per lane, the active-row extremes.
struct RefInputs {
    uint4 s4;
    uint32_t m;
    uint32_t top, bottom1;
};

__device__ inline RefInputs load_ref_inputs(const KFrame* d, int lane) {
    RefInputs r;
    r.s4 = *reinterpret_cast<const uint4*>(d->hist + 4 * lane);
    r.m = d->hist[256 + lane];
    r.top = d->scal[S_TOP];
    r.bottom1 = d->scal[S_BOTTOM1];
    return r;
}

__device__ inline FrameRef frame_reference(const KFrame* d, const RefInputs& in, int lane) {
    FrameRef fr;
    fr.top = static_cast<int>(in.top);
    fr.bottom = static_cast<int>(in.bottom1) - 1;
    // Quartile: the first span bin whose cumulative count exceeds total/4
    // (:184-198); four bins per lane, a wave scan of their sums.
    const uint4 s4 = in.s4;
    const uint32_t m = in.m;
    uint32_t total;
    uint32_t cum = wave_excl_scan(s4.x + s4.y + s4.z + s4.w, lane, &total);
    fr.ref0 = 0;
    if (total > 0) {
        const uint32_t target = total / 4;
        const uint32_t bins[4] = {s4.x, s4.y, s4.z, s4.w};
        int first = 4;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            cum += bins[j];  // cumulative count through bin 4 * lane + j
            if (first == 4 && cum > target) first = j;
        }
        const uint64_t crossing = __ballot(first < 4);
        const int src = __ffsll(static_cast<long long>(crossing)) - 1;
        fr.ref0 = static_cast<int>(wave_lane(static_cast<uint32_t>(4 * lane + first), src));
    }

    // Otsu over the 64 mean bins (:200-229).  The reference accumulates the
    // class weights and sums in FP32 one bin at a time; while every partial
    // sum is an integer below 2^24 those FP32 sums are exact, so each bin's
    // terms equal the integer prefix sums converted once and the 64 bins can
    // be evaluated in parallel.  Larger frames take the serial path.
    int otsu = 0;
    const uint64_t tiles = static_cast<uint64_t>(d->tile_w) * d->tile_h;
    if (tiles * 63u < (1u << 24)) {
        uint32_t tw_total, ts_total;
        const uint32_t wex = wave_excl_scan(m, lane, &tw_total);
        const uint32_t sex = wave_excl_scan(static_cast<uint32_t>(lane) * m, lane, &ts_total);
        const float total_w = static_cast<float>(tw_total);
        const float sum = static_cast<float>(ts_total);
        float between = -1.0f;
        bool valid = false;
        if (total_w > 0.0f) {
            const float weight_below = static_cast<float>(wex + m);
            const float weight_above = total_w - weight_below;
            if (weight_below > 0.0f && weight_above > 0.0f) {
                const float sum_below = static_cast<float>(sex + static_cast<uint32_t>(lane) * m);
                const float mean_below = sum_below / weight_below;
                const float mean_above = (sum - sum_below) / weight_above;
                const float delta = mean_below - mean_above;
                between = weight_below * weight_above * delta * delta;
                valid = true;
            }
        }
        // First bin with the largest between-class variance, by two DPP
        // reductions: `between` is never negative, so its bit pattern
        // (plus one, zero marking an invalid bin) orders like its value.
        const uint32_t key = valid ? __float_as_uint(between) + 1u : 0u;
        const uint32_t best = wave_max(key);
        const uint32_t idx = wave_min(best != 0u && key == best ? static_cast<uint32_t>(lane) : 64u);
        otsu = idx < 64u ? static_cast<int>(idx) : 0;
    } else {
        // Every lane runs the serial loop on the same values (uniform).
        float total_w = 0.0f, sum = 0.0f;
        for (int i = 0; i < 64; ++i) {
            const uint32_t mi = wave_lane(m, i);
            total_w += static_cast<float>(mi);
            sum += static_cast<float>(i) * static_cast<float>(mi);
        }
        if (total_w > 0.0f) {
            float weight_below = 0.0f, sum_below = 0.0f, best = -1.0f;
            for (int i = 0; i < 64; ++i) {
                const uint32_t mi = wave_lane(m, i);
                weight_below += static_cast<float>(mi);
                if (weight_below <= 0.0f) continue;
                const float weight_above = total_w - weight_below;
                if (weight_above <= 0.0f) break;
                sum_below += static_cast<float>(i) * static_cast<float>(mi);
                const float mean_below = sum_below / weight_below;
                const float mean_above = (sum - sum_below) / weight_above;
                const float delta = mean_below - mean_above;
                const float between = weight_below * weight_above * delta * delta;
                if (between > best) {
                    best = between;
                    otsu = i;
                }
            }
        }
    }
    fr.ref1 = otsu * 4 + 2;
    return fr;
}

// ROI of tile row ty, whose active extent is [lo, hi] (hi < 0: none):
// every tile row between the first and last active row (+-1) is searched,
// between its own active extent +-4 tiles, or whole when it has none
// (k_build_row_roi :314-356).  Returns active, start, end.
__device__ inline int3 tile_row_roi(const FrameRef& fr, int ty, int lo, int hi, int tile_w) {
    if (fr.bottom < 0 || ty < fr.top - 1 || ty > fr.bottom + 1) return make_int3(0, 0, -1);
    if (hi < 0) return make_int3(1, 0, tile_w - 1);
    return make_int3(1, max(0, lo - 4), min(tile_w - 1, hi + 4));
}

// ---------------------------------------------------------------------------
// (b) binarization and runs
// ---------------------------------------------------------------------------

// Run words are x0 | x1 << 16; bit 31 marks a band root (see
// k_binarize_runs).  x0, x1 < 8192.
// n / d for 0 <= n < 2^20 and 1 <= d < 2^12 from a float reciprocal of d
// (rd = 1.0f / d): the truncated product is within one of the quotient (its
// relative error is below 2^-22), and one correction each way fixes it.
__device__ inline int div_small(int n, int d, float rd) {
    int q = static_cast<int>(static_cast<float>(n) * rd);
    const int rem = n - q * d;
    q += rem >= d ? 1 : 0;
    q -= rem < 0 ? 1 : 0;
    return q;
}

__device__ inline int run_x0(uint32_t r) { return static_cast<int>(r & 0xffffu); }
__device__ inline int run_x1(uint32_t r) { return static_cast<int>((r >> 16) & 0x7fffu); }
constexpr uint32_t kRootFlag = 0x80000000u;

__device__ inline uint32_t lds_find(uint32_t* lab, uint32_t x) {
    // Path splitting: each visited label is lowered to its grandparent.
    // Parents have smaller indices than their children, so atomicMin only
    // moves a label towards the root and concurrent finds and unions stay
    // consistent.
    uint32_t p = lab[x];
    while (p != x) {
        const uint32_t g = lab[p];
        if (g != p) atomicMin(&lab[x], g);
        x = p;
        p = g;
    }
    return x;
}

// Joins the sets of a and b; returns the joined set's root at the time of
// the join (an ancestor of both from then on).
__device__ inline uint32_t lds_union(uint32_t* lab, uint32_t a, uint32_t b) {
    for (;;) {
        a = lds_find(lab, a);
        b = lds_find(lab, b);
        if (a == b) return a;
        if (a > b) {
            const uint32_t tmp = a;
            a = b;
            b = tmp;
        }
        const uint32_t old = atomicMin(&lab[b], a);
        if (old == b) return a;
        b = old;
    }
}

// Bit b (0-3) set when byte b of x is above t (unsigned), four bytes at
// once: x > t is x >= t + 1, decided per byte by its top bit where those
// differ and else by the low seven bits, whose difference borrows into the
// top bit (bytes cannot borrow from each other: each top bit is set before
// the subtraction).  The four flags are gathered by one multiply.
// No pixel is above 255: that threshold is selected at the end rather than
// branched on, so the loads feeding the compare are never predicated.
__device__ inline uint32_t bytes_above(uint32_t x, uint32_t t) {
    const uint32_t t1 = (min(t, 254u) + 1u) * 0x01010101u;
    const uint32_t low = (x | 0x80808080u) - (t1 & 0x7f7f7f7fu);
    const uint32_t ge = ((x ^ t1) & x) | (~(x ^ t1) & low);
    const uint32_t flags = ((ge & 0x80808080u) * 0x00204081u) >> 28;
    return t >= 255u ? 0u : flags;
}

// One workgroup per tile row (8 pixel rows).
//  1. Pooled 3x3 thresholds of the row's tiles in LDS (k_tile_threshold,
//     reference gpu_detector.cu:237-276).
//  2. Each thread packs 64 pixels of one pixel row into a u64 (1 = white)
//     from two 16-byte loads per 32-pixel group, with the reference's
//     per-32-pixel ROI test (k_binarize :362-391).
//  3. One wave per pixel row turns the masks into black runs with start/end
//     bit masks, popcounts and wave prefix sums: maximal runs, except in
//     image row 0 where runs split at every 32-pixel word as the reference's
//     fragments there are never rejoined (k_find_runs :427-504 emits
//     fragments; k_merge_runs :514 skips row 0, so its same-row joins never
//     run for that row).
//  4. Band-local union-find in LDS over the 8 rows, flattened so every run
//     ends one hop from its band root; band roots are flagged in the run
//     word and only they get component statistics initialised.  A band with
//     more runs than the LDS buffer leaves every run a root and lets k_merge
//     join all of its rows.
// Band ty of frame d (every branch below is uniform across the workgroup).
template <int kChunkWords>
__device__ void binarize_band(const Plan& plan, int cls, const KFrame* d, FrameRef* s_fr, bool ref_here,
                              bool frame_scalars, int ty) {
    const int W = d->width, H = d->height, tw = d->tile_w, th = d->tile_h;
    const int w64 = d->words64, row_cap = d->row_cap, stride = d->stride;
    const int y0 = ty * kTile;
    const int nrows = min(kTile, H - y0);
    const int t = threadIdx.x;
    // Stage capture also keeps the buffers no later kernel reads: the pooled
    // thresholds, the binary image and the runs and labels inside the band.
    const bool capture = plan.capture != 0;
    int32_t* roi = d->row_roi + 4 * ty;
    // The frame's references (first band of the workgroup, wave 0): their
    // inputs are loaded here, computed after the staging loads below.
    const bool ref_wave = ref_here && t < kWave;
    RefInputs ref_in;
    if (ref_wave) ref_in = load_ref_inputs(d, t);
    const int row_lo = kld(d->trow_lo + ty), row_hi = kld(d->trow_hi + ty);

    __shared__ uint8_t thr[1024];
    __shared__ uint32_t s_off[kTile + 1];
    __shared__ uint32_t s_cnt[kTile];  // runs per band row
    // Band arrays in dynamic LDS: the width class's band capacity each.
    extern __shared__ __align__(16) uint32_t band_lds[];
    const int band_cap = plan.class_band_runs[cls];
    uint32_t* b_run = band_lds;
    uint32_t* b_lab = band_lds + band_cap;
    // Short-lived buffers share storage with the band arrays: the tile
    // staging for the pooled thresholds lives in b_run until the runs are
    // written, the pixel masks live in b_lab until the band union starts.
    uint8_t* pool_min = reinterpret_cast<uint8_t*>(b_run);
    uint8_t* pool_max = pool_min + 3 * (tw + 2);
    // Row r's masks at masks + r * w64, its end counts at
    // end_before + r * 2 * w64: 96 * w64 bytes in all, within b_lab's
    // 4 * band_cap (band_cap >= 1.6 * width), as the 6 * (tw + 2) bytes of
    // the pooled staging are within b_run's.
    uint64_t* masks = reinterpret_cast<uint64_t*>(b_lab);
    // Per row and 32-px half word: runs of the row ending before it (pass B),
    // kept until the band union starts, after the masks in b_lab.
    uint16_t* end_before = reinterpret_cast<uint16_t*>(masks + kTile * w64);

    // Pooled 3x3 thresholds for this tile row (k_tile_threshold): the three
    // tile rows are staged in LDS (padded with neutral values), then pooled.
    const float r_tw2 = 1.0f / static_cast<float>(tw + 2);
    for (int i = t; i < 3 * (tw + 2); i += blockDim.x) {
        const int qd = div_small(i, tw + 2, r_tw2);
        const int dy = qd - 1;
        const int xx = i - qd * (tw + 2) - 1;
        const int yy = ty + dy;
        uint8_t mn = 255, mx = 0;
        if (yy >= 0 && yy < th && xx >= 0 && xx < tw) {
            mn = d->tile_min[static_cast<size_t>(yy) * tw + xx];
            mx = d->tile_max[static_cast<size_t>(yy) * tw + xx];
        }
        pool_min[i] = mn;
        pool_max[i] = mx;
    }
    if (ref_wave) {
        const FrameRef r = frame_reference(d, ref_in, t);
        if (t == 0) {
            *s_fr = r;
            // The frame's scalars, once (its first workgroup): the references
            // and the ROI tile-row count (a statistic).
            if (frame_scalars) {
                d->scal[S_REF0] = static_cast<uint32_t>(r.ref0);
                d->scal[S_REF1] = static_cast<uint32_t>(r.ref1);
                d->scal[S_COUNTS + 4] = r.bottom < 0 ? 0u
                    : static_cast<uint32_t>(min(r.bottom + 1, th - 1) - max(r.top - 1, 0) + 1);
            }
        }
    }
    __syncthreads();
    const FrameRef fr = *s_fr;
    const int3 roi3 = tile_row_roi(fr, ty, row_lo, row_hi, tw);
    const int row_on = roi3.x, rs = roi3.y, re = roi3.z;
    // The row's ROI for the later kernels; column 3 (band joined locally)
    // is set once the band's runs are counted.
    if (t == 0) *reinterpret_cast<int4*>(roi) = make_int4(row_on, rs, re, 0);
    const int ref1 = fr.ref1;
    const int floor_contrast = max(kMinTileContrast, (fr.ref0 * 3) / 2);
    for (int tx = t; tx < tw; tx += blockDim.x) {
        int mn = 255, mx = 0;
        for (int dy = 0; dy < 3; ++dy) {
            const uint8_t* rmin = pool_min + dy * (tw + 2) + tx;
            const uint8_t* rmax = pool_max + dy * (tw + 2) + tx;
            mn = min(mn, min(static_cast<int>(rmin[0]), min(static_cast<int>(rmin[1]), static_cast<int>(rmin[2]))));
            mx = max(mx, max(static_cast<int>(rmax[0]), max(static_cast<int>(rmax[1]), static_cast<int>(rmax[2]))));
        }
        const int tv = (mx - mn >= floor_contrast) ? ((mn + mx + 1) >> 1) : ref1;
        thr[tx] = static_cast<uint8_t>(tv);
        if (capture) d->tile_thresh[static_cast<size_t>(ty) * tw + tx] = static_cast<uint8_t>(tv);
    }
    if (!row_on) {
        if (capture) {
            for (int i = t; i < nrows * w64; i += blockDim.x) {
                const int r = i / w64;
                d->bits[static_cast<size_t>(y0 + r) * w64 + (i - r * w64)] = ~0ull;
            }
        }
        if (t < nrows) d->row_count[y0 + t] = 0;
        return;
    }
    if (t < kTile) s_cnt[t] = 0;
    __syncthreads();

    const uint8_t* gray = d->gray;
    const bool vec = ((reinterpret_cast<uintptr_t>(gray) | static_cast<uintptr_t>(stride)) & 15u) == 0;
    // kChunkWords consecutive 64-px words of one row per item (the chunk
    // width launch parameter: 64 or 128 px).
    const int wchunks = (w64 + kChunkWords - 1) / kChunkWords;
    const int items = nrows * wchunks;
    const float r_wchunks = 1.0f / static_cast<float>(wchunks);
    uint32_t frag_acc = 0;
    for (int it = t; it < items; it += blockDim.x) {
        const int r = div_small(it, wchunks, r_wchunks);
        const int y = y0 + r;
#pragma unroll
        for (int cw = 0; cw < kChunkWords; ++cw) {
            const int w = (it - r * wchunks) * kChunkWords + cw;
            if (w >= w64) break;
            // The word's 64 pixels as four 16-byte loads (two per 32-px
            // group inside the row's ROI), all issued before any is used.
            const gbyte* row_px = frame_bytes(gray) + static_cast<size_t>(y) * stride;
            // The pixel left of the word (for its run starts), loaded with
            // the word: thresholded by the same rule, white outside the
            // row's ROI groups.
            const int tile0p = (2 * w - 1) * 4;
            const bool left_on = y != 0 && w > 0 && tile0p <= re && tile0p + 3 >= rs;
            const uint32_t left_px = left_on ? static_cast<uint32_t>(row_px[64 * w - 1]) : 0u;
            bool full[2];
            uint4 px[2][2];
#pragma unroll
            for (int g = 0; g < 2; ++g) {
                const int grp = 2 * w + g;
                const int xs = grp * 32;
                const int tile0 = grp * 4;
                full[g] = vec && xs + 32 <= W && tile0 <= re && tile0 + 3 >= rs;
                px[g][0] = px[g][1] = make_uint4(0u, 0u, 0u, 0u);
                if (full[g]) {
                    px[g][0] = load16(row_px, xs);
                    px[g][1] = load16(row_px, xs + 16);
                }
            }
            uint64_t m = ~0ull;
            for (int g = 0; g < 2; ++g) {
                const int grp = 2 * w + g;
                const int xs = grp * 32;
                if (xs >= W) continue;
                const int tile0 = grp * 4;
                if (!(tile0 <= re && tile0 + 3 >= rs)) continue;  // 32-px ROI test (:376-377)
                const uint8_t* src = gray + static_cast<size_t>(y) * stride + xs;
                uint32_t m32 = 0;
                if (full[g]) {
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        const uint4 v = px[g][q];
                        const uint32_t wd[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const uint32_t tv = thr[(xs >> 3) + 2 * q + (j >> 1)];
                            m32 |= bytes_above(wd[j], tv) << (16 * q + 4 * j);
                        }
                    }
                } else {
                    // Frame edge or unaligned frame: one pixel at a time (kept
                    // rolled so the unrolled form's registers do not spill).
    #pragma unroll 1
                    for (int k = 0; k < 32; ++k) {
                        const int x = xs + k;
                        uint32_t white = 1;
                        if (x < W) white = src[k] > thr[x >> 3] ? 1u : 0u;
                        m32 |= white << k;
                    }
                }
                m = (m & ~(0xffffffffull << (32 * g))) | (static_cast<uint64_t>(m32) << (32 * g));
            }
            masks[r * w64 + w] = m;
            if (capture) d->bits[static_cast<size_t>(y) * w64 + w] = m;
            // The word's run starts, counted here rather than in a pass over
            // the masks: the bit before the word is the pixel left of it,
            // thresholded by the same rule (white outside the row's ROI
            // groups); in image row 0 a fragment also starts at bit 0 and
            // bit 32 of every word (reference k_find_runs :454).  And the
            // reference's fragment count, per 32-px half (:454-468).
            uint64_t prev = 1ull;
            if (y == 0) {
                prev = 1ull | (1ull << 32);
            } else if (left_on) {
                prev = left_px > thr[(64 * w - 1) >> 3] ? 1ull : 0ull;
            }
            atomicAdd(&s_cnt[r], static_cast<uint32_t>(__popcll(~m & ((m << 1) | prev))));
            for (int h = 0; h < 2; ++h) {
                if ((2 * w + h) * 32 >= W) break;
                const uint32_t m32 = static_cast<uint32_t>(m >> (32 * h));
                frag_acc += __popc(~m32 & ((m32 << 1) | 1u));
            }
        }
    }
    __syncthreads();

    // The wave index in a scalar register: rows loop per wave, and the row
    // offsets indexed by it stay scalar.
    const int wave = __builtin_amdgcn_readfirstlane(t / kWave), lane = t % kWave;
    const int nwaves = static_cast<int>(blockDim.x) / kWave;
    // Start/end masks of word w of band row r.
    auto edges = [&](int r, int w, uint64_t* starts, uint64_t* ends) {
        const uint64_t m = masks[r * w64 + w];
        uint64_t prev = w > 0 ? (masks[r * w64 + w - 1] >> 63) : 1ull;
        uint64_t next = w + 1 < w64 ? (masks[r * w64 + w + 1] & 1ull) : 1ull;
        if (y0 + r == 0) {
            // Row 0: a fragment starts at bit 0 of every 32-px word and ends
            // at its last bit (reference k_find_runs :454, :485-487).
            prev = 1ull | (1ull << 32);
            *ends = ~m & ((m >> 1) | (1ull << 63) | (1ull << 31));
        } else {
            *ends = ~m & ((m >> 1) | (next << 63));
        }
        *starts = ~m & ((m << 1) | prev);
    };

    // Row offsets of the band from the counted starts (a wave scan).
    if (t < kWave) {
        const uint32_t c = t < nrows ? s_cnt[t] : 0u;
        uint32_t total;
        const uint32_t ex = wave_excl_scan(c, t, &total);
        if (t <= kTile) s_off[t] = ex;
        if (t < nrows) d->row_count[y0 + t] = c;
    }
    const uint32_t block_frags = wave_sum(frag_acc);
    __syncthreads();
    const uint32_t nb = s_off[nrows];
    const bool local = nb <= static_cast<uint32_t>(band_cap);
    if (lane == 0 && block_frags) atomicAdd(&d->scal[S_COUNTS + 2], block_frags);
    if (t == 0 && nb) atomicAdd(&d->scal[S_MAXRUNS], nb);

    // Pass B: write the runs (LDS when the band fits, else global with every
    // run a root) and initialise the component statistics.
    uint16_t* runs16 = reinterpret_cast<uint16_t*>(d->runs);
    uint16_t* lds16 = reinterpret_cast<uint16_t*>(b_run);
    // The band row of each of the band's first kRowTab runs, in the
    // threshold buffer (no longer read): one LDS read instead of counting
    // row starts wherever a run's row is needed.
    constexpr uint32_t kRowTab = sizeof(thr);
    uint8_t* rowtab = thr;
    for (int r = wave; r < nrows; r += nwaves) {
        const int y = y0 + r;
        const uint32_t base = static_cast<uint32_t>(y) * static_cast<uint32_t>(row_cap);
        const uint32_t lbase = s_off[r];
        uint32_t carry_s = 0, carry_e = 0;
        // One 32-pixel half word per lane, starts and ends extracted in
        // one loop.  The neighbouring halves' edge pixels come from the
        // neighbouring lanes (DPP; the row's ends and halves past it are
        // white, as is everything around image row 0's 32-px fragments,
        // reference k_find_runs :454, :485-487), and one wave scan counts
        // starts and ends together (16 bits each: a pass covers 2048 px).
        const int nh = 2 * w64;
        const uint32_t* masks32 = reinterpret_cast<const uint32_t*>(masks) + r * nh;
        const bool row0 = y == 0;
        uint32_t carry_word = 0xffffffffu;  // the half before the pass
        for (int h0 = 0; h0 < nh; h0 += kWave) {
            const int hw = h0 + lane;
            const uint32_t m32 = hw < nh ? masks32[hw] : 0xffffffffu;
            const uint32_t below = dpp_from_lower(carry_word, m32);
            uint32_t above = dpp_from_upper(0xffffffffu, m32);
            if (lane == kWave - 1 && hw + 1 < nh) above = masks32[hw + 1];
            const uint32_t prev_bit = row0 ? 1u : below >> 31;
            const uint32_t next_bit = row0 ? 1u : above & 1u;
            uint32_t starts = ~m32 & ((m32 << 1) | prev_bit);
            uint32_t ends = ~m32 & ((m32 >> 1) | (next_bit << 31));
            carry_word = __builtin_amdgcn_readlane(m32, kWave - 1);
            uint32_t tot;
            const uint32_t ex = wave_excl_scan(static_cast<uint32_t>(__popc(starts)) |
                                                   (static_cast<uint32_t>(__popc(ends)) << 16), lane, &tot);
            uint32_t ks = carry_s + (ex & 0xffffu);
            uint32_t ke = carry_e + (ex >> 16);
            const uint32_t tot_s = tot & 0xffffu, tot_e = tot >> 16;
            if (hw < nh) end_before[r * 2 * w64 + hw] = static_cast<uint16_t>(ke);
            const int xb = hw * 32;
            if (local) {
                // A run's start and end fill the low and high halves of its
                // LDS word.
                while (starts | ends) {
                    const bool has_s = starts != 0, has_e = ends != 0;
                    const int bs = __ffs(starts) - 1, be = __ffs(ends) - 1;
                    starts &= starts - 1;
                    ends &= ends - 1;
                    if (has_s) {
                        lds16[2 * (lbase + ks)] = static_cast<uint16_t>(xb + bs);
                        if (lbase + ks < kRowTab) rowtab[lbase + ks] = static_cast<uint8_t>(r);
                    }
                    if (has_e) lds16[2 * (lbase + ke) + 1] = static_cast<uint16_t>(xb + be);
                    ks += has_s ? 1u : 0u;
                    ke += has_e ? 1u : 0u;
                }
            } else {
                while (starts | ends) {
                    if (starts) {
                        const int b = __ffs(starts) - 1;
                        starts &= starts - 1;
                        const uint32_t slot = base + ks++;
                        runs16[2 * slot] = static_cast<uint16_t>(xb + b);
                        d->label[slot] = slot;
                        d->area[slot] = 0;
                        uint4* bb = reinterpret_cast<uint4*>(d->bbox + 4 * static_cast<size_t>(slot));
                        *bb = make_uint4(0xffffffffu, 0xffffffffu, 0u, 0u);
                    }
                    if (ends) {
                        const int b = __ffs(ends) - 1;
                        ends &= ends - 1;
                        runs16[2 * (base + ke++) + 1] = static_cast<uint16_t>((xb + b) | 0x8000);
                    }
                }
            }
            carry_s += tot_s;
            carry_e += tot_e;
        }
    }
    if (t == 0) roi[3] = local ? 1 : 0;
    if (!local) return;
    __syncthreads();

    // Band-local union-find (k_merge_runs :525-556 on the band's rows), all
    // row pairs at once: LDS atomics keep it coherent and a band's trees are
    // at most a few hops deep.
    // Row offsets in registers; the row of band run i is the number of row
    // starts at or below i, minus one.
    uint32_t so[kTile + 1];
#pragma unroll
    for (int q = 0; q <= kTile; ++q) so[q] = __builtin_amdgcn_readfirstlane(q <= nrows ? s_off[q] : nb);
    const auto row_of = [&so, rowtab](uint32_t i) {
        if (i < kRowTab) return static_cast<int>(rowtab[i]);
        int r = 0;
#pragma unroll
        for (int q = 1; q < kTile; ++q) r += i >= so[q] ? 1 : 0;
        return r;
    };
    // The first run of the row above that touches run i (x1' >= x0 - 1) is
    // that row's first run plus its runs ending before x0 - 1: the row's end
    // count before x0 - 1's half word plus the end bits below x0 - 1 in it.
    // Computed from the masks (still in b_lab) for a thread's first
    // kFirstAhead runs; later runs (bands of more than 1024 + row-0 runs)
    // search the row above instead.
    constexpr int kFirstAhead = 4;
    int first_ahead[kFirstAhead];
#pragma unroll
    for (int k = 0; k < kFirstAhead; ++k) {
        const uint32_t i = so[1] + t + k * blockDim.x;
        first_ahead[k] = 0;
        if (i < nb) {
            const int r = row_of(i);
            const int p = run_x0(b_run[i]) - 1;
            uint32_t before = 0;
            if (p > 0) {
                const int h = p >> 5, bit = p & 31;
                uint64_t st, en;
                edges(r - 1, h >> 1, &st, &en);
                const uint32_t ends = static_cast<uint32_t>(en >> (32 * (h & 1)));
                before = end_before[(r - 1) * 2 * w64 + h] + static_cast<uint32_t>(__popc(ends & ((1u << bit) - 1u)));
            }
            first_ahead[k] = static_cast<int>(s_off[r - 1] + before);
        }
    }
    __syncthreads();
    for (uint32_t i = t; i < nb; i += blockDim.x) b_lab[i] = i;
    __syncthreads();
    const auto join_above = [&](uint32_t i, int first) {
        const int r = row_of(i);
        const int beg = static_cast<int>(s_off[r]);
        const int x1 = run_x1(b_run[i]);
        // anc: an ancestor of i, its set's root after the last join; a run
        // above whose label is anc is already in the set.
        uint32_t anc = i;
        for (int k = first; k < beg; ++k) {
            if (run_x0(b_run[k]) > x1 + 1) break;
            if (b_lab[k] == anc) continue;
            anc = lds_union(b_lab, anc, static_cast<uint32_t>(k));
        }
    };
#pragma unroll
    for (int k = 0; k < kFirstAhead; ++k) {
        const uint32_t i = so[1] + t + k * blockDim.x;
        if (i < nb) join_above(i, first_ahead[k]);
    }
    for (uint32_t i = so[1] + t + kFirstAhead * blockDim.x; i < nb; i += blockDim.x) {
        const int r = row_of(i);
        const int x0 = run_x0(b_run[i]);
        int lo = static_cast<int>(s_off[r - 1]), hi = static_cast<int>(s_off[r]) - 1, first = hi + 1;
        while (lo <= hi) {
            const int mid = (lo + hi) >> 1;
            if (run_x1(b_run[mid]) >= x0 - 1) {
                first = mid;
                hi = mid - 1;
            } else {
                lo = mid + 1;
            }
        }
        join_above(i, first);
    }
    __syncthreads();
    for (uint32_t i = t; i < nb; i += blockDim.x) b_lab[i] = lds_find(b_lab, i);
    __syncthreads();
    // Write-out, one wave per row, 64 runs at a time.  Global memory gets
    // the runs and labels that later kernels read: the band's first and last
    // rows (k_merge joins them to the neighbouring bands) and every band root
    // (its label is the node the joins and the root walks go through); all
    // runs under stage capture.  The run statistics leave as row items: each
    // stretch of consecutive runs (among the 64) with one band root, folded
    // by one wave scan (its leftmost start is its first run's, its rightmost
    // end its last run's, as a row's runs are x-ordered), so k_stats folds a
    // few dozen items per row instead of every run.  A band root's own run
    // always starts a stretch (no earlier run of the row is in its set).
    for (int r = wave; r < nrows; r += nwaves) {
        const uint32_t rbase = so[r], rn = so[r + 1] - so[r];
        const uint32_t row_slot = static_cast<uint32_t>(y0 + r) * static_cast<uint32_t>(row_cap);
        const bool whole_row = capture || r == 0 || r == nrows - 1;
        uint64_t* row_items = d->items + row_slot;
        uint32_t placed = 0;
        for (uint32_t c = 0; c < rn; c += kWave) {
            const uint32_t k = c + static_cast<uint32_t>(lane);
            const bool valid = k < rn;
            const uint32_t i = rbase + k;
            uint32_t run = 0, root = 0;
            if (valid) {
                run = b_run[i];
                root = b_lab[i];
            }
            const bool is_root = valid && root == i;
            const int rr = row_of(root);
            const uint32_t ridx = root - s_off[rr];  // an LDS read, not a select over so
            if (valid && (whole_row || is_root)) {
                d->runs[row_slot + k] = run | (is_root ? kRootFlag : 0u);
                d->label[row_slot + k] = static_cast<uint32_t>(y0 + rr) * static_cast<uint32_t>(row_cap) + ridx;
            }
            if (is_root) {
                d->area[row_slot + k] = 0;
                uint4* bb = reinterpret_cast<uint4*>(d->bbox + 4 * static_cast<size_t>(row_slot + k));
                *bb = make_uint4(0xffffffffu, 0xffffffffu, 0u, 0u);
            }
            const uint32_t x0 = static_cast<uint32_t>(run_x0(run)), x1 = static_cast<uint32_t>(run_x1(run));
            const uint32_t len = valid ? x1 - x0 + 1u : 0u;
            const uint32_t key = valid ? root : 0xffffffffu;
            // A stretch starts where the band root changes; it ends before
            // the next start or at the last valid lane.
            const bool head = valid && key != dpp_from_lower(0xffffffffu, key);
            const uint64_t heads = __ballot(head);
            const uint32_t incl = wave_incl_scan(len);
            const uint64_t later = lane == kWave - 1 ? 0ull : heads & (~0ull << (lane + 1));
            const int last = static_cast<int>(min(rn - c, static_cast<uint32_t>(kWave))) - 1;
            const int end = later ? __ffsll(static_cast<long long>(later)) - 2 : last;
            const uint32_t end_x1 = static_cast<uint32_t>(__shfl(static_cast<int>(x1), end));
            const uint32_t end_incl = static_cast<uint32_t>(__shfl(static_cast<int>(incl), end));
            if (head) {
                // The item's two words (x0 | end x1 << 13 | (pixels - 1) << 26
                // | root flag << 39 | band row << 40 | index << 43) built in
                // 32-bit halves.
                const uint32_t am1 = end_incl - incl + len - 1u;  // < 2^13
                const uint32_t lo = x0 | (end_x1 << 13) | (am1 << 26);
                const uint32_t hi = (am1 >> 6) | ((is_root ? 1u : 0u) << 7) | (static_cast<uint32_t>(rr) << 8) | (ridx << 11);
                row_items[placed + static_cast<uint32_t>(__popcll(heads & ((1ull << lane) - 1ull)))] =
                    (static_cast<uint64_t>(hi) << 32) | lo;
            }
            placed += static_cast<uint32_t>(__popcll(heads));
        }
        if (lane == 0) d->row_items[y0 + r] = placed;
    }
}

// One workgroup per kBands consecutive bands of one frame (the band-rows
// launch parameter: 8 or 16 pixel rows per workgroup, an execution tiling
// only: the 8x8 threshold tiles and the 8-row union bands are the same),
// the bands in turn, reusing the LDS.  kThreads is 256, or 1024 for a
// launch that fits in one round of 1024-thread workgroups (a single frame,
// a few frames): the per-run phases of a band then run four times as wide.
template <int kChunkWords, int kBands, int kThreads>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(6))) void k_binarize_runs(Plan plan, int cls) {
    int unit;
    const KFrame* d =
        (const KFrame*)plan.frames + frame_of(plan, (kBands == 1 ? SH_CLASS : SH_CLASS_PAIR) + cls, blockIdx.x, &unit);
    if (!frame_head(d).active) return;
    __shared__ FrameRef s_fr;
#pragma unroll 1
    for (int b = 0; b < kBands; ++b) {
        const int ty = unit * kBands + b;
        if (ty >= d->tile_h) break;
        if (b > 0) __syncthreads();  // the previous band's LDS reads are done
        binarize_band<kChunkWords>(plan, cls, d, &s_fr, b == 0, unit == 0 && b == 0, ty);
    }
}

// ---------------------------------------------------------------------------
// (b) union-find over runs
// ---------------------------------------------------------------------------


// Hooks the larger root under the smaller one with atomicMin; labels only
// decrease, so the root of a component is its smallest slot, i.e. its first
// run in raster order (the reference's uf_union, gpu_detector.cu:408-421).
__device__ inline void uf_union(uint32_t* label, uint32_t a, uint32_t b) {
    for (;;) {
        // Both roots by plain reads, the two walks in lockstep so that their
        // loads overlap (a finished walk re-reads its root).
        uint32_t pa = label[a], pb = label[b];
        while (pa != a || pb != b) {
            a = pa;
            b = pb;
            pa = label[a];
            pb = label[b];
        }
        if (a == b) return;
        if (a > b) {
            const uint32_t tmp = a;
            a = b;
            b = tmp;
        }
        const uint32_t old = atomicMin(&label[b], a);
        if (old == b) return;
        b = old;
    }
}

// k_merge LDS: a pair hash set of kPairSlots entries (8 KB) and, in dynamic
// LDS, the two boundary rows' runs and labels staged for up to
// Plan::class_merge_stage runs per row (rows with more runs are read from
// global memory).
constexpr int kPairSlots = 1024;
constexpr int kMergeStageNarrow = 256;
// Rows staged by the 256-thread form: up to 768 runs (noisy 4096-px rows
// hold about 500, at most ~620), 20 KB of LDS with the pair table, seven
// resident workgroups per CU; longer rows are read in place.
constexpr int kMergeStageWide = 768;

// Records the (smaller, larger) label pair in the LDS hash set of kSlots
// entries, or joins the pair directly when 16 probes find no room.
template <int kSlots>
__device__ inline void merge_pair(unsigned long long* pairs, uint32_t* label, uint32_t a, uint32_t b) {
    const unsigned long long key = (static_cast<unsigned long long>(min(a, b)) << 32) | max(a, b);
    const uint32_t h = static_cast<uint32_t>((key * 0x9e3779b97f4a7c15ull) >> 54);
    for (int probe = 0; probe < 16; ++probe) {
        const int e = static_cast<int>((h + static_cast<uint32_t>(probe)) & (kSlots - 1));
        unsigned long long old = pairs[e];
        if (old == ~0ull) old = atomicCAS(&pairs[e], ~0ull, key);
        if (old == ~0ull || old == key) return;
    }
    uf_union(label, a, b);
}

// Joins run i of a row (runs cur, labels cur_lab) to the touching runs of
// the row above (prev, prev_lab, pn runs): 8-connectivity, runs touching
// [x0 - 1, x1 + 1] (reference k_merge_runs :509-557).  The arrays are LDS
// copies or the global rows.
template <int kSlots, class RunPtr, class LabPtr>
__device__ inline void merge_run(unsigned long long* pairs, uint32_t* label, RunPtr cur, LabPtr cur_lab,
                                 RunPtr prev, LabPtr prev_lab, int pn, int i) {
    const uint32_t run = cur[i];
    const int x0 = run_x0(run);
    const int x1 = run_x1(run);
    int lo = 0, hi = pn - 1, first = pn;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        if (run_x1(prev[mid]) >= x0 - 1) {
            first = mid;
            hi = mid - 1;
        } else {
            lo = mid + 1;
        }
    }
    const uint32_t a = cur_lab[i];
    uint32_t last_b = 0xffffffffu;
    for (int k = first; k < pn; ++k) {
        if (run_x0(prev[k]) > x1 + 1) break;
        const uint32_t b = prev_lab[k];
        if (b == a || b == last_b) continue;
        last_b = b;
        merge_pair<kSlots>(pairs, label, a, b);
    }
}

// One workgroup per band, one launch per width class.  Its first row is
// joined to the last row of the band above (8-connectivity: runs touching
// [x0-1, x1+1]; reference k_merge_runs :509-557); a band that was not joined
// in LDS also joins all of its own rows here.  Each pair of rows is staged
// in LDS (runs and labels, coalesced loads) and the row's runs are spread
// over the whole workgroup.  The (label, label) pairs of all touching runs are
// first de-duplicated in an LDS hash set, so a large component whose runs
// cross the boundary hundreds of times costs one union per distinct pair
// instead of one per run.  Labels read at any time are ancestors of the
// runs' roots, so the pairs join the same components whenever they are
// read.  Runs of one row never need same-row joins: they are maximal, or in
// row 0 they are fragments the reference leaves unjoined.
// kThreads 128 with a 512-entry pair table and rows of up to kMergeStageNarrow
// runs staged serves the narrow width class in launches over more than two
// frames: 16 resident workgroups per CU instead of 8 (a band's ~140 runs per
// row at about one per thread), so the 64-frame 1280x800 batch needs fewer
// rounds of this latency-bound chain (52 -> 38 us); a single frame's bands
// all fit in one round of the 256-thread form.
template <int kThreads, int kSlots>
__global__ __launch_bounds__(kThreads) void k_merge(Plan plan, int cls) {
    int ty;
    const KFrame* d = (const KFrame*)plan.frames + frame_of(plan, SH_CLASS + cls, blockIdx.x, &ty);
    if (!frame_head(d).active) return;
    const int H = d->height;
    const int y0 = ty * kTile;
    const int nrows = min(kTile, H - y0);
    const int first_row = ty == 0 ? 1 : 0;
    // The band's ROI word and its first row pair's run counts, loaded
    // together (the counts do not depend on the ROI; frames have at least
    // two rows).
    typedef int i4v __attribute__((ext_vector_type(4)));
    const i4v roi = kld(reinterpret_cast<const i4v*>(d->row_roi + 4 * ty));
    const int roi_on = roi[0], roi_local = roi[3];
    const int n_first = static_cast<int>(kld(d->row_count + y0 + first_row));
    const int pn_first = static_cast<int>(kld(d->row_count + y0 + first_row - 1));
    if (!roi_on) return;
    const bool local = roi_local != 0;
    const int last_row = local ? 1 : nrows;  // exclusive, band-relative
    if (first_row >= last_row) return;
    const int t = threadIdx.x;
    __shared__ unsigned long long pairs[kSlots];
    extern __shared__ uint32_t merge_stage[];
    const int stage_cap = min(plan.class_merge_stage[cls], kThreads == 256 ? kMergeStageWide : kMergeStageNarrow);
    uint32_t* s_cur = merge_stage;
    uint32_t* s_cur_lab = merge_stage + stage_cap;
    uint32_t* s_prev = merge_stage + 2 * stage_cap;
    uint32_t* s_prev_lab = merge_stage + 3 * stage_cap;
    for (int e = t; e < kSlots; e += kThreads) pairs[e] = ~0ull;
    const uint32_t row_cap = static_cast<uint32_t>(d->row_cap);
    uint32_t* label = d->label;
    __syncthreads();
    // Row by row (the band's first row against the band above; every row of
    // a band not joined in LDS against the row above it), the whole
    // workgroup on each pair of rows: both rows' runs and labels staged in
    // LDS by coalesced loads when they fit, else read from global memory.
    for (int r = first_row; r < last_row; ++r) {
        const int y = y0 + r;
        const int n = r == first_row ? n_first : static_cast<int>(kld(d->row_count + y));
        const int pn = r == first_row ? pn_first : static_cast<int>(kld(d->row_count + y - 1));
        if (n == 0 || pn == 0) continue;
        const uint32_t cur_base = static_cast<uint32_t>(y) * row_cap;
        const uint32_t prev_base = static_cast<uint32_t>(y - 1) * row_cap;
        if (n <= stage_cap && pn <= stage_cap) {
            __syncthreads();  // the previous pair's readers are done
            // Both rows in one pass: a thread's four loads are issued
            // before any of its LDS stores waits for them.
            for (int i = t; i < max(n, pn); i += kThreads) {
                uint32_t cr = 0, cl = 0, pr = 0, pl = 0;
                if (i < n) {
                    cr = d->runs[cur_base + i];
                    cl = label[cur_base + i];
                }
                if (i < pn) {
                    pr = d->runs[prev_base + i];
                    pl = label[prev_base + i];
                }
                if (i < n) {
                    s_cur[i] = cr;
                    s_cur_lab[i] = cl;
                }
                if (i < pn) {
                    s_prev[i] = pr;
                    s_prev_lab[i] = pl;
                }
            }
            __syncthreads();
            for (int i = t; i < n; i += kThreads)
                merge_run<kSlots>(pairs, label, s_cur, s_cur_lab, s_prev, s_prev_lab, pn, i);
        } else {
            for (int i = t; i < n; i += kThreads)
                merge_run<kSlots>(pairs, label, d->runs + cur_base, label + cur_base, d->runs + prev_base,
                                  label + prev_base, pn, i);
        }
    }
    __syncthreads();
    // Distinct pairs are gathered to the front of the table (each thread
    // holds its slots' keys in registers meanwhile), then joined one per
    // thread.
    __shared__ uint32_t s_npairs;
    if (t == 0) s_npairs = 0;
    __syncthreads();
    unsigned long long mine[kSlots / kThreads];
    int nmine = 0;
    for (int e = t; e < kSlots; e += kThreads) {
        const unsigned long long key = pairs[e];
        if (key != ~0ull) mine[nmine++] = key;
    }
    __syncthreads();
    for (int q = 0; q < nmine; ++q) pairs[atomicAdd(&s_npairs, 1u)] = mine[q];
    __syncthreads();
    const uint32_t npairs = s_npairs;
    for (uint32_t k = t; k < npairs; k += kThreads) {
        const unsigned long long key = pairs[k];
        uf_union(label, static_cast<uint32_t>(key >> 32), static_cast<uint32_t>(key & 0xffffffffu));
    }
}

constexpr int kStatSlots = 256;
constexpr uint32_t kEmpty = 0xffffffffu;

// One workgroup per band of 8 rows, after all joins, over the band's band
// roots and items (k_binarize_runs: stretches of consecutive runs of one
// row and one band-local component, folded; a band joined in global memory
// has every run an item and a band root of its own).
//  1. Each band root walks to its component root, which is written to its
//     label; a band root that is its own component root is a component
//     root: its row is the component's top row (the root is the component's
//     first run in raster order) and it joins the band's root list (the
//     roots k_collect filters).
//  2. Every item takes its band root's component root; its pixel count
//     goes into an LDS table keyed by root and its extent into a per-row LDS
//     table keyed by root (the row's components with their leftmost start
//     and rightmost end).  The root table takes each root's x extent and
//     last row from the row tables and is flushed with one set of global
//     atomics per (band, root) (reference k_region_stats :562-586 does one
//     atomic per run); the row tables are written out for k_fit's row
//     extremes (1 << kRowLog2 slots per row).
// kThreads: 256, or 512 for the 256-slot tables (frames over 2560 px), whose
// 33 KB of LDS leaves four workgroups per CU.
template <int kRowLog2, int kThreads>
__global__ __launch_bounds__(kThreads) void k_stats(Plan plan) {
    constexpr int kRowSlots = 1 << kRowLog2;
    int ty;
    const KFrame* d = (const KFrame*)plan.frames + frame_of(plan, SH_CLASS + (kRowLog2 - kRowExtLog2Min), blockIdx.x, &ty);
    if (!frame_head(d).active) return;
    typedef int i4v __attribute__((ext_vector_type(4)));
    const i4v roi = kld(reinterpret_cast<const i4v*>(d->row_roi + 4 * ty));
    if (!roi[0]) {
        if (threadIdx.x == 0) d->band_nroots[ty] = 0;  // an inactive band has no runs
        return;
    }
    const bool local = roi[3] != 0;
    const int H = d->height;
    const int y0 = ty * kTile;
    const int nrows = min(kTile, H - y0);
    const uint32_t row_cap = static_cast<uint32_t>(d->row_cap);
    const int t = threadIdx.x;
    __shared__ uint32_t hkey[kStatSlots], harea[kStatSlots], hx0[kStatSlots], hx1[kStatSlots], hy1[kStatSlots];
    __shared__ uint32_t s_off[kTile + 1];
    __shared__ uint32_t s_flush[kStatSlots];
    __shared__ uint32_t s_nflush, s_nroots, s_nwalk;
    // Per band row, the components in the row with their leftmost start and
    // rightmost end (k_fit's row extremes).
    __shared__ uint32_t s_rkey[kTile][kRowSlots], s_rx0[kTile][kRowSlots], s_rx1[kTile][kRowSlots];
    __shared__ uint32_t s_rfull[kTile];
    for (int e = t; e < kTile * kRowSlots; e += blockDim.x) {
        (&s_rkey[0][0])[e] = kEmpty;
        (&s_rx0[0][0])[e] = 0xffffffffu;
        (&s_rx1[0][0])[e] = 0u;
    }
    if (t < kTile) s_rfull[t] = 0;
    for (int e = t; e < kStatSlots; e += blockDim.x) {
        hkey[e] = kEmpty;
        harea[e] = 0;
        hx0[e] = 0xffffffffu;
        hx1[e] = 0;
        hy1[e] = 0;
    }
    if (t < kWave) {
        // Item offsets of the band's rows (a band joined in global memory:
        // its runs): one lane per row, a wave prefix sum.
        const uint32_t c = t < nrows ? (local ? d->row_items[y0 + t] : d->row_count[y0 + t]) : 0u;
        uint32_t total;
        const uint32_t ex = wave_excl_scan(c, t, &total);
        if (t <= kTile) s_off[t] = ex;
        if (t == 0) {
            s_nflush = 0;
            s_nroots = 0;
            s_nwalk = 0;
        }
    }
    __syncthreads();
    uint32_t so[kTile + 1];
#pragma unroll
    for (int q = 0; q <= kTile; ++q) so[q] = s_off[q];
    const uint32_t band_base = static_cast<uint32_t>(y0) * row_cap;
    const uint32_t nitems = so[kTile];
    uint32_t* label = d->label;
    // Item j of the band: x0, x1, pixels, band row, and its band root's slot
    // (flagged on the band root's own item).  A band joined in global memory
    // has every run an item and a band root of its own.
    struct Item {
        uint32_t x0, x1, area, broot;
        int row;
        bool flag;
    };
    const auto read_item = [&](uint32_t j) {
        Item it;
        int r = 0;
#pragma unroll
        for (int q = 1; q < kTile; ++q) r += j >= so[q] ? 1 : 0;
        const uint32_t slot = band_base + static_cast<uint32_t>(r) * row_cap + (j - so[r]);
        it.row = r;
        if (local) {
            const uint64_t v = d->items[slot];
            it.x0 = static_cast<uint32_t>(v) & 0x1fffu;
            it.x1 = static_cast<uint32_t>(v >> 13) & 0x1fffu;
            it.area = (static_cast<uint32_t>(v >> 26) & 0x1fffu) + 1u;
            it.flag = ((v >> 39) & 1u) != 0;
            it.broot = band_base + static_cast<uint32_t>((v >> 40) & 7u) * row_cap + static_cast<uint32_t>(v >> 43);
        } else {
            const uint32_t run = d->runs[slot];
            it.x0 = static_cast<uint32_t>(run_x0(run));
            it.x1 = static_cast<uint32_t>(run_x1(run));
            it.area = it.x1 - it.x0 + 1u;
            it.flag = true;
            it.broot = slot;
        }
        return it;
    };

    // Band roots walk to their root.  Every join is done (k_merge), so a
    // label names an ancestor and the walk only reads; lowering the visited
    // labels on the way (path splitting with atomicMin) was measured slower:
    // the chains are short and the atomics cost more than they save.  A
    // thread advances two walks in lockstep so their loads overlap.  Walks
    // only visit band roots (a band root's label always names another band
    // root).  On return root[u] is the component root of band root slot[u],
    // which is written to its label (a plain store: only this walk writes
    // that label, and concurrent walks through it read either value, both
    // ancestors).
    constexpr int kItems = 2;
    const auto walk = [&](const uint32_t* slot, const bool* on, uint32_t* root) {
        uint32_t x[kItems], p[kItems];
#pragma unroll
        for (int u = 0; u < kItems; ++u) {
            x[u] = on[u] ? slot[u] : 0u;  // x == p ends a walk
            p[u] = on[u] ? label[slot[u]] : 0u;
        }
        bool live = true;
        while (live) {
            live = false;
#pragma unroll
            for (int u = 0; u < kItems; ++u) {
                if (p[u] == x[u]) continue;
                x[u] = p[u];
                p[u] = label[p[u]];
                live |= p[u] != x[u];
            }
        }
#pragma unroll
        for (int u = 0; u < kItems; ++u) {
            if (on[u] && x[u] != slot[u]) label[slot[u]] = x[u];
            root[u] = x[u];
        }
    };
    // After its walk: a band root that is its own component root is a
    // component root, whose row is the component's top row (the root is the
    // component's first run in raster order); it joins the band's root list
    // (the roots k_collect filters).
    const auto settle = [&](uint32_t slot, uint32_t root) {
        if (root == slot) {
            d->bbox[4 * static_cast<size_t>(root) + 1] = slot / row_cap;
            d->roots[band_base + atomicAdd(&s_nroots, 1u)] = root;
        }
    };
    // The band roots (a few dozen per band) are gathered into an LDS list
    // and walk in one batch, so a band pays the walk's load latency once.
    // Band roots beyond the list's capacity walk where they are found.
    constexpr uint32_t kWalkCap = kItems * kThreads;
    __shared__ uint32_t s_wslot[kWalkCap];
    for (uint32_t j0 = 0; j0 < nitems; j0 += blockDim.x) {
        const uint32_t j = j0 + static_cast<uint32_t>(t);
        bool spill = false;
        uint32_t spill_slot = 0;
        if (j < nitems) {
            const Item it = read_item(j);
            if (it.flag) {
                const uint32_t k = atomicAdd(&s_nwalk, 1u);
                if (k < kWalkCap) {
                    s_wslot[k] = it.broot;
                } else {
                    spill = true;
                    spill_slot = it.broot;
                }
            }
        }
        if (__any(spill)) {
            const uint32_t sl[kItems] = {spill_slot, 0u};
            const bool on[kItems] = {spill, false};
            uint32_t rt[kItems];
            walk(sl, on, rt);
            if (spill) settle(spill_slot, rt[0]);
        }
    }
    __syncthreads();
    const uint32_t nwalk = min(s_nwalk, kWalkCap);
    if (static_cast<uint32_t>(t) < nwalk) {
        uint32_t sl[kItems], rt[kItems];
        bool on[kItems];
#pragma unroll
        for (int u = 0; u < kItems; ++u) {
            const uint32_t k = t + u * blockDim.x;
            on[u] = k < nwalk;
            sl[u] = on[u] ? s_wslot[k] : 0u;
        }
        walk(sl, on, rt);
#pragma unroll
        for (int u = 0; u < kItems; ++u) {
            if (on[u]) settle(sl[u], rt[u]);
        }
    }
    __syncthreads();

    // Every item is accounted under its component root.  Pixels per root
    // here; a root's x extent and last row are folded from the row tables at
    // the end (or straight into global memory for an item whose row table is
    // full).  A root that finds no table slot goes to global memory too.
    const auto root_slot = [&](uint32_t root, bool insert) {
        const uint32_t h = (root * 2654435761u) >> 24;
        for (int probe = 0; probe < 32; ++probe) {
            const int e = static_cast<int>((h + static_cast<uint32_t>(probe)) & (kStatSlots - 1));
            uint32_t k = hkey[e];
            if (k == kEmpty && insert) k = atomicCAS(&hkey[e], kEmpty, root);
            if (k == root || (k == kEmpty && insert)) return e;
            if (k == kEmpty) return -1;
        }
        return -1;
    };
    const auto fold = [&](uint32_t root, uint32_t area) {
        const int e = root_slot(root, true);
        if (e >= 0) atomicAdd(&harea[e], area); else atomicAdd(&d->area[root], area);
    };
    const auto fold_extent = [&](uint32_t root, uint32_t x0, uint32_t x1, uint32_t y) {
        const int e = root_slot(root, false);
        if (e >= 0) {
            atomicMin(&hx0[e], x0);
            atomicMax(&hx1[e], x1);
            atomicMax(&hy1[e], y);
        } else {
            atomicMin(&d->bbox[4 * static_cast<size_t>(root) + 0], x0);
            atomicMax(&d->bbox[4 * static_cast<size_t>(root) + 2], x1);
            atomicMax(&d->bbox[4 * static_cast<size_t>(root) + 3], y);
        }
    };
    const auto fold_row = [&](uint32_t root, int r, uint32_t x0, uint32_t x1) {
        const uint32_t h = (root * 2654435761u) >> (32 - kRowLog2);
        for (int probe = 0; probe < kRowSlots; ++probe) {
            const int e = static_cast<int>((h + static_cast<uint32_t>(probe)) & (kRowSlots - 1));
            uint32_t k = s_rkey[r][e];
            if (k == kEmpty) k = atomicCAS(&s_rkey[r][e], kEmpty, root);
            if (k == kEmpty || k == root) {
                atomicMin(&s_rx0[r][e], x0);
                atomicMax(&s_rx1[r][e], x1);
                return;
            }
        }
        s_rfull[r] = 1;
        atomicMin(&d->bbox[4 * static_cast<size_t>(root) + 0], x0);
        atomicMax(&d->bbox[4 * static_cast<size_t>(root) + 2], x1);
        atomicMax(&d->bbox[4 * static_cast<size_t>(root) + 3], static_cast<uint32_t>(y0 + r));
    };
    for (uint32_t j = t; j < nitems; j += blockDim.x) {
        const Item it = read_item(j);
        // The band root's label is its component root, written by its walk
        // before the barrier above; it is read past the vector L1, where the
        // walks' plain loads may have left the label's old value.
        const uint32_t root = __hip_atomic_load(&label[it.broot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        fold(root, it.area);
        fold_row(root, it.row, it.x0, it.x1);
    }
    __syncthreads();
    if (t == 0) d->band_nroots[ty] = s_nroots;
    // Each row's occupied entries, packed to the front of the row's slots in
    // slot order (a row holds 9 components on average: the empty slots are
    // not written); the row's count, or kRowTableFull when its table
    // overflowed, goes to row_ext_full.  One wave per row.
    {
        const int wave = t / kWave, lane = t % kWave;
        const int nwaves = static_cast<int>(blockDim.x) / kWave;
        for (int r = wave; r < nrows; r += nwaves) {
            uint64_t* row_out = d->row_ext + (static_cast<size_t>(y0 + r) << kRowLog2);
            uint32_t placed = 0;
#pragma unroll
            for (int e0 = 0; e0 < kRowSlots; e0 += kWave) {
                const int e = e0 + lane;
                const uint32_t key = s_rkey[r][e];
                const bool occupied = key != kEmpty;
                const uint64_t m = __ballot(occupied);
                if (occupied) {
                    const uint32_t rx0 = s_rx0[r][e], rx1 = s_rx1[r][e];
                    const uint32_t pos = placed + static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)));
                    row_out[pos] = (static_cast<uint64_t>(key) << 32) | (rx0 << 16) | rx1;
                    fold_extent(key, rx0, rx1, static_cast<uint32_t>(y0 + r));
                }
                placed += static_cast<uint32_t>(__popcll(m));
            }
            if (lane == 0) d->row_ext_full[y0 + r] = s_rfull[r] ? kRowTableFull : placed;
        }
    }
    __syncthreads();
    // Occupied table entries are compacted first so every thread flushes
    // about the same number of roots.
    for (int e = t; e < kStatSlots; e += blockDim.x) {
        if (hkey[e] != kEmpty) s_flush[atomicAdd(&s_nflush, 1u)] = static_cast<uint32_t>(e);
    }
    __syncthreads();
    const uint32_t nflush = s_nflush;
    for (uint32_t k = t; k < nflush; k += blockDim.x) {
        const uint32_t e = s_flush[k];
        const uint32_t root = hkey[e];
        atomicAdd(&d->area[root], harea[e]);
        atomicMin(&d->bbox[4 * static_cast<size_t>(root) + 0], hx0[e]);
        atomicMax(&d->bbox[4 * static_cast<size_t>(root) + 2], hx1[e]);
        atomicMax(&d->bbox[4 * static_cast<size_t>(root) + 3], hy1[e]);
    }
}

// ---------------------------------------------------------------------------
// (c) candidates
// ---------------------------------------------------------------------------

// One workgroup per band over the band's component roots (listed by
// k_stats): roots that pass the tag-border filter append the sort key
// (~area << 32 | root slot): ascending keys are largest first, ties by
// raster order (reference k_collect_candidates :591-638, which visits every
// run and keeps the roots).
// The sort key of component root `idx` when it passes the tag-border
// filter (reference k_collect_candidates :600-636).
__device__ inline bool candidate_key(const KFrame* d, uint32_t idx, uint64_t* key) {
    const uint32_t row_cap = static_cast<uint32_t>(d->row_cap);
    const int W = d->width, H = d->height;
    const int y = static_cast<int>(idx / row_cap);
    const int a = static_cast<int>(d->area[idx]);
    if (a < d->min_cluster) return false;
    const uint4 bb = *reinterpret_cast<const uint4*>(d->bbox + 4 * static_cast<size_t>(idx));
    const int x0 = static_cast<int>(bb.x), y0 = y;
    const int x1 = static_cast<int>(bb.z), y1 = static_cast<int>(bb.w);
    const int bw = x1 - x0 + 1;
    const int bh = y1 - y0 + 1;
    if (bw < d->min_tag || bh < d->min_tag) return false;
    if (x0 <= 0 || y0 <= 0 || x1 >= W - 1 || y1 >= H - 1) return false;
    if (bw > 6 * bh || bh > 6 * bw) return false;
    const float fill = static_cast<float>(a) / static_cast<float>(bw * bh);
    if (fill < 0.18f || fill > 0.95f) return false;
    *key = (static_cast<uint64_t>(~static_cast<uint32_t>(a)) << 32) | idx;
    return true;
}

// One wave per band: a band lists a few dozen component roots, and 64-thread
// workgroups put every band of the 64-frame batch in one round.
__global__ __launch_bounds__(64) void k_collect(Plan plan) {
    int ty;
    const KFrame* d = (const KFrame*)plan.frames + frame_of(plan, SH_TILE_ROW, blockIdx.x, &ty);
    if (!frame_head(d).active) return;
    const uint32_t base = static_cast<uint32_t>(ty * kTile) * static_cast<uint32_t>(d->row_cap);
    const uint32_t n = kld(d->band_nroots + ty);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        uint64_t key;
        if (!candidate_key(d, d->roots[base + i], &key)) continue;
        const uint32_t pos = atomicAdd(&d->scal[S_CAND_ACC], 1u);
        d->cand_keys[pos] = key;
    }
}

__device__ inline void bitonic_sort_lds(uint64_t* keys, int n) {
    for (int k = 2; k <= n; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const uint64_t a = keys[i], b = keys[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        keys[i] = b;
                        keys[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// One workgroup per frame: sorts the candidate keys and keeps the first
// max_candidates (reference k_order_candidates :642-654, with the in-bucket
// order fixed to (area desc, root slot asc)).  Lists longer than the LDS
// buffer first find the key of rank max_candidates by an 8-bit radix select.
__global__ __launch_bounds__(256) void k_select(Plan plan) {
    const KFrame* d = (const KFrame*)plan.frames + blockIdx.x;
    if (!frame_head(d).active) return;
    __shared__ uint64_t keys[kSortCap];
    __shared__ uint32_t hist[256];
    __shared__ uint64_t s_prefix, s_mask;
    __shared__ uint32_t s_rank, s_fill;
    const int t = threadIdx.x;
    const uint32_t n = kld(d->scal + S_CAND_ACC);
    const uint32_t nsel = min(n, static_cast<uint32_t>(d->max_candidates));

    __shared__ uint32_t s_head, s_tail, s_nhead;
    // All keys fit one per thread: the head count needs no sorted order.
    const bool early = n <= blockDim.x;
    uint32_t m = n;
    if (n > static_cast<uint32_t>(kSortCap)) {
        if (t == 0) {
            s_prefix = 0;
            s_mask = 0;
            s_rank = nsel;  // 1-based rank of the last kept key
            s_fill = 0;
        }
        __syncthreads();
        for (int shift = 56; shift >= 0; shift -= 8) {
            for (int i = t; i < 256; i += blockDim.x) hist[i] = 0;
            __syncthreads();
            const uint64_t prefix = s_prefix, mask = s_mask;
            for (uint32_t i = t; i < n; i += blockDim.x) {
                const uint64_t key = d->cand_keys[i];
                if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
            }
            __syncthreads();
            if (t == 0) {
                uint32_t cum = 0, rank = s_rank;
                int digit = 255;
                for (int b = 0; b < 256; ++b) {
                    if (cum + hist[b] >= rank) {
                        digit = b;
                        break;
                    }
                    cum += hist[b];
                }
                s_rank = rank - cum;
                s_prefix = prefix | (static_cast<uint64_t>(digit) << shift);
                s_mask = mask | (255ull << shift);
            }
            __syncthreads();
        }
        const uint64_t limit = s_prefix;
        for (uint32_t i = t; i < n; i += blockDim.x) {
            const uint64_t key = d->cand_keys[i];
            if (key <= limit) {
                const uint32_t pos = atomicAdd(&s_fill, 1u);
                if (pos < static_cast<uint32_t>(kSortCap)) keys[pos] = key;
            }
        }
        __syncthreads();
        m = min(s_fill, static_cast<uint32_t>(kSortCap));
    } else {
        for (uint32_t i = t; i < n; i += blockDim.x) keys[i] = d->cand_keys[i];
    }
    if (m <= blockDim.x) {
        // A key per thread (the usual case: a frame keeps a few hundred
        // candidates at most): each key's rank is the number of smaller
        // keys, all distinct (the low word is the root's slot), so one pass
        // and one barrier place them.
        __syncthreads();
        const uint64_t mine = static_cast<uint32_t>(t) < m ? keys[t] : ~0ull;
        if (early) {
            // Every key is here, one per thread: the head count (the
            // selection's keys of at least head_area pixels, i.e. all such
            // keys up to nsel) and the work-list reservations go out before
            // the ranking, which hides the atomics' round trip.
            const bool big = static_cast<uint32_t>(t) < m &&
                             ~static_cast<uint32_t>(mine >> 32) >= static_cast<uint32_t>(plan.tune.head_area);
            const uint32_t nbig = min(static_cast<uint32_t>(__syncthreads_count(big)), nsel);
            const uint32_t nhead = min(min(nsel, static_cast<uint32_t>(kHeadSlots)),
                                       max(nbig, static_cast<uint32_t>(plan.tune.head_ranks)));
            if (t == 0) {
                s_nhead = nhead;
                s_head = atomicAdd(&plan.work[kWorkHeadCount], nhead);
                s_tail = atomicAdd(&plan.work[kWorkCount], nsel - nhead);
            }
        }
        uint32_t rank = 0;
        for (uint32_t j = 0; j < m; ++j) rank += keys[j] < mine ? 1u : 0u;
        __syncthreads();
        if (static_cast<uint32_t>(t) < m) keys[rank] = mine;
        __syncthreads();
    } else {
        int p2 = 1;
        while (p2 < static_cast<int>(m)) p2 <<= 1;
        for (int i = static_cast<int>(m) + t; i < p2; i += blockDim.x) keys[i] = ~0ull;
        __syncthreads();
        bitonic_sort_lds(keys, p2);
    }
    // The frame's largest candidates go to the head of the batch list, so
    // that the longest fits of the batch start first: its first head_ranks
    // and every other one of at least head_area pixels (a prefix of the
    // sorted keys, whose high word is the complemented area).
    if (!early) {
        uint32_t nbig = 0;
        for (uint32_t i0 = 0; i0 < nsel; i0 += blockDim.x) {
            const uint32_t i = i0 + t;
            const bool big = i < nsel && ~static_cast<uint32_t>(keys[i] >> 32) >= static_cast<uint32_t>(plan.tune.head_area);
            nbig += static_cast<uint32_t>(__syncthreads_count(big));
        }
        const uint32_t nhead = min(min(nsel, static_cast<uint32_t>(kHeadSlots)),
                                   max(nbig, static_cast<uint32_t>(plan.tune.head_ranks)));
        if (t == 0) {
            s_nhead = nhead;
            s_head = atomicAdd(&plan.work[kWorkHeadCount], nhead);
            s_tail = atomicAdd(&plan.work[kWorkCount], nsel - nhead);
        }
    }
    __syncthreads();
    const uint32_t nhead = s_nhead;
    uint32_t* head = plan.work + kWorkItems;
    uint32_t* tail = head + kHeadSlots * plan.nframes;
    for (uint32_t i = t; i < nsel; i += blockDim.x) {
        d->candidates[i] = static_cast<uint32_t>(keys[i]);
        const uint32_t code = (blockIdx.x << 10) | i;
        if (i < nhead) head[s_head + i] = code; else tail[s_tail + i - nhead] = code;
    }
    // Slots past the selection read as "no quad" and "not decoded".
    const uint32_t maxc = static_cast<uint32_t>(d->max_candidates);
    for (uint32_t i = nsel * 8 + t; i < maxc * 8; i += blockDim.x) {
        d->quads[i] = __int_as_float(0x7fc00000);
        d->hull_quads[i] = __int_as_float(0x7fc00000);
    }
    for (uint32_t i = nsel + t; i < maxc; i += blockDim.x) d->dec[i].id = -1;
    if (t == 0) {
        d->scal[S_COUNTS + 0] = n;
        d->scal[S_NCAND] = nsel;
        d->scal[S_CAND_ACC] = 0;
    }
}

// ---------------------------------------------------------------------------
// (d) quad fit
// ---------------------------------------------------------------------------


// Bytes of a frame from its base to the last pixel, (height - 1) * stride +
// width.
__device__ inline uint32_t frame_span(int width, int height, int stride) {
    return static_cast<uint32_t>(height - 1) * static_cast<uint32_t>(stride) + static_cast<uint32_t>(width);
}

// Pixels o and o + 1 of a frame (low byte first) from one 8-byte load with
// no branch, so that several pairs' loads can be in flight together.  The
// window is the 4-byte-aligned one at or below pixel o, moved down one word
// where its second word would start past the frame's last byte: both of its
// words then hold a frame byte (bytes of those words outside the frame lie
// in the same page as a frame byte and are discarded), and the pair sits at
// byte 0-6 of it (frames are at least 16x16, so the moved window still
// starts at or after the frame's first word).  `span` is the frame's byte
// extent from its first pixel through its last.
__device__ inline uint32_t pixel_pair(const gbyte* image, uint32_t o, uint32_t span) {
    const uint32_t mis = static_cast<uint32_t>(reinterpret_cast<uintptr_t>(image)) & 3u;
    const gbyte* base = image - mis;  // the frame's first word
    const uint32_t p = o + mis;       // pixel o, counted from base
    const uint32_t at = min(p & ~3u, ((span - 1u + mis) & ~3u) - 4u);
    const uint64_t w = *reinterpret_cast<const gword64*>(base + at);
    return static_cast<uint32_t>(w >> (8u * (p - at))) & 0xffffu;
}

// sample_bilinear split into address and value halves so the loads of
// several samples can be in flight at once; tap_value evaluates the
// reference's bilinear expression.
struct Tap {
    uint32_t off;  // byte offset of the top-left neighbour (0 when outside)
    float fx, fy;
    bool ok;
};

__device__ inline Tap make_tap(int width, int height, int stride, float x, float y) {
    Tap t;
    t.ok = x >= 0.0f && y >= 0.0f && x < static_cast<float>(width - 1) && y < static_cast<float>(height - 1);
    const int ix = t.ok ? static_cast<int>(x) : 0;
    const int iy = t.ok ? static_cast<int>(y) : 0;
    t.fx = x - static_cast<float>(ix);
    t.fy = y - static_cast<float>(iy);
    t.off = static_cast<uint32_t>(iy) * static_cast<uint32_t>(stride) + static_cast<uint32_t>(ix);
    return t;
}

// The value is computed whether or not the tap is inside and then selected
// (a branch here would let the compiler sink the tap's loads into it, one
// round trip per tap).
__device__ inline float tap_value(const Tap& t, float a, float b, float c, float e) {
    const float v = (a * (1.0f - t.fx) + b * t.fx) * (1.0f - t.fy) + (c * (1.0f - t.fx) + e * t.fx) * t.fy;
    return t.ok ? v : -1.0f;
}

// reference sample_bilinear: -1 outside [0, W-1) x [0, H-1).  Both pixel
// pairs are loaded whether or not the point is inside (a point outside
// loads the frame's first pixels), so the loads carry no branch.
__device__ inline float sample_bilinear(const gbyte* image, int width, int height, int stride,
                                        float x, float y) {
    const Tap tp = make_tap(width, height, stride, x, y);
    const uint32_t span = frame_span(width, height, stride);
    const uint32_t top = pixel_pair(image, tp.off, span), bottom = pixel_pair(image, tp.off + stride, span);
    return tap_value(tp, static_cast<float>(top & 0xffu), static_cast<float>(top >> 8),
                     static_cast<float>(bottom & 0xffu), static_cast<float>(bottom >> 8));
}

__device__ inline float directional_rise(const gbyte* image, int width, int height, int stride,
                                         float x, float y, float dx, float dy) {
    const float ahead = sample_bilinear(image, width, height, stride, x + dx * 0.7f, y + dy * 0.7f);
    const float behind = sample_bilinear(image, width, height, stride, x - dx * 0.7f, y - dy * 0.7f);
    if (ahead < 0.0f || behind < 0.0f) return 0.0f;
    return ahead - behind;
}

__device__ inline float hyp(float x, float y) { return sqrtf(x * x + y * y); }

// Major-axis normal of the moments by the half-angle identities (the same
// closed form as the oracle; see the note on axis_normal there).
__device__ inline void axis_normal(float cxx, float cxy, float cyy, float* nx, float* ny) {
    const float a = cxx - cyy;
    const float b = 2.0f * cxy;
    const float r = sqrtf(a * a + b * b);
    float c, s;
    if (!(r > 0.0f)) {
        c = 1.0f;
        s = 0.0f;
    } else if (a >= 0.0f) {
        c = sqrtf((r + a) / (2.0f * r));
        s = b / (2.0f * r * c);
    } else {
        const float sa = sqrtf((r - a) / (2.0f * r));
        s = signbit(b) ? -sa : sa;
        c = b / (2.0f * r * s);
    }
    *nx = -s;
    *ny = c;
}

__device__ inline bool intersect_lines(const float* li, const float* lj, float* x, float* y) {
    const float det = li[2] * lj[3] - li[3] * lj[2];
    if (fabsf(det) < 0.2f) return false;
    const float di = li[2] * li[0] + li[3] * li[1];
    const float dj = lj[2] * lj[0] + lj[3] * lj[1];
    *x = (di * lj[3] - li[3] * dj) / det;
    *y = (li[2] * dj - di * lj[2]) / det;
    return true;
}

__device__ inline float lane_f(float v, int k) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

// reference apply_refined_lines (gpu_detector.cu:1021-1076) on lanes 0-3 of
// one wave (the whole wave calls it; lane i refines corner (i + 1) & 3 from
// lines i and i + 1, the reference's loop step i), reading the quad, lines
// and line flags from LDS and writing the refined quad back when it is
// accepted.  Every value is the one the reference's loop computes: the
// extremes are folded in the loop's order and the area terms summed in
// corner order, from lane reads.  Returns the largest corner move (0 when
// the refinement is rejected).
__device__ float apply_refined_lines(float (*quad)[2], const float (*lines)[4], const int* have, int lane) {
    const int i = lane & 3, j = (i + 1) & 3;
    const float qix = quad[i][0], qiy = quad[i][1], qjx = quad[j][0], qjy = quad[j][1];
    float li[4], lj[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        li[k] = lines[i][k];
        lj[k] = lines[j][k];
    }
    const bool hi = have[i] != 0, hj = have[j] != 0;
    if (__popcll(__ballot(lane < 4 && hi)) < 3) return 0.0f;
    const float side = hyp(qjx - qix, qjy - qiy);
    float extent = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) extent = fmaxf(extent, lane_f(side, k));
    const float max_shift = fminf(fmaxf(0.15f * extent, 4.0f), 16.0f);
    float rx = qjx, ry = qjy, mv = 0.0f;  // corner j, refined or kept
    float x, y;
    if (hi && hj && intersect_lines(li, lj, &x, &y)) {
        const float dx = x - qjx;
        const float dy = y - qjy;
        if (!(dx * dx + dy * dy > max_shift * max_shift)) {
            rx = x;
            ry = y;
            mv = sqrtf(dx * dx + dy * dy);
        }
    }
    // Corner c is refined by lane (c + 3) & 3.
    float cx[4], cy[4];
    float moved = 0.0f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        cx[c] = lane_f(rx, (c + 3) & 3);
        cy[c] = lane_f(ry, (c + 3) & 3);
        moved = fmaxf(moved, lane_f(mv, c));
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int d = (c + 1) & 3;
        if (hyp(cx[d] - cx[c], cy[d] - cy[c]) < 3.0f) return 0.0f;
    }
    float area2 = 0.0f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int d = (c + 1) & 3;
        area2 += cx[c] * cy[d] - cx[d] * cy[c];
    }
    if (area2 <= 0.0f) return 0.0f;
    if (lane < 4) {
        quad[j][0] = rx;
        quad[j][1] = ry;
    }
    return moved;
}

// Geometry of one edge fit (reference fit_edge_line_warp :884-920).
struct EdgeGeom {
    float ax, ay, ex, ey, nx, ny, step;
    int half, probes, samples;
    bool ok;
};

__device__ inline EdgeGeom edge_geometry(const float (*quad)[2], int edge, int dim, bool first_pass) {
    EdgeGeom g;
    const int b = (edge + 1) & 3;
    g.ax = quad[edge][0];
    g.ay = quad[edge][1];
    const float bx = quad[b][0], by = quad[b][1];
    g.ex = bx - g.ax;
    g.ey = by - g.ay;
    const float len = sqrtf(g.ex * g.ex + g.ey * g.ey);
    g.ok = !(len < 8.0f);
    g.nx = g.ey / len;
    g.ny = -g.ex / len;
    const float cell = len / static_cast<float>(dim);
    const float range = first_pass ? fminf(fmaxf(1.15f * cell, 1.5f), 2.5f * cell)
                                   : fminf(fmaxf(0.30f * cell, 0.8f), 2.0f);
    g.half = min(max(2, static_cast<int>(ceilf(range / 0.75f))), kFitProbeHalf);
    g.probes = 2 * g.half + 1;
    g.step = range / static_cast<float>(g.half);
    g.samples = min(max(static_cast<int>(len / 6.0f), 8), kMaxSamples);
    return g;
}

// Sampled bbox row s of `kept` (all rows when rows <= kMaxBoundaryRows;
// collect_row_extremes :676-715).
__device__ inline int fit_row(int s, int kept, int rows, int by0) {
    return by0 + (kept > 1 ? (s * (rows - 1)) / (kept - 1) : 0);
}

// One workgroup (4 waves) per candidate.
//  1. Row extremes of the component over (at most 1024 sampled) bbox rows,
//     one lookup per row in the row's component table (collect_row_extremes
//     :676-715 scans the row's runs from a binary search to bx0).
//  2. The extreme boundary point in each of 32 directions: 8 slices of the
//     boundary x 32 directions, first strictly larger projection wins, so the
//     result is the first maximum in boundary order as in :1132-1170.
//  3. Hull reduction to four corners by the smallest-triangle rule, one lane
//     per vertex (hull_to_quad :736-831).
//  4. Two passes of edge refinement: every (edge, sample, probe) rise is
//     evaluated in parallel into LDS, then one lane per (edge, sample) runs
//     the peak rules and the parabola, and the moments are summed in the
//     reference's 32-lane shuffle-down order (fit_edge_line_warp :881-1016).
template <int kThreads>
__device__ void fit_candidate(const KFrame* d, int ci, int t) {
    // kThreads / 32 slices of the boundary in the direction probe; kThreads
    // / 256 waves per edge in the refinement's rise loop.
    constexpr int kParts = kThreads / 32, kSub = kThreads / 256;
    const int wave = t / kWave, lane = t % kWave;
    float* out = d->quads + 8 * static_cast<size_t>(ci);
    float* hull_out = d->hull_quads + 8 * static_cast<size_t>(ci);
    if (t < 8) {
        out[t] = __int_as_float(0x7fc00000);
        hull_out[t] = __int_as_float(0x7fc00000);
    }

    __shared__ float pdot[kParts][kFitDirections];
    __shared__ int prow[kParts][kFitDirections];
    __shared__ uint32_t extreme[kFitDirections];
    __shared__ uint32_t hull_pts[kFitDirections];
    // One LDS buffer holds the boundary points (until the hull is built) and
    // then the refinement's probe magnitudes, rows padded to a multiple of
    // four floats for 16-byte LDS reads.
    constexpr int kProbeStride = (kMaxProbes + 3) & ~3;
    constexpr int kFitLdsWords = 2 * kMaxBoundaryRows > 4 * kMaxSamples * kProbeStride
                                     ? 2 * kMaxBoundaryRows : 4 * kMaxSamples * kProbeStride;
    __shared__ __align__(16) uint32_t work_lds[kFitLdsWords];
    uint32_t* boundary = work_lds;
    float (*mags)[kMaxSamples][kProbeStride] = reinterpret_cast<float (*)[kMaxSamples][kProbeStride]>(work_lds);
    __shared__ float fit_quad[4][2];
    __shared__ float edge_lines[4][4];
    __shared__ int edge_have[4];
    __shared__ int quad_ok;
    __shared__ float refine_moved;

    const uint32_t root = d->candidates[ci];
    const uint32_t row_cap = static_cast<uint32_t>(d->row_cap);
    // The root is the component's first run, so its row is the top row.
    const int by0 = static_cast<int>(root / row_cap);
    const int by1 = static_cast<int>(d->bbox[4 * static_cast<size_t>(root) + 3]);
    const int rows = by1 - by0 + 1;
    const int kept = min(rows, kMaxBoundaryRows);

    // Row extremes: the component's leftmost start and rightmost end in
    // each sampled row, looked up in the row's component table (written by
    // k_stats from every run of the row); a row whose table overflowed is
    // scanned whole.  Every row of the bbox holds a run of the (connected)
    // component.
    for (int s = t; s < kept; s += blockDim.x) {
        const int y = fit_row(s, kept, rows, by0);
        uint32_t lo = 0xffffffffu, hi = 0u;  // min x0, max x1 + 1 (0: none)
        const uint32_t entries = d->row_ext_full[y];
        // The row's packed entries, four per step (two 16-byte loads in
        // flight; slots past `entries` hold no entry of this row).  The first
        // step's loads go out with the entry count (a table has at least 64
        // slots) and its match is taken whatever the count says, so the
        // loads are not held back behind it; a full table ignores it.
        const uint64_t* tab = d->row_ext + (static_cast<size_t>(y) << d->row_ext_log2);
        const auto match4 = [&](uint32_t k, const ulonglong2& e01, const ulonglong2& e23) {
            const uint64_t ev[4] = {e01.x, e01.y, e23.x, e23.y};
            int hit = -1;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                if (hit < 0 && k + c < entries && static_cast<uint32_t>(ev[c] >> 32) == root) hit = c;
            }
            return hit < 0 ? 0ull : ev[hit] | (1ull << 63);
        };
        const ulonglong2 f01 = *reinterpret_cast<const ulonglong2*>(tab);
        const ulonglong2 f23 = *reinterpret_cast<const ulonglong2*>(tab + 2);
        // An empty asm statement that takes the three loaded values keeps
        // the compiler from sinking the table loads into the branch below,
        // behind the count's round trip.
        asm volatile("" : : "v"(entries), "v"(f01.x), "v"(f23.x));
        const uint64_t first = match4(0u, f01, f23);
        if (entries != kRowTableFull) {
            uint64_t found = first;
            for (uint32_t k = 4; found == 0ull && k < entries; k += 4)
                found = match4(k, *reinterpret_cast<const ulonglong2*>(tab + k),
                               *reinterpret_cast<const ulonglong2*>(tab + k + 2));
            if (found != 0ull) {
                lo = static_cast<uint32_t>(found >> 16) & 0xffffu;
                hi = (static_cast<uint32_t>(found) & 0xffffu) + 1u;
            }
        } else if (d->row_roi[4 * (y / kTile) + 3]) {
            // The row's items whose band root's component root (its label,
            // final after k_stats) is this root.
            const uint32_t band_base = static_cast<uint32_t>(y / kTile) * kTile * row_cap;
            const uint64_t* items = d->items + static_cast<size_t>(y) * row_cap;
            const uint32_t n = d->row_items[y];
            for (uint32_t k = 0; k < n; ++k) {
                const uint64_t it = items[k];
                const uint32_t broot =
                    band_base + static_cast<uint32_t>((it >> 40) & 7u) * row_cap + static_cast<uint32_t>(it >> 43);
                if (d->label[broot] != root) continue;
                lo = min(lo, static_cast<uint32_t>(it) & 0x1fffu);
                hi = max(hi, (static_cast<uint32_t>(it >> 13) & 0x1fffu) + 1u);
            }
        } else {
            // A band joined in global memory: every run of the row is a band
            // root, its label its component root.
            const uint32_t base = static_cast<uint32_t>(y) * row_cap;
            const uint32_t n = d->row_count[y];
            for (uint32_t k = 0; k < n; ++k) {
                if (d->label[base + k] != root) continue;
                const uint32_t run = d->runs[base + k];
                lo = min(lo, static_cast<uint32_t>(run_x0(run)));
                hi = max(hi, static_cast<uint32_t>(run_x1(run)) + 1u);
            }
        }
        // Boundary points x | y << 16, left then right; ~0 when absent.
        const uint32_t yy = static_cast<uint32_t>(y) << 16;
        boundary[2 * s] = lo == 0xffffffffu ? 0xffffffffu : (lo | yy);
        boundary[2 * s + 1] = hi == 0u ? 0xffffffffu : ((hi - 1u) | yy);
    }
    __syncthreads();

    {
        // kParts consecutive slices of the boundary rows x 32 directions.
        // Only one side of each row is scanned: projections are rounded
        // monotonically, so for ux <= 0 a row's left point projects at least
        // as far as its right point (and comes first in boundary order), and
        // for ux > 0 the right point projects at least as far.  The slices
        // are merged in order, so each direction keeps the first row holding
        // the largest projection; for ux > 0 that row's left point replaces
        // the right one when it projects as far, being first.  The result is
        // the first maximum in boundary order, as in :1132-1170.
        const int dir = t & 31, part = t >> 5;
        const int chunk = (((kept + kParts - 1) / kParts) + 1) & ~1;  // rows, even
        const int begin = part * chunk;
        const int end = min(kept, begin + chunk);
        const float ux = c_dir_cos[dir], uy = c_dir_sin[dir];
        const bool right = ux > 0.0f;
        float best = -FLT_MAX;
        int brow = -1;
        for (int r = begin; r < end; r += 2) {
            const uint4 q = *reinterpret_cast<const uint4*>(boundary + 2 * r);
            const uint32_t pv[2] = {right ? q.y : q.x, right ? q.w : q.z};
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const uint32_t p = pv[c];
                if (r + c >= end || p == 0xffffffffu) continue;
                const float dd = static_cast<float>(p & 0xffffu) * ux + static_cast<float>(p >> 16) * uy;
                if (dd > best) {
                    best = dd;
                    brow = r + c;
                }
            }
        }
        pdot[part][dir] = best;
        prow[part][dir] = brow;
    }
    __syncthreads();
    if (t < kFitDirections) {
        float best = -FLT_MAX;
        int brow = -1;
        // Unrolled so that the parts' LDS reads are issued together.
#pragma unroll
        for (int p = 0; p < kParts; ++p) {
            if (pdot[p][t] > best) {
                best = pdot[p][t];
                brow = prow[p][t];
            }
        }
        uint32_t point = 0xffffffffu;
        if (brow >= 0) {
            const float ux = c_dir_cos[t], uy = c_dir_sin[t];
            const uint32_t left = boundary[2 * brow];
            point = left;
            if (ux > 0.0f) {
                const float dl = static_cast<float>(left & 0xffffu) * ux + static_cast<float>(left >> 16) * uy;
                if (!(dl == best)) point = boundary[2 * brow + 1];
            }
        }
        extreme[t] = point;
    }
    __syncthreads();

    // Hull reduction on wave 0, one hull vertex (packed x | y << 16) per lane.
    if (wave == 0) {
        // Ordered hull without repeats (:1172-1180): keep a valid extreme
        // unless it equals the previous valid one, compact by ballot, then
        // drop a last vertex equal to the first (consecutive kept vertices
        // differ, so at most one).
        const uint32_t p = lane < kFitDirections ? extreme[lane] : 0xffffffffu;
        const bool valid = p != 0xffffffffu;
        const uint64_t below = __ballot(valid) & ((1ull << lane) - 1ull);
        const uint32_t pp = __shfl(p, below ? 63 - __clzll(static_cast<long long>(below)) : 0, kWave);
        const bool keep = valid && (below == 0 || pp != p);
        const uint64_t kmask = __ballot(keep);
        int hc = __popcll(kmask);
        if (keep) hull_pts[__popcll(kmask & ((1ull << lane) - 1ull))] = p;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        uint32_t hv = lane < hc ? hull_pts[lane] : 0xffffffffu;
        if (hc > 1 && wave_lane(hv, 0) == wave_lane(hv, hc - 1)) {
            --hc;
            if (lane == hc) hv = 0xffffffffu;
        }
        // Smallest-triangle removal (hull_to_quad :736-831): drop the vertex
        // whose triangle with its neighbours has the least area, the lowest
        // index on ties, until four remain.  hc <= 32, so lanes 0-31 hold
        // the polygon.
        int ok = hc >= 4 && hc <= kMaxHullVertices;
        int count = hc;
        while (ok && count > 4) {
            const uint32_t first = wave_lane(hv, 0);
            const uint32_t prev = dpp_from_lower(wave_lane(hv, count - 1), hv);
            // The DPP move runs on every lane: a lane masked off in EXEC
            // would read as out of range to its lower neighbour.
            const uint32_t upper = dpp_from_upper(first, hv);
            const uint32_t next = lane == count - 1 ? first : upper;
            const float vx = static_cast<float>(hv & 0xffffu), vy = static_cast<float>(hv >> 16);
            const float ax = vx - static_cast<float>(prev & 0xffffu), ay = vy - static_cast<float>(prev >> 16);
            const float bx = static_cast<float>(next & 0xffffu) - vx, by = static_cast<float>(next >> 16) - vy;
            const float area = fabsf(ax * by - ay * bx) * 0.5f;
            const bool live = lane < count && area < FLT_MAX;
            // Non-negative floats order as their bit patterns.
            const uint32_t key = live ? __float_as_uint(area) : 0xffffffffu;
            const uint32_t least = min_lanes_0_31(key);
            const uint64_t tie = __ballot(live && key == least);
            if (tie == 0) {
                ok = 0;
                break;
            }
            const int idx = __ffsll(static_cast<long long>(tie)) - 1;
            const uint32_t shifted = dpp_from_upper(hv, hv);
            if (lane >= idx) hv = shifted;
            --count;
        }
        float qx[4], qy[4];
        for (int i = 0; i < 4; ++i) {
            const uint32_t q = wave_lane(hv, i);
            qx[i] = static_cast<float>(q & 0xffffu);
            qy[i] = static_cast<float>(q >> 16);
        }
        if (lane == 0) {
            if (ok) {
                float area2 = 0.0f;
                for (int i = 0; i < 4; ++i) {
                    const int j = (i + 1) & 3;
                    area2 += qx[i] * qy[j] - qx[j] * qy[i];
                }
                if (fabsf(area2) < 24.0f) ok = 0;
                if (ok && area2 < 0.0f) {
                    for (int i = 0; i < 2; ++i) {
                        const float tx = qx[i + 1], tyy = qy[i + 1];
                        qx[i + 1] = qx[3 - i];
                        qy[i + 1] = qy[3 - i];
                        qx[3 - i] = tx;
                        qy[3 - i] = tyy;
                    }
                }
            }
            if (ok) {
                float min_side = FLT_MAX, max_side = 0.0f;
                for (int i = 0; i < 4 && ok; ++i) {
                    const int j = (i + 1) & 3;
                    const float ex = qx[j] - qx[i];
                    const float ey = qy[j] - qy[i];
                    const float len = sqrtf(ex * ex + ey * ey);
                    if (len < 4.0f) {
                        ok = 0;
                        break;
                    }
                    min_side = fminf(min_side, len);
                    max_side = fmaxf(max_side, len);
                    const int p = (i + 3) & 3;
                    const float ax = qx[i] - qx[p];
                    const float ay = qy[i] - qy[p];
                    const float bx = qx[j] - qx[i];
                    const float by = qy[j] - qy[i];
                    if (ax * by - ay * bx <= 0.0f) {
                        ok = 0;
                        break;
                    }
                    const float dot = ax * bx + ay * by;
                    const float n1 = sqrtf(ax * ax + ay * ay);
                    const float n2 = sqrtf(bx * bx + by * by);
                    if (n1 < 1e-3f || n2 < 1e-3f || fabsf(dot / (n1 * n2)) > 0.97f) {
                        ok = 0;
                        break;
                    }
                }
                if (ok && max_side > 12.0f * min_side) ok = 0;
            }
            if (ok) {
                const float cx = 0.25f * (qx[0] + qx[1] + qx[2] + qx[3]);
                const float cy = 0.25f * (qy[0] + qy[1] + qy[2] + qy[3]);
                for (int i = 0; i < 4; ++i) {
                    const float dx = qx[i] - cx;
                    const float dy = qy[i] - cy;
                    const float len = sqrtf(dx * dx + dy * dy);
                    if (len > 1e-3f) {
                        qx[i] += dx / len * 0.65f;
                        qy[i] += dy / len * 0.65f;
                    }
                }
                for (int i = 0; i < 4; ++i) {
                    fit_quad[i][0] = qx[i];
                    fit_quad[i][1] = qy[i];
                    hull_out[2 * i] = qx[i];
                    hull_out[2 * i + 1] = qy[i];
                }
            }
            quad_ok = ok;
        }
    }
    __syncthreads();
    if (!quad_ok) return;

    if (d->refine) {
        const gbyte* gray = frame_bytes(d->gray);
        const int W = d->width, H = d->height, stride = d->stride;
        const uint32_t span = frame_span(W, H, stride);
        for (int iteration = 0; iteration < kFitIterations; ++iteration) {
            const bool first_pass = iteration == 0;
            const int edge = wave & 3, sub = wave >> 2;
            const EdgeGeom g = edge_geometry(fit_quad, edge, d->dim, first_pass);
            if (g.ok) {
                // Every (sample, probe) rise of this edge, one per lane per
                // round (its four pixel-pair loads issued together); more per
                // round measured slower: the larger code and registers cost
                // more than the extra loads in flight gain.
                const int total = g.samples * g.probes;
                for (int it0 = sub * kWave + lane; it0 < total; it0 += kSub * kWave) {
                    Tap ta[1], tb[1];
                    int slot[1];
#pragma unroll
                    for (int u = 0; u < 1; ++u) {
                        const int it = it0 + u * kWave;
                        const bool live = it < total;
                        const int s = live ? it / g.probes : 0;
                        const int i = live ? it - s * g.probes : 0;
                        slot[u] = s * kProbeStride + i;
                        const float alpha = static_cast<float>(s + 1) / static_cast<float>(g.samples + 1);
                        const float px = g.ax + alpha * g.ex;
                        const float py = g.ay + alpha * g.ey;
                        const float tt = g.step * static_cast<float>(i - g.half);
                        const float x = px + g.nx * tt;
                        const float y = py + g.ny * tt;
                        ta[u] = make_tap(W, H, stride, x + g.nx * 0.7f, y + g.ny * 0.7f);
                        tb[u] = make_tap(W, H, stride, x - g.nx * 0.7f, y - g.ny * 0.7f);
                        ta[u].ok = ta[u].ok && live;
                        tb[u].ok = tb[u].ok && live;
                    }
                    // Each tap's 2x2 neighbourhood as two horizontal pixel
                    // pairs (one request each).
                    uint32_t pa[1][2], pb[1][2];
#pragma unroll
                    for (int u = 0; u < 1; ++u) {
                        // Loaded whatever the taps' bounds (a tap outside
                        // the frame reads its first pixels): no branch
                        // between the four loads.
                        const uint32_t p = ta[u].off, q = tb[u].off;
                        pa[u][0] = pixel_pair(gray, p, span);
                        pa[u][1] = pixel_pair(gray, p + stride, span);
                        pb[u][0] = pixel_pair(gray, q, span);
                        pb[u][1] = pixel_pair(gray, q + stride, span);
                    }
                    float va[1][4], vb[1][4];
#pragma unroll
                    for (int u = 0; u < 1; ++u) {
                        va[u][0] = static_cast<float>(pa[u][0] & 0xffu);
                        va[u][1] = static_cast<float>(pa[u][0] >> 8);
                        va[u][2] = static_cast<float>(pa[u][1] & 0xffu);
                        va[u][3] = static_cast<float>(pa[u][1] >> 8);
                        vb[u][0] = static_cast<float>(pb[u][0] & 0xffu);
                        vb[u][1] = static_cast<float>(pb[u][0] >> 8);
                        vb[u][2] = static_cast<float>(pb[u][1] & 0xffu);
                        vb[u][3] = static_cast<float>(pb[u][1] >> 8);
                    }
#pragma unroll
                    for (int u = 0; u < 1; ++u) {
                        if (it0 + u * kWave >= total) continue;
                        const float ahead = tap_value(ta[u], va[u][0], va[u][1], va[u][2], va[u][3]);
                        const float behind = tap_value(tb[u], vb[u][0], vb[u][1], vb[u][2], vb[u][3]);
                        (&mags[edge][0][0])[slot[u]] = (ahead < 0.0f || behind < 0.0f) ? 0.0f : ahead - behind;
                    }
                }
            }
            // mags[edge] is written by the edge's waves and read by its first.
            if constexpr (kSub == 1) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            } else {
                __syncthreads();
            }
            if (sub == 0) {
                float mx = 0.0f, my = 0.0f, mxx = 0.0f, mxy = 0.0f, myy = 0.0f;
                int used = 0;
                // Moments about the edge midpoint (see the oracle's note).
                const float rx = g.ax + 0.5f * g.ex, ry = g.ay + 0.5f * g.ey;
                // First strictly largest rise (:934-945): four lanes per
                // sample (lane = sample + 16 * part), each the first strict
                // maximum of one quarter of the probes (four per LDS read),
                // folded into the sample's lane in probe order.
                float best_mag = 0.0f;
                int best = -1;
                {
                    static_assert(kWave == 4 * kMaxSamples, "four quarters of the probes per sample");
                    const int ls = lane & (kMaxSamples - 1), part = lane / kMaxSamples;
                    if (g.ok && ls < g.samples) {
                        const float* mg = mags[edge][ls];
                        const int quarter = ((g.probes + 15) >> 4) << 2;
                        const int q0 = part * quarter, q1 = min(g.probes, q0 + quarter);
#pragma unroll 4
                        for (int q = q0; q < q1; q += 4) {
                            const float4 v = *reinterpret_cast<const float4*>(mg + q);
                            const float vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                            for (int c = 0; c < 4; ++c) {
                                if (q + c < q1 && vv[c] > best_mag) {
                                    best_mag = vv[c];
                                    best = q + c;
                                }
                            }
                        }
                    }
#pragma unroll
                    for (int k = 1; k < kWave / kMaxSamples; ++k) {
                        const float m = __shfl(best_mag, lane + kMaxSamples * k);
                        const int b = __shfl(best, lane + kMaxSamples * k);
                        if (m > best_mag) {
                            best_mag = m;
                            best = b;
                        }
                    }
                }
                if (g.ok && lane < g.samples) {
                    const float* mg = mags[edge][lane];
                    const float alpha = static_cast<float>(lane + 1) / static_cast<float>(g.samples + 1);
                    const float px = g.ax + alpha * g.ex;
                    const float py = g.ay + alpha * g.ey;
                    // The nearest-peak rule (:952-966): the first i (in
                    // increasing order) below best - 1 that is a local
                    // maximum of at least 0.7 best_mag with a valley under
                    // half its height before best.  Walking down from
                    // best - 2 with the running minimum of mg[i+1 .. best-1]
                    // finds the same i in one pass.
                    if (first_pass && best > 2) {
                        const float accept = 0.70f * best_mag;
                        float valley = FLT_MAX;
                        float right = mg[best - 1], here = mg[best - 2];
                        int nearest = -1;
#pragma unroll 4
                        for (int i = best - 2; i >= 1; --i) {
                            const float left = mg[i - 1];
                            valley = fminf(valley, right);
                            if (!(here < accept || here < left || here < right) && valley < 0.5f * here) nearest = i;
                            right = here;
                            here = left;
                        }
                        if (nearest >= 0) {
                            best = nearest;
                            best_mag = mg[nearest];
                        }
                    }
                    if (best >= 0 && best_mag >= 6.0f) {
                        float tt = g.step * static_cast<float>(best - g.half);
                        if (best > 0 && best < g.probes - 1) {
                            const float m0 = directional_rise(gray, W, H, stride, px + g.nx * (tt - g.step),
                                                              py + g.ny * (tt - g.step), g.nx, g.ny);
                            const float m2 = directional_rise(gray, W, H, stride, px + g.nx * (tt + g.step),
                                                              py + g.ny * (tt + g.step), g.nx, g.ny);
                            const float den = m0 - 2.0f * best_mag + m2;
                            if (den < -1e-3f) {
                                tt += fminf(fmaxf(0.5f * (m0 - m2) / den, -0.5f), 0.5f) * g.step;
                            }
                        }
                        const float x = px + g.nx * tt - rx;
                        const float y = py + g.ny * tt - ry;
                        mx = x;
                        my = y;
                        mxx = x * x;
                        mxy = x * y;
                        myy = y * y;
                        used = 1;
                    }
                }
                // The reference's shuffle-down tree over 32 lanes (lane i
                // adds lane i + off, off = 16 .. 1); lanes 32-63 hold zeros
                // and reduce in their own 32-lane segment.
                for (int off = 16; off > 0; off >>= 1) {
                    mx += __shfl_down(mx, off, 32);
                    my += __shfl_down(my, off, 32);
                    mxx += __shfl_down(mxx, off, 32);
                    mxy += __shfl_down(mxy, off, 32);
                    myy += __shfl_down(myy, off, 32);
                    used += __shfl_down(used, off, 32);
                }
                if (lane == 0) {
                    int have = 0;
                    if (g.ok && used >= 4) {
                        const float inv = 1.0f / static_cast<float>(used);
                        const float cx = mx * inv;
                        const float cy = my * inv;
                        const float cxx = mxx * inv - cx * cx;
                        const float cxy = mxy * inv - cx * cy;
                        const float cyy = myy * inv - cy * cy;
                        float lnx, lny;
                        axis_normal(cxx, cxy, cyy, &lnx, &lny);
                        edge_lines[edge][0] = rx + cx;
                        edge_lines[edge][1] = ry + cy;
                        edge_lines[edge][2] = lnx;
                        edge_lines[edge][3] = lny;
                        have = 1;
                    }
                    edge_have[edge] = have;
                }
            }
            __syncthreads();
            if (wave == 0) {
                const float moved = apply_refined_lines(fit_quad, edge_lines, edge_have, lane);
                if (lane == 0) refine_moved = moved;
            }
            __syncthreads();
            if (refine_moved < 0.15f) break;
        }
    }
    if (t == 0) {
        for (int i = 0; i < 4; ++i) {
            out[2 * i] = fit_quad[i][0];
            out[2 * i + 1] = fit_quad[i][1];
        }
        atomicAdd(&d->scal[S_COUNTS + 3], 1u);
    }
}

// Next work-list item of a persistent workgroup: the first is its block
// index, later ones come from the kernel's shared cursor, claimed only when
// the previous item is done (a claim made earlier would queue an item behind
// a long one while other workgroups run dry).
__device__ inline uint32_t next_item(uint32_t* cursor) {
    __shared__ uint32_t s_next;
    __syncthreads();
    if (threadIdx.x == 0) s_next = gridDim.x + atomicAdd(cursor, 1u);
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(s_next);  // uniform: a scalar loop counter
}

// Entry `item` of the batch candidate list: the head region (each frame's
// largest candidates) first, then the tail region.
__device__ inline uint32_t work_item(const Plan& plan, uint32_t item, uint32_t nhead) {
    const uint32_t* head = plan.work + kWorkItems;
    return kld(item < nhead ? head + item : head + kHeadSlots * plan.nframes + (item - nhead));
}

// Persistent workgroups over the batch's candidate list (k_select appends
// frame << 10 | candidate for every selected candidate of every frame):
// 256 threads, or for a launch over one or two frames (one round of
// candidates) 1024, four waves per edge in the refinement's rise loop.
template <int kThreads>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kThreads == 256 ? 8 : 4))) void k_fit(Plan plan) {
    const uint32_t nhead = kld(plan.work + kWorkHeadCount);
    const uint32_t total = nhead + kld(plan.work + kWorkCount);
    for (uint32_t item = blockIdx.x; item < total; item = next_item(&plan.work[kWorkFitCursor])) {
        // Uniform across the workgroup: scalar registers for the descriptor.
        const uint32_t code = __builtin_amdgcn_readfirstlane(work_item(plan, item, nhead));
        // An opaque copy of the thread index keeps the compiler from hoisting
        // every per-thread address out of the item loop (which would hold
        // them in registers across it and halve occupancy).
        int t = threadIdx.x;
        asm volatile("" : "+v"(t));
        fit_candidate<kThreads>((const KFrame*)plan.frames + (code >> 10), static_cast<int>(code & 1023u), t);
    }
}

// ---------------------------------------------------------------------------
// (e) decode
// ---------------------------------------------------------------------------

// reference homography_from_unit_quad (gpu_detector.cu:1249-1272).
__device__ inline bool homography_from_unit_quad(const float (*q)[2], float h[9]) {
    const float x0 = q[0][0], y0 = q[0][1];
    const float x1 = q[1][0], y1 = q[1][1];
    const float x2 = q[2][0], y2 = q[2][1];
    const float x3 = q[3][0], y3 = q[3][1];
    const float dx1 = x1 - x2, dx2 = x3 - x2, dx3 = x0 - x1 + x2 - x3;
    const float dy1 = y1 - y2, dy2 = y3 - y2, dy3 = y0 - y1 + y2 - y3;
    const float det = dx1 * dy2 - dx2 * dy1;
    if (fabsf(det) < 1e-9f) return false;
    const float g = (dx3 * dy2 - dx2 * dy3) / det;
    const float k = (dx1 * dy3 - dx3 * dy1) / det;
    h[0] = x1 - x0 + g * x1;
    h[1] = x3 - x0 + k * x3;
    h[2] = x0;
    h[3] = y1 - y0 + g * y1;
    h[4] = y3 - y0 + k * y3;
    h[5] = y0;
    h[6] = g;
    h[7] = k;
    h[8] = 1.0f;
    return true;
}

// Precision policy of the decode homography (no counterpart in the
// reference, which solves it in FP32 always): FP32 rounding in the sample
// positions grows with the determinant's cancellation (|dx1 dy2| +
// |dx2 dy1|) / |det| and with (1 + |g| + |k|) / min |g u + k v + 1| over the
// corners; when their product exceeds 16 (0.2 % of the quads of heavily
// tilted scenes) the homography and the sample positions are computed in
// FP64.  The measure is FP32, evaluated exactly as the oracle does.
// g and k are the FP32 homography's h[6] and h[7] (the same expressions).
__device__ inline bool decode_needs_fp64(const float (*q)[2], float g, float k) {
    const float dx1 = q[1][0] - q[2][0], dx2 = q[3][0] - q[2][0];
    const float dy1 = q[1][1] - q[2][1], dy2 = q[3][1] - q[2][1];
    const float det = dx1 * dy2 - dx2 * dy1;
    const float kappa = (fabsf(dx1 * dy2) + fabsf(dx2 * dy1)) / fabsf(det);
    const float den = fminf(1.0f, fminf(fabsf(1.0f + g), fminf(fabsf(1.0f + g + k), fabsf(1.0f + k))));
    const float amplification = kappa * (1.0f + fabsf(g) + fabsf(k)) / den;
    return !(amplification <= 16.0f);
}

// homography_from_unit_quad in FP64 from FP32 corners.
__device__ inline void homography_from_unit_quad_f64(const float (*q)[2], double h[9]) {
    const double x0 = q[0][0], y0 = q[0][1], x1 = q[1][0], y1 = q[1][1];
    const double x2 = q[2][0], y2 = q[2][1], x3 = q[3][0], y3 = q[3][1];
    const double dx1 = x1 - x2, dx2 = x3 - x2, dx3 = x0 - x1 + x2 - x3;
    const double dy1 = y1 - y2, dy2 = y3 - y2, dy3 = y0 - y1 + y2 - y3;
    const double det = dx1 * dy2 - dx2 * dy1;
    const double g = (dx3 * dy2 - dx2 * dy3) / det;
    const double k = (dx1 * dy3 - dx3 * dy1) / det;
    h[0] = x1 - x0 + g * x1;
    h[1] = x3 - x0 + k * x3;
    h[2] = x0;
    h[3] = y1 - y0 + g * y1;
    h[4] = y3 - y0 + k * y3;
    h[5] = y0;
    h[6] = g;
    h[7] = k;
    h[8] = 1.0;
}

// Result of one decode variant (reference decode_quad :1291-1430).
struct VariantResult {
    int id, ham;
    float margin;
    float corners[4][2];
};

// Variant v of candidate ci on one wave: the refined quad for v < 4, the
// hull quad otherwise, at rotation v & 3; one lane per grid cell, the code
// matched with one lane per group of codewords.  `diff` is the wave's own
// 64-entry LDS row (`reused`: the wave's previous variant read it).  Lane 0
// writes the result.
__device__ void decode_variant(const KFrame* d, int ci, int v, int lane, float* diff_row, bool reused,
                               const uint64_t* codes, int ncodes, VariantResult* out) {
    const int dim = d->dim;
    const int cells = dim * dim;
    const float* base = (v < 4 ? d->quads : d->hull_quads) + 8 * static_cast<size_t>(ci);
    const int rot = v & 3;
    float corners[4][2];
    for (int k = 0; k < 4; ++k) {
        const int src = (rot + k) & 3;
        corners[k][0] = kld(base + 2 * src);
        corners[k][1] = kld(base + 2 * src + 1);
    }
    float h[9];
    bool ok = homography_from_unit_quad(corners, h);
    const bool wide = ok && decode_needs_fp64(corners, h[6], h[7]);
    const int cx = lane % dim, cy = lane / dim;
    float raw = 0.0f;
    bool cell_ok = true;
    if (ok && lane < cells) {
        const float u = (static_cast<float>(cx) + 0.5f) / static_cast<float>(dim);
        const float vv = (static_cast<float>(cy) + 0.5f) / static_cast<float>(dim);
        float x = 0.0f, y = 0.0f;
        if (wide) {
            double h64[9];
            homography_from_unit_quad_f64(corners, h64);
            const double du = u, dv = vv;
            const double den = h64[6] * du + h64[7] * dv + 1.0;
            cell_ok = !(fabs(den) < 1e-9);
            if (cell_ok) {
                x = static_cast<float>((h64[0] * du + h64[1] * dv + h64[2]) / den);
                y = static_cast<float>((h64[3] * du + h64[4] * dv + h64[5]) / den);
            }
        } else {
            const float den = h[6] * u + h[7] * vv + 1.0f;
            cell_ok = !(fabsf(den) < 1e-9f);
            if (cell_ok) {
                x = (h[0] * u + h[1] * vv + h[2]) / den;
                y = (h[3] * u + h[4] * vv + h[5]) / den;
            }
        }
        if (cell_ok) {
            raw = sample_bilinear(frame_bytes(d->gray), d->width, d->height, d->stride, x - 0.0f, y - 0.0f);
            if (raw < 0.0f) cell_ok = false;
        }
    }
    const uint64_t bad_cells = __ballot(!cell_ok);
    ok = ok && bad_cells == 0ull;
    // With every cell inside the frame the samples are in [0, 255], where
    // float order is the order of the bit patterns (otherwise ok is false
    // and the extremes are not used).
    const bool sampled = lane < cells && raw >= 0.0f;
    const float mn = fminf(255.0f, __uint_as_float(wave_min(sampled ? __float_as_uint(raw) : __float_as_uint(255.0f))));
    const float mx = fmaxf(0.0f, __uint_as_float(wave_max(sampled ? __float_as_uint(raw) : 0u)));
    ok = ok && !(mx - mn < 10.0f);
    const float thresh = 0.5f * (mn + mx);
    const bool border = lane < cells && (cx == 0 || cy == 0 || cx == dim - 1 || cy == dim - 1);
    const uint64_t bad_border = __ballot(border && raw >= thresh);
    ok = ok && bad_border == 0ull;
    const float diff = raw - thresh;
    if (reused) {
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    }
    diff_row[lane] = diff;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    float sharp = diff;
    if (lane < cells && cx >= 1 && cx <= dim - 2 && cy >= 1 && cy <= dim - 2) {
        const float lap = 4.0f * diff - diff_row[lane - dim] - diff_row[lane + dim] - diff_row[lane - 1] - diff_row[lane + 1];
        sharp = diff + d->sharpening * lap;
    }
    // Code bits and class sums in the reference's row-major order: every
    // lane reads the interior cells' values by readlane and adds them in
    // that order (the same roundings as one serial loop).
    uint64_t code = 0;
    float margin = 0.0f;
    if (ok) {
        float white_sum = 0.0f, black_sum = 0.0f;
        int white_count = 0, black_count = 0;
        for (int yy = 1; yy <= dim - 2; ++yy) {
            for (int xx = 1; xx <= dim - 2; ++xx) {
                const float dv = __uint_as_float(wave_lane(__float_as_uint(sharp), yy * dim + xx));
                code = (code << 1) | (dv > 0.0f ? 1ull : 0ull);
                if (dv > 0.0f) {
                    white_sum += dv;
                    ++white_count;
                } else {
                    black_sum -= dv;
                    ++black_count;
                }
            }
        }
        const float white_avg = white_sum / static_cast<float>(max(white_count, 1));
        const float black_avg = black_sum / static_cast<float>(max(black_count, 1));
        margin = fminf(white_avg, black_avg);
    }
    // First codeword at the smallest distance (codebook in LDS), if within
    // max_hamming: (distance << 16 | index) orders by distance, then index.
    uint32_t key = 0xffffffffu;
    if (ok) {
        for (int i = lane; i < ncodes; i += kWave) {
            const uint32_t k = (static_cast<uint32_t>(__popcll(codes[i] ^ code)) << 16) | static_cast<uint32_t>(i);
            key = min(key, k);
        }
    }
    key = wave_min(key);
    const int best_d = key == 0xffffffffu ? 1000 : static_cast<int>(key >> 16);
    const int best_i = key == 0xffffffffu ? (1 << 30) : static_cast<int>(key & 0xffffu);
    if (lane == 0) {
        int id = -1;
        if (ok && best_d <= d->max_hamming) {
            if (!(margin < d->min_margin)) {
                const float score = static_cast<float>(4 - best_d) * 1000.0f + margin;
                if (score > -1.0f) id = best_i;
            }
        }
        out->id = id;
        out->ham = best_d;
        out->margin = margin;
        for (int k = 0; k < 4; ++k) {
            out->corners[k][0] = corners[k][0];
            out->corners[k][1] = corners[k][1];
        }
    }
}

// The reference's preference among the 8 variants (refined first, highest
// score, first on ties; k_decode_quads :1438-1537), written to the
// candidate's decode record.
__device__ void decode_choose(const KFrame* d, int ci, bool have_quad, const VariantResult* r) {
    bd_detection* res = d->dec + ci;
    int best = -1;
    float best_score = -1.0f;
    for (int variant = 0; variant < 2 && best < 0 && have_quad; ++variant) {
        for (int rotation = 0; rotation < 4; ++rotation) {
            const int v = variant * 4 + rotation;
            if (r[v].id < 0) continue;
            const float score = static_cast<float>(4 - r[v].ham) * 1000.0f + r[v].margin;
            if (score > best_score) {
                best_score = score;
                best = v;
            }
        }
    }
    if (best < 0) {
        res->id = -1;
    } else {
        res->id = r[best].id;
        res->hamming = r[best].ham;
        res->decision_margin = r[best].margin;
        for (int k = 0; k < 4; ++k) {
            res->corners[2 * k] = r[best].corners[k][0];
            res->corners[2 * k + 1] = r[best].corners[k][1];
        }
        res->reserved = 0;
        atomicAdd(&d->scal[S_DECODED], 1u);
    }
}

// Persistent decode workgroups of kWaves waves over the work list (the
// decode_waves launch parameter): 4 waves decode one candidate two variants
// per wave, 8 waves one candidate one variant per wave, 16 waves two
// candidates at once.  Decoding costs about the same per candidate, so
// workgroups take items by a fixed stride (no cursor).  The codebook (shared
// by every frame of the batch) is copied to LDS once per workgroup.
template <int kWaves>
__global__ __launch_bounds__(kWaves * 64) __attribute__((amdgpu_waves_per_eu(8))) void k_decode(Plan plan) {
    constexpr int kPerBlock = kWaves >= 16 ? kWaves / 8 : 1;   // candidates per workgroup
    constexpr int kWavesPerCand = kWaves / kPerBlock;           // 4 or 8
    constexpr int kVariantsPerWave = 8 / kWavesPerCand;         // 2 or 1
    const uint32_t nhead = kld(plan.work + kWorkHeadCount);
    const uint32_t total = nhead + kld(plan.work + kWorkCount);
    if (blockIdx.x * kPerBlock >= total) return;
    __shared__ uint64_t s_codes[kMaxCodes];
    __shared__ float s_diff[kWaves][kWave];
    __shared__ VariantResult s_res[kPerBlock][8];
    const KFrame* d0 = (const KFrame*)plan.frames;
    const int ncodes = min(d0->code_count, kMaxCodes);
    for (int i = threadIdx.x; i < ncodes; i += blockDim.x) s_codes[i] = d0->codes[i];
    __syncthreads();
    // The wave index in a scalar register, so that the item, its frame and
    // the descriptor fields are scalar too.
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
    const int lane = static_cast<int>(threadIdx.x) % kWave;
    const int slot = wave / kWavesPerCand, wslot = wave % kWavesPerCand;
    // Item, frame and descriptor stay in scalar registers: the step is read
    // once into one, and a slot past the list reads the last item (ignored)
    // rather than selecting a value in a divergent branch.
    const uint32_t step = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(gridDim.x)) * kPerBlock;
    for (uint32_t first = blockIdx.x * kPerBlock; first < total; first += step) {
        const uint32_t item = first + static_cast<uint32_t>(slot);
        const bool have = item < total;
        const uint32_t code = work_item(plan, have ? item : total - 1u, nhead);
        const KFrame* d = (const KFrame*)plan.frames + (code >> 10);
        const int ci = static_cast<int>(code & 1023u);
        // A candidate without a quad decodes to nothing.
        const float q0 = kld(d->quads + 8 * static_cast<size_t>(ci));
        const bool have_quad = have && q0 == q0;
        if (have_quad) {
#pragma unroll
            for (int k = 0; k < kVariantsPerWave; ++k) {
                const int v = wslot + k * kWavesPerCand;
                decode_variant(d, ci, v, lane, s_diff[wave], k > 0, s_codes, ncodes, &s_res[slot][v]);
            }
        }
        __syncthreads();
        if (have && wslot == 0 && lane == 0) decode_choose(d, ci, have_quad, s_res[slot]);
        __syncthreads();  // s_res of this item is no longer read
    }
}

// reference fill_homography (gpu_detector.cu:1739-1774), FP64 on the device.
__device__ inline void fill_homography(const float* corners, float* H) {
    const double x0 = corners[0], y0 = corners[1];
    const double x1 = corners[2], y1 = corners[3];
    const double x2 = corners[4], y2 = corners[5];
    const double x3 = corners[6], y3 = corners[7];
    const double dx1 = x1 - x2, dx2 = x3 - x2;
    const double dx3 = x0 - x1 + x2 - x3;
    const double dy1 = y1 - y2, dy2 = y3 - y2;
    const double dy3 = y0 - y1 + y2 - y3;
    double g = 0.0, hh = 0.0;
    const double det = dx1 * dy2 - dx2 * dy1;
    if (fabs(det) > 1e-12) {
        g = (dx3 * dy2 - dx2 * dy3) / det;
        hh = (dx1 * dy3 - dx3 * dy1) / det;
    }
    const double a = x1 - x0 + g * x1;
    const double b = x3 - x0 + hh * x3;
    const double dd = y1 - y0 + g * y1;
    const double e = y3 - y0 + hh * y3;
    const double w = g + hh + 2.0;
    const double inv = (fabs(w) > 1e-12) ? 1.0 / w : 0.0;
    H[0] = static_cast<float>(a * inv);
    H[1] = static_cast<float>(b * inv);
    H[2] = static_cast<float>((a + b + 2.0 * x0) * inv);
    H[3] = static_cast<float>(dd * inv);
    H[4] = static_cast<float>(e * inv);
    H[5] = static_cast<float>((dd + e + 2.0 * y0) * inv);
    H[6] = static_cast<float>(g * inv);
    H[7] = static_cast<float>(hh * inv);
    H[8] = static_cast<float>(w * inv);
}

__device__ inline void centre_side(const float* c, float* side, float* cx, float* cy) {
    float s = 0.0f, x = 0.0f, y = 0.0f;
    for (int k = 0; k < 4; ++k) {
        const int j = (k + 1) & 3;
        s += hyp(c[j * 2] - c[k * 2], c[j * 2 + 1] - c[k * 2 + 1]);
        x += c[k * 2];
        y += c[k * 2 + 1];
    }
    *side = s * 0.25f;
    *cx = x * 0.25f;
    *cy = y * 0.25f;
}

// One workgroup per frame: decoded candidates in candidate order (the first
// max_detections), de-duplicated as the reference's k_finalize (:1542-1601)
// does it serially, then placed by a stable rank on id, each with its FP64
// homography; the records and the frame's counts go to pinned host memory.
__global__ __launch_bounds__(256) void k_finalize(Plan plan) {
    const KFrame* d = (const KFrame*)plan.frames + blockIdx.x;
    if (!frame_head(d).active) return;
    const int t = threadIdx.x;
    // k_tile_stats' accumulators (histograms, active tile rows and each
    // tile row's active extent) start cleared for the frame's next pass;
    // k_binarize_runs has read them.
    for (int i = t; i < 320; i += blockDim.x) d->hist[i] = 0;
    for (int i = t; i < d->tile_h; i += blockDim.x) {
        d->trow_lo[i] = d->tile_w;
        d->trow_hi[i] = -1;
    }
    if (t == 0) {
        d->scal[S_TOP] = 0xffffffffu;
        d->scal[S_BOTTOM1] = 0;
    }
    const int maxd = d->max_detections;
    const int ncand = static_cast<int>(kld(d->scal + S_NCAND));
    // The decoded candidates' records, compacted in candidate order into LDS
    // as they are loaded (one global round trip for ids and records).
    __shared__ bd_detection s_rec[256];
    __shared__ int k_pos[256];      // record position of each kept entry
    __shared__ float k_cx[256], k_cy[256], k_margin[256];
    __shared__ int k_id[256];
    __shared__ int s_kept, s_dup, s_any_shared;
    __shared__ uint32_t s_base;
    if (t == 0) {
        s_kept = 0;
        s_any_shared = 0;
        s_base = 0;
    }
    __syncthreads();
    // Compaction in candidate order, chunks of 256.
    for (int c0 = 0; c0 < ncand; c0 += blockDim.x) {
        const int ci = c0 + t;
        bd_detection rec;
        rec.id = -1;
        if (ci < ncand) rec = d->dec[ci];
        const bool valid = rec.id >= 0;
        // block-wide exclusive scan of `valid`
        __shared__ uint32_t wsum[4];
        const int lane = t % kWave, wave = t / kWave;
        const uint64_t bal = __ballot(valid);
        const uint32_t before = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wave] = __popcll(bal);
        __syncthreads();
        uint32_t off = s_base;
        for (int w = 0; w < wave; ++w) off += wsum[w];
        if (valid) {
            const uint32_t pos = off + before;
            if (pos < static_cast<uint32_t>(maxd)) s_rec[pos] = rec;
        }
        __syncthreads();
        if (t == 0) s_base += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
    const int n = min(static_cast<int>(s_base), maxd);
    // Any id shared by two entries?  If not, de-duplication keeps every entry.
    for (int i = t; i < n; i += blockDim.x) {
        const int id = s_rec[i].id;
        for (int j = 0; j < n; ++j) {
            if (j != i && s_rec[j].id == id) {
                s_any_shared = 1;
                break;
            }
        }
    }
    __syncthreads();
    if (!s_any_shared) {
        for (int i = t; i < n; i += blockDim.x) {
            k_pos[i] = i;
            k_id[i] = s_rec[i].id;
        }
        if (t == 0) s_kept = n;
    } else {
        for (int i = 0; i < n; ++i) {
            const bd_detection& e = s_rec[i];
            float side, cx, cy;
            centre_side(e.corners, &side, &cx, &cy);
            if (t == 0) s_dup = 0x7fffffff;
            __syncthreads();
            const int kept = s_kept;
            const float limit = 0.35f * side;
            for (int k = t; k < kept; k += blockDim.x) {
                if (k_id[k] == e.id && fabsf(k_cx[k] - cx) < limit && fabsf(k_cy[k] - cy) < limit) {
                    atomicMin(&s_dup, k);
                }
            }
            __syncthreads();
            if (t == 0) {
                const int dup = s_dup;
                if (dup != 0x7fffffff) {
                    if (e.decision_margin > k_margin[dup]) {
                        float sd, ox, oy;
                        centre_side(e.corners, &sd, &ox, &oy);
                        k_pos[dup] = i;
                        k_id[dup] = e.id;
                        k_margin[dup] = e.decision_margin;
                        k_cx[dup] = ox;
                        k_cy[dup] = oy;
                    }
                } else if (kept < maxd) {
                    k_pos[kept] = i;
                    k_id[kept] = e.id;
                    k_margin[kept] = e.decision_margin;
                    k_cx[kept] = cx;
                    k_cy[kept] = cy;
                    s_kept = kept + 1;
                }
            }
            __syncthreads();
        }
    }
    __syncthreads();
    const int kept = s_kept;
    // Stable placement by id (the reference's insertion sort); k_id holds
    // the kept entries' ids.
    for (int i = t; i < kept; i += blockDim.x) {
        const int id = k_id[i];
        int rank = 0;
        for (int j = 0; j < kept; ++j) {
            const int oj = k_id[j];
            rank += (oj < id) || (oj == id && j < i);
        }
        bd_detection o = s_rec[k_pos[i]];
        o.reserved = 0;
        fill_homography(o.corners, o.homography);
        d->out[rank] = o;
    }
    if (t == 0) d->scal[S_COUNTS + 1] = static_cast<uint32_t>(kept);
    // The counts for the host (every other count is final: its kernels ran).
    if (!d->host_counts) return;
    if (t < 5) d->host_counts[t] = t == 1 ? static_cast<uint32_t>(kept) : d->scal[S_COUNTS + t];
    if (t == 5) *reinterpret_cast<uint64_t*>(d->host_counts + kHostEnd) = __builtin_amdgcn_s_memrealtime();
    // Then the pass number, after every wave's records and counts: each wave
    // waits for its stores and writes them back (system-scope release), the
    // barrier orders those before the one flag store.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    __syncthreads();
    if (t == 0) {
        const uint32_t seq = d->scal[S_SEQ] + 1u;
        d->scal[S_SEQ] = seq;
        __hip_atomic_store(&d->host_counts[kHostSeq], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// ---------------------------------------------------------------------------
// device-resident serving
// ---------------------------------------------------------------------------

// Admission of one lane's launch (one wave): waits, bounded by a wall-clock
// budget, a poll count and the host's stop flag, for the class's next
// request to be published; claims the run of published requests from the
// class head, up to the lane's positions, with one compare-and-swap on the
// head (lanes of a class admit concurrently); and points the claimed
// positions' descriptors at their frames.  Unclaimed positions are inactive
// for this launch.  Every path through the loop ends: a claim, the stop
// flag, the budget or the poll cap.
__global__ __launch_bounds__(64) void k_admit(Plan plan, ResidentArgs a) {
    const int lane = threadIdx.x;
    const int cap = plan.nframes;
    uint32_t first = 0, count = 0;
    if (lane == 0) {
        const uint64_t start = wall_clock64();
        for (uint32_t polls = 0;; ++polls) {
            const uint32_t h = __hip_atomic_load(a.head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint32_t n = 0;
            while (static_cast<int>(n) < cap &&
                   __hip_atomic_load(&a.slots[(h + n) % static_cast<uint32_t>(a.depth)].seq, __ATOMIC_ACQUIRE,
                                     __HIP_MEMORY_SCOPE_SYSTEM) == h + n + 1u)
                ++n;
            if (n > 0) {
                uint32_t expected = h;
                if (__hip_atomic_compare_exchange_strong(a.head, &expected, h + n, __ATOMIC_ACQ_REL,
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    first = h;
                    count = n;
                    break;
                }
                // Another lane took them: look again from the new head, after
                // the same checks as an empty poll.
            }
            if (__hip_atomic_load(a.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u ||
                wall_clock64() - start > a.wait_ticks || polls >= a.max_polls)
                break;
            if (n == 0) __builtin_amdgcn_s_sleep(8);
        }
    }
    first = static_cast<uint32_t>(__shfl(static_cast<int>(first), 0));
    count = static_cast<uint32_t>(__shfl(static_cast<int>(count), 0));
    for (int k = lane; k < cap; k += 64) {
        FrameDesc* d = const_cast<FrameDesc*>(plan.frames) + k;
        const bool on = static_cast<uint32_t>(k) < count;
        if (on) {
            const uint32_t t = first + static_cast<uint32_t>(k);
            const RequestSlot* r = a.slots + t % static_cast<uint32_t>(a.depth);
            d->gray = r->gray;
            d->stride = r->stride;
            d->out = a.lane_out + static_cast<size_t>(k) * a.max_det;
            a.lane_tickets[k] = t;
        }
        d->active = on ? 1 : 0;
    }
    if (lane == 0) a.lane_tickets[cap] = count;
}

// Publication of one lane's launch (one wave): each admitted request's
// records (from the lane's device output) and count into its ticket's
// slots in pinned host memory, a system-scope release fence and a barrier,
// then each request's completion flag.
__global__ __launch_bounds__(64) void k_publish(Plan plan, ResidentArgs a) {
    const uint32_t count = a.lane_tickets[plan.nframes];
    const uint32_t words = sizeof(bd_detection) / sizeof(uint32_t);
    static_assert(sizeof(bd_detection) % sizeof(uint32_t) == 0, "records copy as words");
    for (uint32_t k = 0; k < count; ++k) {
        const KFrame* d = (const KFrame*)plan.frames + k;
        const uint32_t t = a.lane_tickets[k];
        const uint32_t slot = t % static_cast<uint32_t>(a.depth);
        const int n = min(static_cast<int>(d->scal[S_COUNTS + 1]), d->max_detections);
        const uint32_t* src = reinterpret_cast<const uint32_t*>(d->out);
        uint32_t* dst = reinterpret_cast<uint32_t*>(a.out + static_cast<size_t>(slot) * a.max_det);
        for (uint32_t w = threadIdx.x; w < static_cast<uint32_t>(n) * words; w += 64) dst[w] = src[w];
        if (threadIdx.x == 0) {
            a.done[slot].count = n;
            a.done[slot].batch = static_cast<int32_t>(count);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < count; k += 64) {
        const uint32_t t = a.lane_tickets[k];
        __hip_atomic_store(&a.done[t % static_cast<uint32_t>(a.depth)].seq, t + 1u, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Frame bindings of a device-path call (one wave): pointer, stride and
// active flag of up to kBindFrames positions from the kernel arguments into
// their descriptors.
__global__ __launch_bounds__(64) void k_bind(FrameDesc* frames, BindArgs a) {
    const int k = threadIdx.x;
    if (k >= a.count) return;
    FrameDesc* d = frames + a.first + k;
    d->gray = a.gray[k];
    d->stride = a.stride[k];
    d->active = a.active[k];
}

}  // namespace

hipError_t launch_bind(FrameDesc* frames, const BindArgs& args, hipStream_t stream) {
    hipLaunchKernelGGL(k_bind, dim3(1), dim3(64), 0, stream, frames, args);
    return hipGetLastError();
}

hipError_t launch_admit(const Plan& plan, const ResidentArgs& args, hipStream_t stream) {
    hipLaunchKernelGGL(k_admit, dim3(1), dim3(64), 0, stream, plan, args);
    return hipGetLastError();
}

hipError_t launch_publish(const Plan& plan, const ResidentArgs& args, hipStream_t stream) {
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, stream, plan, args);
    return hipGetLastError();
}

int32_t binarize_wide_blocks() {
    static const int32_t blocks = [] {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            return 0;
        return 2 * cus;
    }();
    return blocks;
}

// 1024 threads per workgroup when the launch fits in one round of them, or
// when the class's band capacity leaves room for at most two workgroups per
// CU (wide frames): two 1024-thread workgroups keep a CU's 32 wave slots
// busy where two 256-thread ones would hold eight.
template <int kChunkWords, int kBands>
void launch_binarize(const Plan& plan, int cls, int blocks, size_t lds, hipStream_t stream) {
    constexpr size_t kLdsPerCu = 160 * 1024;
    if (blocks <= plan.wide_blocks || kLdsPerCu / (lds + 2048) <= 2)
        hipLaunchKernelGGL((k_binarize_runs<kChunkWords, kBands, 1024>), dim3(blocks), dim3(1024), lds, stream, plan,
                           cls);
    else
        hipLaunchKernelGGL((k_binarize_runs<kChunkWords, kBands, 256>), dim3(blocks), dim3(256), lds, stream, plan,
                           cls);
}

Tune default_tune() {
    // The persistent grids default to one round of resident workgroups: as
    // many per CU as one CU holds of the compiled kernel (registers and LDS
    // decide).  Computed once per process.
    static const Tune tuned = [] {
        Tune t;
        int dev = 0, cus = 0, fit = 0, dec = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            return t;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, k_fit<256>, 256, 0) == hipSuccess && fit > 0)
            t.fit_grid = std::min(std::max(cus * fit, 64), 16384);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&dec, k_decode<8>, 512, 0) == hipSuccess && dec > 0)
            t.decode_grid = std::min(std::max(cus * dec, 64), 16384);
        return t;
    }();
    return tuned;
}

hipError_t launch_stage(const Plan& plan, int stage, hipStream_t stream) {
    const int nf = plan.nframes;
    switch (stage) {
        case 0: hipLaunchKernelGGL(k_tile_stats, dim3(plan.blocks[SH_TSTAT]), dim3(kTileStatBlocks), 0, stream, plan); break;
        case 1:
            // One launch per width class present, with the class's band capacity.
            for (int c = 0; c < kWidthClasses; ++c) {
                const bool pairs = plan.tune.band_rows == 16;
                const int blocks = plan.blocks[(pairs ? SH_CLASS_PAIR : SH_CLASS) + c];
                if (blocks == 0) continue;
                const size_t lds = static_cast<size_t>(plan.class_band_runs[c]) * 8;
                const bool wide = plan.tune.chunk_px == 128;
                if (wide && pairs)
                    launch_binarize<2, 2>(plan, c, blocks, lds, stream);
                else if (wide)
                    launch_binarize<2, 1>(plan, c, blocks, lds, stream);
                else if (pairs)
                    launch_binarize<1, 2>(plan, c, blocks, lds, stream);
                else
                    launch_binarize<1, 1>(plan, c, blocks, lds, stream);
            }
            break;
        case 2:
            // One launch per width class present, staging its rows in LDS.
            for (int c = 0; c < kWidthClasses; ++c) {
                const int blocks = plan.blocks[SH_CLASS + c];
                if (blocks == 0) continue;
                if (c == 0 && nf > 2) {
                    const size_t lds = static_cast<size_t>(std::min(plan.class_merge_stage[c], kMergeStageNarrow)) * 16;
                    hipLaunchKernelGGL((k_merge<128, kPairSlots / 2>), dim3(blocks), dim3(128), lds, stream, plan, c);
                } else {
                    const size_t lds = static_cast<size_t>(std::min(plan.class_merge_stage[c], kMergeStageWide)) * 16;
                    hipLaunchKernelGGL((k_merge<256, kPairSlots>), dim3(blocks), dim3(256), lds, stream, plan, c);
                }
            }
            break;
        case 3: {
            // One launch per width class present, with the class's row tables.
            static_assert(kWidthClasses == 3, "one k_stats instance per width class");
            if (plan.blocks[SH_CLASS + 0])
            {
                // 128 threads for launches over more than two frames: more
                // resident workgroups of this latency-bound kernel per CU.
                if (nf > 2)
                    hipLaunchKernelGGL((k_stats<kRowExtLog2Min + 0, 128>), dim3(plan.blocks[SH_CLASS + 0]), dim3(128), 0, stream, plan);
                else
                    hipLaunchKernelGGL((k_stats<kRowExtLog2Min + 0, 256>), dim3(plan.blocks[SH_CLASS + 0]), dim3(256), 0, stream, plan);
            }
            if (plan.blocks[SH_CLASS + 1])
                hipLaunchKernelGGL((k_stats<kRowExtLog2Min + 1, 256>), dim3(plan.blocks[SH_CLASS + 1]), dim3(256), 0, stream, plan);
            if (plan.blocks[SH_CLASS + 2])
                hipLaunchKernelGGL((k_stats<kRowExtLog2Min + 2, 512>), dim3(plan.blocks[SH_CLASS + 2]), dim3(512), 0, stream, plan);
            break;
        }
        case 4: hipLaunchKernelGGL(k_collect, dim3(plan.blocks[SH_TILE_ROW]), dim3(64), 0, stream, plan); break;
        case 5: hipLaunchKernelGGL(k_select, dim3(nf), dim3(256), 0, stream, plan); break;
        case 6:
            if (plan.nframes <= 2 && plan.wide_blocks >= 2)
                hipLaunchKernelGGL(k_fit<1024>, dim3(min(plan.work_cap, plan.wide_blocks / 2)), dim3(1024), 0, stream,
                                   plan);
            else
                hipLaunchKernelGGL(k_fit<256>, dim3(min(plan.work_cap, plan.tune.fit_grid)), dim3(256), 0, stream, plan);
            break;
        case 7: {
            const int w = plan.tune.decode_waves;
            const int per = w >= 16 ? w / 8 : 1;
            const dim3 grid(min((plan.work_cap + per - 1) / per, plan.tune.decode_grid));
            if (w == 4) hipLaunchKernelGGL(k_decode<4>, grid, dim3(256), 0, stream, plan);
            else if (w == 16) hipLaunchKernelGGL(k_decode<16>, grid, dim3(1024), 0, stream, plan);
            else hipLaunchKernelGGL(k_decode<8>, grid, dim3(512), 0, stream, plan);
            break;
        }
        case 8: hipLaunchKernelGGL(k_finalize, dim3(nf), dim3(256), 0, stream, plan); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_pipeline(const Plan& plan, hipStream_t stream) {
    for (int stage = 0; stage < kStageCount; ++stage) {
        const hipError_t e = launch_stage(plan, stage, stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace bd
