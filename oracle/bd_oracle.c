/*
 * Scalar CPU oracle of the bdcudatag pipeline -- TEST INFRASTRUCTURE ONLY.
 *
 * A restatement of the reference kernels in detector/src/gpu_detector.cu,
 * executed serially.  Every function names the reference lines it follows.
 * Build with -ffp-contract=off so that every a*b+c is a rounded product and a
 * rounded sum, as written.
 *
 * Where the reference is order-dependent on atomics, the oracle fixes one
 * order that the reference allows:
 *   - candidates inside an area bucket are ordered by (area desc, root slot
 *     asc); the root slot is the component's first fragment in raster order;
 *   - decoded detections enter de-duplication in candidate order.
 * Where the reference calls a CUDA transcendental, the oracle uses a closed
 * form that both this file and the HIP kernels evaluate with correctly
 * rounded +, -, *, / and sqrt only:
 *   - hull probe directions: a 32-entry table of cos/sin of the reference's
 *     FP32 angle 6.28318530718f * d / 32, correctly rounded to FP32;
 *   - edge-line direction: the major-axis normal of the 2x2 moment matrix by
 *     the half-angle identities instead of atan2f/sinf/cosf;
 *   - hypotf(x, y) is sqrtf(x * x + y * y).
 * Built with -DOR_LITERAL (FP32 only), these three places and the edge-fit
 * moments follow the reference literally instead: libm cosf/sinf for the
 * hull directions, raw moments about the image origin with
 * 0.5f * atan2f(2 cxy, cxx - cyy) and -sinf/cosf for the edge line
 * (gpu_detector.cu:981-1015), and hypotf (:1035, :1061, :1557).  That build
 * bounds how far the closed forms above move the results; it is not the
 * parity target of the HIP kernels.
 */
#include "bd_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../bd-cudatag_amd/csrc/tag_codes.inc"

#ifdef OR_FP64
typedef double real;
#define RL(x) (x)
#define R_SQRT sqrt
#define R_FABS fabs
#define R_CEIL ceil
#define R_FMIN fmin
#define R_FMAX fmax
#define R_MAX DBL_MAX
#define R_BITS 64
#else
typedef float real;
#define RL(x) (x##f)
#define R_SQRT sqrtf
#define R_FABS fabsf
#define R_CEIL ceilf
#define R_FMIN fminf
#define R_FMAX fmaxf
#define R_MAX FLT_MAX
#define R_BITS 32
#endif

/* Constants of the reference, gpu_detector.cu:38-50. */
enum {
    kTile = 8,
    kMinTileContrast = 12,
    kFitDirections = 32,
    kMaxBoundaryRows = 1024,
    kMaxHullVertices = 64,
    kFitIterations = 2,
    kFitProbeHalf = 24,
    kFitWarps = 4,
};

static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }
static uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
static uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }
static int popc32(uint32_t v) { return __builtin_popcount(v); }
static int popc64(uint64_t v) { return __builtin_popcountll(v); }

#if defined(OR_LITERAL) && defined(OR_FP64)
#error "OR_LITERAL restates the FP32 reference"
#endif

#ifdef OR_LITERAL
static real or_hypot(real x, real y) { return hypotf(x, y); }
#else
static real or_hypot(real x, real y) { return R_SQRT(x * x + y * y); }
#endif

#if !defined(OR_FP64) && !defined(OR_LITERAL)
/* cos/sin of fl32(fl32(6.28318530718f * d) / 32), correctly rounded to FP32
 * (the reference computes cosf/sinf of that angle at gpu_detector.cu:1139-1142). */
static const float kDirCos[32] = {
    0x1p+0f,          0x1.f6297cp-1f,   0x1.d906bcp-1f,   0x1.a9b662p-1f,
    0x1.6a09e6p-1f,   0x1.1c73b2p-1f,   0x1.87de2ap-2f,   0x1.8f8b78p-3f,
    -0x1.777a5cp-25f, -0x1.8f8b84p-3f,  -0x1.87de30p-2f,  -0x1.1c73b8p-1f,
    -0x1.6a09e6p-1f,  -0x1.a9b664p-1f,  -0x1.d906c0p-1f,  -0x1.f6297ep-1f,
    -0x1p+0f,         -0x1.f6297cp-1f,  -0x1.d906bcp-1f,  -0x1.a9b660p-1f,
    -0x1.6a09e2p-1f,  -0x1.1c73acp-1f,  -0x1.87de16p-2f,  -0x1.8f8b8cp-3f,
    0x1.99bc5cp-27f,  0x1.8f8b90p-3f,   0x1.87de36p-2f,   0x1.1c73bap-1f,
    0x1.6a09eep-1f,   0x1.a9b662p-1f,   0x1.d906bep-1f,   0x1.f6297ep-1f,
};
static const float kDirSin[32] = {
    0x0p+0f,          0x1.8f8b84p-3f,   0x1.87de2cp-2f,   0x1.1c73b4p-1f,
    0x1.6a09e6p-1f,   0x1.a9b664p-1f,   0x1.d906bcp-1f,   0x1.f6297ep-1f,
    0x1p+0f,          0x1.f6297cp-1f,   0x1.d906bcp-1f,   0x1.a9b660p-1f,
    0x1.6a09e6p-1f,   0x1.1c73b2p-1f,   0x1.87de20p-2f,   0x1.8f8b82p-3f,
    -0x1.777a5cp-24f, -0x1.8f8b9ap-3f,  -0x1.87de2ap-2f,  -0x1.1c73b6p-1f,
    -0x1.6a09eap-1f,  -0x1.a9b668p-1f,  -0x1.d906c2p-1f,  -0x1.f6297cp-1f,
    -0x1p+0f,         -0x1.f6297cp-1f,  -0x1.d906bap-1f,  -0x1.a9b65ep-1f,
    -0x1.6a09dep-1f,  -0x1.1c73b6p-1f,  -0x1.87de2ap-2f,  -0x1.8f8b78p-3f,
};
#endif

static void hull_direction(int d, real* ux, real* uy) {
#if defined(OR_LITERAL)
    const float angle = 6.28318530718f * (float)d / (float)kFitDirections;
    *ux = cosf(angle);
    *uy = sinf(angle);
#elif defined(OR_FP64)
    const double a = 6.283185307179586 * (double)d / 32.0;
    *ux = cos(a);
    *uy = sin(a);
#else
    *ux = kDirCos[d];
    *uy = kDirSin[d];
#endif
}

/* Normal (-sin t, cos t) of the line through the moments' major axis,
 * t = atan2(2 cxy, cxx - cyy) / 2 (reference gpu_detector.cu:1010-1014).
 * With a = cxx - cyy, b = 2 cxy and r = |(a, b)|, the half-angle identities
 * give cos t = sqrt((r + a) / 2r) when a >= 0 and |sin t| = sqrt((r - a) / 2r)
 * otherwise, each with the other from sin 2t = 2 sin t cos t; both branches
 * avoid cancellation.  t lies in (-pi/2, pi/2], so cos t >= 0 and sin t takes
 * the sign of b (a signed zero b selects +-pi/2 as atan2 does). */
static void axis_normal(real cxx, real cxy, real cyy, real* nx, real* ny) {
    const real a = cxx - cyy;
    const real b = RL(2.0) * cxy;
    const real r = R_SQRT(a * a + b * b);
    real c;
    real s;
    if (!(r > RL(0.0))) {
        c = RL(1.0);
        s = RL(0.0);
    } else if (a >= RL(0.0)) {
        c = R_SQRT((r + a) / (RL(2.0) * r));
        s = b / (RL(2.0) * r * c);
    } else {
        const real sa = R_SQRT((r - a) / (RL(2.0) * r));
        s = signbit(b) ? -sa : sa;
        c = b / (RL(2.0) * r * s);
    }
    *nx = -s;
    *ny = c;
}

/* ------------------------------------------------------------------------ */
/* Frame state                                                               */
/* ------------------------------------------------------------------------ */

typedef struct dec_result {
    int id;
    int hamming;
    real margin;
    real corners[4][2];
} dec_result;

struct or_frame {
    or_config cfg;
    /* resolved, gpu_detector.cu:1904-1943 */
    int width, height, tile_w, tile_h, words, row_cap, total_slots;
    int max_candidates, max_detections, hamming, dim, bits_per_side, code_count;
    int min_cluster_pixels, min_tag_pixels, refine_edges;
    float min_margin, sharpening;
    const uint64_t* codes;

    uint8_t *tile_min, *tile_max, *tile_mean, *tile_thresh, *tile_active;
    uint32_t span_hist[256];
    uint32_t mean_hist[64];
    uint32_t ref[2];
    uint8_t* row_active;
    int32_t *row_start, *row_end;
    uint32_t* bits;
    uint32_t* row_count;
    uint32_t *runs, *parent, *area, *bbox;
    uint32_t counts[5];
    uint64_t* cand_keys;
    uint32_t* candidates;
    int n_candidates;
    double *quads64, *hull64;
    int32_t *dec_id, *dec_hamming;
    double *dec_margin, *dec_corners;
    or_detection* detections;
    double *det_corners, *det_h64;
    int n_detections;
    or_fit_trace* traces; /* per candidate slot, for tests */
};

static int next_power_of_two(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

const uint64_t* or_family_codes(int32_t family, int32_t* count, int32_t* bits_per_side) {
    if (family == 0) {
        if (count) *count = 587;
        if (bits_per_side) *bits_per_side = 6;
        return bd_codes_tag36h11;
    }
    if (family == 1) {
        if (count) *count = 30;
        if (bits_per_side) *bits_per_side = 4;
        return bd_codes_tag16h5;
    }
    return NULL;
}

int32_t or_real_bits(void) { return R_BITS; }

int32_t or_literal(void) {
#ifdef OR_LITERAL
    return 1;
#else
    return 0;
#endif
}

or_frame* or_create(const or_config* c) {
    if (!c || c->family < 0 || c->family > 1 || c->width < 16 || c->height < 16 ||
        c->width > 8192 || c->height > 16384)
        return NULL;
    or_frame* f = (or_frame*)calloc(1, sizeof(or_frame));
    if (!f) return NULL;
    f->cfg = *c;
    /* gpu_detector.cu:1905-1923 */
    f->max_detections = imin(imax(c->max_detections, 1), 256);
    f->hamming = c->max_hamming >= 0 ? imin(c->max_hamming, 3) : (c->family == 0 ? 2 : 1);
    const int w = c->width, h = c->height;
    f->max_candidates = c->max_candidates > 0
                            ? imin(imax(c->max_candidates, 16), 1024)
                            : imin(imax(next_power_of_two((w * h) / 4096) * 3 / 2, 96), 1024);
    /* gpu_detector.cu:1925-1943 */
    f->width = w;
    f->height = h;
    f->tile_w = (w + kTile - 1) / kTile;
    f->tile_h = (h + kTile - 1) / kTile;
    f->words = (w + 31) / 32;
    f->row_cap = w / 2 + 2;
    f->total_slots = h * f->row_cap;
    f->min_cluster_pixels = imax(c->min_cluster_pixels, 1);
    f->min_tag_pixels = imax(c->min_tag_pixels, 4);
    f->codes = or_family_codes(c->family, &f->code_count, &f->bits_per_side);
    f->dim = f->bits_per_side + 2; /* one border cell each side */
    f->refine_edges = c->refine_edges ? 1 : 0;
    f->min_margin = c->min_decision_margin;
    f->sharpening = c->decode_sharpening;

    const size_t tiles = (size_t)f->tile_w * f->tile_h;
    const size_t slots = (size_t)f->total_slots;
    const size_t mc = (size_t)f->max_candidates;
    f->tile_min = (uint8_t*)calloc(tiles, 1);
    f->tile_max = (uint8_t*)calloc(tiles, 1);
    f->tile_mean = (uint8_t*)calloc(tiles, 1);
    f->tile_thresh = (uint8_t*)calloc(tiles, 1);
    f->tile_active = (uint8_t*)calloc(tiles, 1);
    f->row_active = (uint8_t*)calloc((size_t)f->tile_h, 1);
    f->row_start = (int32_t*)calloc((size_t)f->tile_h, sizeof(int32_t));
    f->row_end = (int32_t*)calloc((size_t)f->tile_h, sizeof(int32_t));
    f->bits = (uint32_t*)calloc((size_t)f->words * h, sizeof(uint32_t));
    f->row_count = (uint32_t*)calloc((size_t)h, sizeof(uint32_t));
    f->runs = (uint32_t*)calloc(slots, sizeof(uint32_t));
    f->parent = (uint32_t*)calloc(slots, sizeof(uint32_t));
    f->area = (uint32_t*)calloc(slots, sizeof(uint32_t));
    f->bbox = (uint32_t*)calloc(slots * 4, sizeof(uint32_t));
    f->cand_keys = (uint64_t*)calloc(slots + 1, sizeof(uint64_t));
    f->candidates = (uint32_t*)calloc(mc, sizeof(uint32_t));
    f->quads64 = (double*)calloc(mc * 8, sizeof(double));
    f->hull64 = (double*)calloc(mc * 8, sizeof(double));
    f->dec_id = (int32_t*)calloc(mc, sizeof(int32_t));
    f->dec_hamming = (int32_t*)calloc(mc, sizeof(int32_t));
    f->dec_margin = (double*)calloc(mc, sizeof(double));
    f->dec_corners = (double*)calloc(mc * 8, sizeof(double));
    f->traces = (or_fit_trace*)calloc(mc, sizeof(or_fit_trace));
    f->detections = (or_detection*)calloc((size_t)f->max_detections, sizeof(or_detection));
    f->det_corners = (double*)calloc((size_t)f->max_detections * 8, sizeof(double));
    f->det_h64 = (double*)calloc((size_t)f->max_detections * 9, sizeof(double));
    if (!f->tile_min || !f->tile_max || !f->tile_mean || !f->tile_thresh || !f->tile_active ||
        !f->row_active || !f->row_start || !f->row_end || !f->bits || !f->row_count ||
        !f->runs || !f->parent || !f->area || !f->bbox || !f->cand_keys || !f->candidates ||
        !f->quads64 || !f->hull64 || !f->dec_id || !f->dec_hamming || !f->dec_margin ||
        !f->dec_corners || !f->detections || !f->det_corners || !f->det_h64) {
        or_destroy(f);
        return NULL;
    }
    return f;
}

void or_destroy(or_frame* f) {
    if (!f) return;
    free(f->tile_min);
    free(f->tile_max);
    free(f->tile_mean);
    free(f->tile_thresh);
    free(f->tile_active);
    free(f->row_active);
    free(f->row_start);
    free(f->row_end);
    free(f->bits);
    free(f->row_count);
    free(f->runs);
    free(f->parent);
    free(f->area);
    free(f->bbox);
    free(f->cand_keys);
    free(f->candidates);
    free(f->quads64);
    free(f->hull64);
    free(f->dec_id);
    free(f->dec_hamming);
    free(f->dec_margin);
    free(f->dec_corners);
    free(f->traces);
    free(f->detections);
    free(f->det_corners);
    free(f->det_h64);
    free(f);
}

/* ------------------------------------------------------------------------ */
/* (a) tile statistics, threshold, ROI                                      */
/* ------------------------------------------------------------------------ */

/* gpu_detector.cu:104-135 (k_tile_stats) and :140-166 (k_tile_histograms). */
static void tile_stats(or_frame* f, const uint8_t* gray, int stride) {
    memset(f->span_hist, 0, sizeof f->span_hist);
    memset(f->mean_hist, 0, sizeof f->mean_hist);
    for (int ty = 0; ty < f->tile_h; ++ty) {
        for (int tx = 0; tx < f->tile_w; ++tx) {
            const int idx = ty * f->tile_w + tx;
            const int x0 = tx * kTile, y0 = ty * kTile;
            const int x1 = imin(x0 + kTile, f->width), y1 = imin(y0 + kTile, f->height);
            int mn = 255, mx = 0, sum = 0, count = 0;
            for (int y = y0; y < y1; ++y) {
                const uint8_t* row = gray + (size_t)y * stride;
                for (int x = x0; x < x1; ++x) {
                    const int v = row[x];
                    mn = imin(mn, v);
                    mx = imax(mx, v);
                    sum += v;
                    ++count;
                }
            }
            f->tile_min[idx] = (uint8_t)mn;
            f->tile_max[idx] = (uint8_t)mx;
            f->tile_mean[idx] = (uint8_t)(count > 0 ? (sum + count / 2) / count : 0);
            f->span_hist[(mx - mn) & 0xff] += 1;
            f->mean_hist[f->tile_mean[idx] >> 2] += 1;
        }
    }
}

/* gpu_detector.cu:174-230 (k_tile_reference): lower-quartile span and FP32
 * Otsu split of the 64 tile-mean bins. */
static void tile_reference(or_frame* f) {
    uint32_t span_total = 0;
    for (int i = 0; i < 256; ++i) span_total += f->span_hist[i];
    uint32_t span = 0;
    if (span_total > 0) {
        const uint32_t target = span_total / 4;
        uint32_t seen = 0;
        for (int i = 0; i < 256; ++i) {
            seen += f->span_hist[i];
            if (seen > target) {
                span = (uint32_t)i;
                break;
            }
        }
    }
    f->ref[0] = span;

    float total = 0.0f;
    float sum = 0.0f;
    for (int i = 0; i < 64; ++i) {
        total += (float)f->mean_hist[i];
        sum += (float)i * (float)f->mean_hist[i];
    }
    int otsu = 0;
    if (total > 0.0f) {
        float weight_below = 0.0f;
        float sum_below = 0.0f;
        float best = -1.0f;
        for (int i = 0; i < 64; ++i) {
            weight_below += (float)f->mean_hist[i];
            if (weight_below <= 0.0f) continue;
            const float weight_above = total - weight_below;
            if (weight_above <= 0.0f) break;
            sum_below += (float)i * (float)f->mean_hist[i];
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
    f->ref[1] = (uint32_t)(otsu * 4 + 2);
}

/* gpu_detector.cu:237-276 (k_tile_threshold). */
static void tile_threshold(or_frame* f) {
    const int tw = f->tile_w, th = f->tile_h;
    const int floor_contrast = imax(kMinTileContrast, (int)((f->ref[0] * 3u) / 2u));
    for (int idx = 0; idx < tw * th; ++idx) {
        const int tx = idx % tw, ty = idx / tw;
        int mn = 255, mx = 0;
        for (int dy = -1; dy <= 1; ++dy) {
            const int y = ty + dy;
            if (y < 0 || y >= th) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                const int x = tx + dx;
                if (x < 0 || x >= tw) continue;
                mn = imin(mn, f->tile_min[y * tw + x]);
                mx = imax(mx, f->tile_max[y * tw + x]);
            }
        }
        int t = (int)f->ref[1];
        if (mx - mn >= floor_contrast) t = (mn + mx + 1) >> 1;
        f->tile_thresh[idx] = (uint8_t)t;
        const int own = (int)f->tile_max[idx] - (int)f->tile_min[idx];
        f->tile_active[idx] = own >= kMinTileContrast ? 1u : 0u;
    }
}

/* gpu_detector.cu:282-357 (k_build_row_roi). */
static void build_row_roi(or_frame* f) {
    const int tw = f->tile_w, th = f->tile_h;
    int top = th, bottom = -1;
    for (int i = 0; i < tw * th; ++i) {
        if (f->tile_active[i]) {
            top = imin(top, i / tw);
            bottom = imax(bottom, i / tw);
        }
    }
    f->counts[4] = 0;
    for (int ty = 0; ty < th; ++ty) {
        if (bottom < 0 || ty < top - 1 || ty > bottom + 1) {
            f->row_active[ty] = 0;
            f->row_start[ty] = 0;
            f->row_end[ty] = -1;
            continue;
        }
        int mn = tw, mx = -1;
        for (int x = 0; x < tw; ++x) {
            if (f->tile_active[ty * tw + x]) {
                mn = imin(mn, x);
                mx = imax(mx, x);
            }
        }
        f->row_active[ty] = 1;
        f->counts[4] += 1;
        if (mx < 0) {
            f->row_start[ty] = 0;
            f->row_end[ty] = tw - 1;
        } else {
            f->row_start[ty] = imax(0, mn - 4);
            f->row_end[ty] = imin(tw - 1, mx + 4);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* (b) binarization, runs, union-find, component statistics                 */
/* ------------------------------------------------------------------------ */

/* gpu_detector.cu:362-391 (k_binarize): 1 = white, 32-px words, LSB first. */
static void binarize(or_frame* f, const uint8_t* gray, int stride) {
    for (int y = 0; y < f->height; ++y) {
        const int ty = y >> 3;
        const int row_on = f->row_active[ty] != 0;
        for (int wx = 0; wx < f->words; ++wx) {
            const int tile0 = wx << 2;
            const int overlap =
                row_on && tile0 <= f->row_end[ty] && tile0 + 3 >= f->row_start[ty];
            uint32_t mask = 0xffffffffu;
            if (overlap) {
                mask = 0;
                for (int lane = 0; lane < 32; ++lane) {
                    const int x = (wx << 5) + lane;
                    int white = 1;
                    if (x < f->width) {
                        const int t = f->tile_thresh[ty * f->tile_w + (x >> 3)];
                        white = gray[(size_t)y * stride + x] > t;
                    }
                    if (white) mask |= 1u << lane;
                }
            }
            f->bits[(size_t)y * f->words + wx] = mask;
        }
    }
}

/* gpu_detector.cu:427-504 (k_find_runs): black fragments, split at 32-px words. */
static void find_runs(or_frame* f) {
    f->counts[2] = 0;
    uint32_t* offsets = (uint32_t*)malloc((size_t)f->words * sizeof(uint32_t));
    for (int y = 0; y < f->height; ++y) {
        if (!f->row_active[y >> 3]) {
            f->row_count[y] = 0;
            continue;
        }
        const uint32_t* row_bits = f->bits + (size_t)y * f->words;
        uint32_t off = 0;
        for (int c = 0; c < f->words; ++c) {
            const uint32_t mask = row_bits[c];
            const uint32_t starts = ~mask & ((mask << 1) | 1u);
            offsets[c] = off;
            off += (uint32_t)popc32(starts);
        }
        const uint32_t total = umin(off, (uint32_t)f->row_cap);
        f->row_count[y] = total;
        f->counts[2] += total;
        for (int c = 0; c < f->words; ++c) {
            const uint32_t mask = row_bits[c];
            const uint32_t starts = ~mask & ((mask << 1) | 1u);
            for (int s = 0; s < 32; ++s) {
                if ((starts & (1u << s)) == 0u) continue;
                const int rank = popc32(starts & ((1u << s) - 1u));
                const int x0 = (c << 5) + s;
                if (x0 >= f->width) continue;
                const int last = imin(31, f->width - 1 - (c << 5));
                const int next_white = __builtin_ffs((int)(mask >> s));
                int e = next_white ? (s + next_white - 2) : last;
                e = imin(e, last);
                const int x1 = (c << 5) + e;
                const int index = (int)offsets[c] + rank;
                if (index >= f->row_cap) continue;
                const uint32_t id = (uint32_t)(y * f->row_cap + index);
                f->runs[id] = ((uint32_t)x0 & 0xffffu) | (((uint32_t)x1 & 0xffffu) << 16);
                f->parent[id] = id;
                f->area[id] = 0;
                f->bbox[id * 4 + 0] = 0xffffffffu;
                f->bbox[id * 4 + 1] = 0xffffffffu;
                f->bbox[id * 4 + 2] = 0;
                f->bbox[id * 4 + 3] = 0;
            }
        }
    }
    free(offsets);
}

/* gpu_detector.cu:398-421 (uf_find with path halving, uf_union hooking the
 * larger root under the smaller one), executed serially. */
static uint32_t uf_find(uint32_t* parent, uint32_t x) {
    for (;;) {
        const uint32_t p = parent[x];
        if (p == x) return x;
        const uint32_t gp = parent[p];
        parent[x] = gp;
        x = gp;
    }
}

static void uf_union(uint32_t* parent, uint32_t a, uint32_t b) {
    a = uf_find(parent, a);
    b = uf_find(parent, b);
    while (a != b) {
        if (a > b) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        const uint32_t old = parent[b];
        if (a < old) parent[b] = a;
        if (old == a) break;
        b = uf_find(parent, old);
    }
}

/* gpu_detector.cu:509-557 (k_merge_runs): same-row fragments that touch, and
 * 8-connected overlaps with the previous row. */
static void merge_runs(or_frame* f) {
    for (int y = 1; y < f->height; ++y) {
        if (!f->row_active[y >> 3]) continue;
        const int n = (int)f->row_count[y];
        const int prev_n = (int)f->row_count[y - 1];
        const uint32_t* cur_runs = f->runs + (size_t)y * f->row_cap;
        const uint32_t* prev_runs = f->runs + (size_t)(y - 1) * f->row_cap;
        const uint32_t cur_base = (uint32_t)(y * f->row_cap);
        const uint32_t prev_base = (uint32_t)((y - 1) * f->row_cap);
        for (int i = 0; i < n; ++i) {
            const uint32_t cur = cur_runs[i];
            const int x0 = (int)(cur & 0xffffu);
            const int x1 = (int)(cur >> 16);
            if (i > 0) {
                const int px1 = (int)(cur_runs[i - 1] >> 16);
                if (px1 + 1 >= x0) uf_union(f->parent, cur_base + i, cur_base + i - 1);
            }
            if (prev_n == 0) continue;
            int lo = 0, hi = prev_n - 1, first = prev_n;
            while (lo <= hi) {
                const int mid = (lo + hi) >> 1;
                if ((int)(prev_runs[mid] >> 16) >= x0 - 1) {
                    first = mid;
                    hi = mid - 1;
                } else {
                    lo = mid + 1;
                }
            }
            for (int k = first; k < prev_n; ++k) {
                const int px0 = (int)(prev_runs[k] & 0xffffu);
                if (px0 > x1 + 1) break;
                uf_union(f->parent, cur_base + i, prev_base + k);
            }
        }
    }
    /* Compress every slot to its root so the view exposes the partition. */
    for (int y = 0; y < f->height; ++y) {
        for (int i = 0; i < (int)f->row_count[y]; ++i) {
            const uint32_t id = (uint32_t)(y * f->row_cap + i);
            f->parent[id] = uf_find(f->parent, id);
        }
    }
}

/* gpu_detector.cu:562-586 (k_region_stats). */
static void region_stats(or_frame* f) {
    for (int y = 0; y < f->height; ++y) {
        if (!f->row_active[y >> 3]) continue;
        const uint32_t base = (uint32_t)(y * f->row_cap);
        for (int i = 0; i < (int)f->row_count[y]; ++i) {
            const uint32_t idx = base + (uint32_t)i;
            const uint32_t root = uf_find(f->parent, idx);
            const uint32_t run = f->runs[idx];
            const uint32_t x0 = run & 0xffffu, x1 = run >> 16;
            f->area[root] += x1 - x0 + 1;
            f->bbox[root * 4 + 0] = umin(f->bbox[root * 4 + 0], x0);
            f->bbox[root * 4 + 1] = umin(f->bbox[root * 4 + 1], (uint32_t)y);
            f->bbox[root * 4 + 2] = umax(f->bbox[root * 4 + 2], x1);
            f->bbox[root * 4 + 3] = umax(f->bbox[root * 4 + 3], (uint32_t)y);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* (c) candidates                                                           */
/* ------------------------------------------------------------------------ */

static int cmp_u64(const void* a, const void* b) {
    const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* gpu_detector.cu:591-638 (k_collect_candidates) and :642-654
 * (k_order_candidates).  Buckets are 63 - floor(log2 area), flattened in
 * bucket order; inside a bucket the oracle orders by (area desc, root slot
 * asc), so the whole list is sorted by that key. */
static void collect_candidates(or_frame* f) {
    size_t n = 0;
    const int W = f->width, H = f->height;
    for (int y = 0; y < H; ++y) {
        if (!f->row_active[y >> 3]) continue;
        const uint32_t base = (uint32_t)(y * f->row_cap);
        for (int i = 0; i < (int)f->row_count[y]; ++i) {
            const uint32_t idx = base + (uint32_t)i;
            if (f->parent[idx] != idx) continue;
            const int a = (int)f->area[idx];
            if (a < f->min_cluster_pixels) continue;
            const int x0 = (int)f->bbox[idx * 4 + 0];
            const int y0 = (int)f->bbox[idx * 4 + 1];
            const int x1 = (int)f->bbox[idx * 4 + 2];
            const int y1 = (int)f->bbox[idx * 4 + 3];
            const int bw = x1 - x0 + 1;
            const int bh = y1 - y0 + 1;
            if (bw < f->min_tag_pixels || bh < f->min_tag_pixels) continue;
            if (x0 <= 0 || y0 <= 0 || x1 >= W - 1 || y1 >= H - 1) continue;
            if (bw > 6 * bh || bh > 6 * bw) continue;
            const float fill = (float)a / (float)(bw * bh);
            if (fill < 0.18f || fill > 0.95f) continue;
            f->cand_keys[n++] = ((uint64_t)(~(uint32_t)a) << 32) | idx;
        }
    }
    qsort(f->cand_keys, n, sizeof(uint64_t), cmp_u64);
    f->counts[0] = (uint32_t)n;
    const uint32_t limit = umin((uint32_t)n, (uint32_t)f->max_candidates);
    for (uint32_t i = 0; i < limit; ++i) f->candidates[i] = (uint32_t)(f->cand_keys[i] & 0xffffffffu);
    f->n_candidates = (int)limit;
}

/* ------------------------------------------------------------------------ */
/* (d) boundary, hull, quad, edge refinement                                */
/* ------------------------------------------------------------------------ */

static uint32_t pack_point(int x, int y) {
    return ((uint32_t)x & 0xffffu) | ((uint32_t)y << 16);
}

/* gpu_detector.cu:676-715 (collect_row_extremes). */
static void collect_row_extremes(const or_frame* f, uint32_t root, int bx0, int bx1, int y0,
                                 int rows, int kept, uint32_t* out) {
    for (int s = 0; s < kept; ++s) {
        const int y = y0 + (kept > 1 ? (s * (rows - 1)) / (kept - 1) : 0);
        const int n = (int)f->row_count[y];
        const uint32_t* row_runs = f->runs + (size_t)y * f->row_cap;
        int lo = 0, hi = n - 1, first = n;
        while (lo <= hi) {
            const int mid = (lo + hi) >> 1;
            if ((int)(row_runs[mid] >> 16) >= bx0) {
                first = mid;
                hi = mid - 1;
            } else {
                lo = mid + 1;
            }
        }
        const uint32_t base = (uint32_t)(y * f->row_cap);
        int left = -1, right = -1;
        for (int k = first; k < n; ++k) {
            const int x0 = (int)(row_runs[k] & 0xffffu);
            const int x1 = (int)(row_runs[k] >> 16);
            if (x0 > bx1) break;
            if (f->parent[base + (uint32_t)k] != root) continue;
            left = left < 0 ? x0 : imin(left, x0);
            right = right < 0 ? x1 : imax(right, x1);
        }
        out[2 * s] = left < 0 ? 0xffffffffu : pack_point(left, y);
        out[2 * s + 1] = right < 0 ? 0xffffffffu : pack_point(right, y);
    }
}

/* gpu_detector.cu:736-831 (hull_to_quad). */
static int hull_to_quad(const uint32_t* hull, int count, real quad[4][2]) {
    if (count < 4 || count > kMaxHullVertices) return 0;
    real vx[kMaxHullVertices], vy[kMaxHullVertices];
    for (int i = 0; i < count; ++i) {
        vx[i] = (real)(hull[i] & 0xffffu);
        vy[i] = (real)(hull[i] >> 16);
    }
    while (count > 4) {
        int best = -1;
        real best_area = R_MAX;
        for (int i = 0; i < count; ++i) {
            const int prev = (i + count - 1) % count;
            const int next = (i + 1) % count;
            const real ax = vx[i] - vx[prev];
            const real ay = vy[i] - vy[prev];
            const real bx = vx[next] - vx[i];
            const real by = vy[next] - vy[i];
            const real area = R_FABS(ax * by - ay * bx) * RL(0.5);
            if (area < best_area) {
                best_area = area;
                best = i;
            }
        }
        if (best < 0) return 0;
        for (int i = best; i + 1 < count; ++i) {
            vx[i] = vx[i + 1];
            vy[i] = vy[i + 1];
        }
        --count;
    }
    real area2 = RL(0.0);
    for (int i = 0; i < 4; ++i) {
        const int j = (i + 1) & 3;
        area2 += vx[i] * vy[j] - vx[j] * vy[i];
    }
    if (R_FABS(area2) < RL(24.0)) return 0;
    if (area2 < RL(0.0)) {
        for (int i = 0; i < 2; ++i) {
            const real tx = vx[i + 1], ty = vy[i + 1];
            vx[i + 1] = vx[3 - i];
            vy[i + 1] = vy[3 - i];
            vx[3 - i] = tx;
            vy[3 - i] = ty;
        }
    }
    real min_side = R_MAX, max_side = RL(0.0), perimeter = RL(0.0);
    for (int i = 0; i < 4; ++i) {
        const int j = (i + 1) & 3;
        const real ex = vx[j] - vx[i];
        const real ey = vy[j] - vy[i];
        const real len = R_SQRT(ex * ex + ey * ey);
        if (len < RL(4.0)) return 0;
        min_side = R_FMIN(min_side, len);
        max_side = R_FMAX(max_side, len);
        perimeter += len;
        const int p = (i + 3) & 3;
        const real ax = vx[i] - vx[p];
        const real ay = vy[i] - vy[p];
        const real bx = vx[j] - vx[i];
        const real by = vy[j] - vy[i];
        if (ax * by - ay * bx <= RL(0.0)) return 0;
        const real dot = ax * bx + ay * by;
        const real n1 = R_SQRT(ax * ax + ay * ay);
        const real n2 = R_SQRT(bx * bx + by * by);
        if (n1 < RL(1e-3) || n2 < RL(1e-3)) return 0;
        if (R_FABS(dot / (n1 * n2)) > RL(0.97)) return 0;
    }
    (void)perimeter;
    if (max_side > RL(12.0) * min_side) return 0;
    const real cx = RL(0.25) * (vx[0] + vx[1] + vx[2] + vx[3]);
    const real cy = RL(0.25) * (vy[0] + vy[1] + vy[2] + vy[3]);
    for (int i = 0; i < 4; ++i) {
        const real dx = vx[i] - cx;
        const real dy = vy[i] - cy;
        const real len = R_SQRT(dx * dx + dy * dy);
        if (len > RL(1e-3)) {
            vx[i] += dx / len * RL(0.65);
            vy[i] += dy / len * RL(0.65);
        }
    }
    for (int i = 0; i < 4; ++i) {
        quad[i][0] = vx[i];
        quad[i][1] = vy[i];
    }
    return 1;
}

/* gpu_detector.cu:833-849 (sample_bilinear). */
static real sample_bilinear(const uint8_t* image, int width, int height, int stride, real x,
                            real y) {
    if (!(x >= RL(0.0) && y >= RL(0.0) && x < (real)(width - 1) && y < (real)(height - 1))) {
        return RL(-1.0);
    }
    const int ix = (int)x;
    const int iy = (int)y;
    const real fx = x - (real)ix;
    const real fy = y - (real)iy;
    const uint8_t* p = image + (size_t)iy * stride + ix;
    const real a = p[0], b = p[1], c = p[stride], d = p[stride + 1];
    return (a * (RL(1.0) - fx) + b * fx) * (RL(1.0) - fy) + (c * (RL(1.0) - fx) + d * fx) * fy;
}

/* gpu_detector.cu:851-859 (intersect_lines). */
static int intersect_lines(const real li[4], const real lj[4], real* x, real* y) {
    const real det = li[2] * lj[3] - li[3] * lj[2];
    if (R_FABS(det) < RL(0.2)) return 0;
    const real di = li[2] * li[0] + li[3] * li[1];
    const real dj = lj[2] * lj[0] + lj[3] * lj[1];
    *x = (di * lj[3] - li[3] * dj) / det;
    *y = (li[2] * dj - di * lj[2]) / det;
    return 1;
}

/* gpu_detector.cu:865-874 (directional_rise). */
static real directional_rise(const uint8_t* image, int width, int height, int stride, real x,
                             real y, real dx, real dy) {
    const real ahead =
        sample_bilinear(image, width, height, stride, x + dx * RL(0.7), y + dy * RL(0.7));
    const real behind =
        sample_bilinear(image, width, height, stride, x - dx * RL(0.7), y - dy * RL(0.7));
    if (ahead < RL(0.0) || behind < RL(0.0)) return RL(0.0);
    return ahead - behind;
}

/* gpu_detector.cu:881-1016 (fit_edge_line_warp), with the 32 lanes of the
 * warp executed in turn and the shuffle-down tree of :993-1000 replayed in
 * the same pairing order.  Returns 1 and writes line when lane 0 fits. */
static int fit_edge_line(const real quad[4][2], int edge, const uint8_t* image, int width,
                         int height, int stride, int dim, int first_pass, real line[4],
                         int32_t* best_out) {
    const int b = (edge + 1) & 3;
    const real ax = quad[edge][0], ay = quad[edge][1];
    const real bx = quad[b][0], by = quad[b][1];
    const real ex = bx - ax;
    const real ey = by - ay;
    const real len = R_SQRT(ex * ex + ey * ey);
    if (len < RL(8.0)) return 0;
    const real nx = ey / len;
    const real ny = -ex / len;
    const real cell = len / (real)dim;
    const real range = first_pass ? R_FMIN(R_FMAX(RL(1.15) * cell, RL(1.5)), RL(2.5) * cell)
                                  : R_FMIN(R_FMAX(RL(0.30) * cell, RL(0.8)), RL(2.0));
    const int half = imin(imax(2, (int)R_CEIL(range / RL(0.75))), kFitProbeHalf);
    const int probes = 2 * half + 1;
    const real step = range / (real)half;
    const int samples = imin(imax((int)(len / RL(6.0)), 8), 16);
    /* Moments are taken about the edge midpoint instead of the image origin
     * (the reference sums raw x, x^2 at gpu_detector.cu:983-987): the same
     * line, without the cancellation that raw second moments of coordinates
     * near 1000 suffer in FP32. */
#ifdef OR_LITERAL
    const real rx = RL(0.0), ry = RL(0.0);
#else
    const real rx = ax + RL(0.5) * ex, ry = ay + RL(0.5) * ey;
#endif

    real lm[32][5];
    int lused[32];
    for (int lane = 0; lane < 32; ++lane) {
        real mx = RL(0.0), my = RL(0.0), mxx = RL(0.0), mxy = RL(0.0), myy = RL(0.0);
        int used = 0;
        if (lane < samples) {
            const real alpha = (real)(lane + 1) / (real)(samples + 1);
            const real px = ax + alpha * ex;
            const real py = ay + alpha * ey;
            real mags[kFitProbeHalf * 2 + 1];
            real best_mag = RL(0.0);
            int best = -1;
            for (int i = 0; i < probes; ++i) {
                const real t = step * (real)(i - half);
                const real mag = directional_rise(image, width, height, stride, px + nx * t,
                                                  py + ny * t, nx, ny);
                mags[i] = mag;
                if (mag > best_mag) {
                    best_mag = mag;
                    best = i;
                }
            }
            if (first_pass && best > 2) {
                const real accept = RL(0.70) * best_mag;
                for (int i = 1; i + 1 < best; ++i) {
                    if (mags[i] < accept || mags[i] < mags[i - 1] || mags[i] < mags[i + 1]) {
                        continue;
                    }
                    real valley = R_MAX;
                    for (int k = i + 1; k < best; ++k) valley = R_FMIN(valley, mags[k]);
                    if (valley < RL(0.5) * mags[i]) {
                        best = i;
                        best_mag = mags[i];
                        break;
                    }
                }
            }
            if (best_out) best_out[lane] = (best >= 0 && best_mag >= RL(6.0)) ? best : -1;
            if (best >= 0 && best_mag >= RL(6.0)) {
                real t = step * (real)(best - half);
                if (best > 0 && best < probes - 1) {
                    const real m0 = directional_rise(image, width, height, stride,
                                                     px + nx * (t - step), py + ny * (t - step),
                                                     nx, ny);
                    const real m2 = directional_rise(image, width, height, stride,
                                                     px + nx * (t + step), py + ny * (t + step),
                                                     nx, ny);
                    const real den = m0 - RL(2.0) * best_mag + m2;
                    if (den < RL(-1e-3)) {
                        t += R_FMIN(R_FMAX(RL(0.5) * (m0 - m2) / den, RL(-0.5)), RL(0.5)) * step;
                    }
                }
                const real x = px + nx * t - rx;
                const real y = py + ny * t - ry;
                mx = x;
                my = y;
                mxx = x * x;
                mxy = x * y;
                myy = y * y;
                used = 1;
            }
        }
        lm[lane][0] = mx;
        lm[lane][1] = my;
        lm[lane][2] = mxx;
        lm[lane][3] = mxy;
        lm[lane][4] = myy;
        lused[lane] = used;
    }
    /* Shuffle-down tree: after the step with offset `off`, lane i holds
     * lane i + lane i+off for every i + off < 32; lane 0 only ever reads
     * lanes that are updated this way. */
    for (int off = 16; off > 0; off >>= 1) {
        for (int i = 0; i + off < 32; ++i) {
            if (i >= off) break;
            for (int k = 0; k < 5; ++k) lm[i][k] = lm[i][k] + lm[i + off][k];
            lused[i] += lused[i + off];
        }
    }
    if (lused[0] < 4) return 0;
    const real inv = RL(1.0) / (real)lused[0];
    const real cx = lm[0][0] * inv;
    const real cy = lm[0][1] * inv;
    const real cxx = lm[0][2] * inv - cx * cx;
    const real cxy = lm[0][3] * inv - cx * cy;
    const real cyy = lm[0][4] * inv - cy * cy;
    real lnx, lny;
#ifdef OR_LITERAL
    const float theta = 0.5f * atan2f(2.0f * cxy, cxx - cyy);
    lnx = -sinf(theta);
    lny = cosf(theta);
#else
    axis_normal(cxx, cxy, cyy, &lnx, &lny);
#endif
    line[0] = rx + cx;
    line[1] = ry + cy;
    line[2] = lnx;
    line[3] = lny;
    return 1;
}

/* gpu_detector.cu:1021-1076 (apply_refined_lines). */
static real apply_refined_lines(real quad[4][2], real lines[4][4], const int have[4]) {
    int line_count = 0;
    for (int i = 0; i < 4; ++i) line_count += have[i] ? 1 : 0;
    if (line_count < 3) return RL(0.0);
    real refined[4][2];
    for (int i = 0; i < 4; ++i) {
        refined[i][0] = quad[i][0];
        refined[i][1] = quad[i][1];
    }
    real extent = RL(0.0);
    for (int i = 0; i < 4; ++i) {
        const int j = (i + 1) & 3;
        extent = R_FMAX(extent, or_hypot(quad[j][0] - quad[i][0], quad[j][1] - quad[i][1]));
    }
    const real max_shift = R_FMIN(R_FMAX(RL(0.15) * extent, RL(4.0)), RL(16.0));
    real moved = RL(0.0);
    for (int i = 0; i < 4; ++i) {
        const int j = (i + 1) & 3;
        if (!have[i] || !have[j]) continue;
        real x, y;
        if (!intersect_lines(lines[i], lines[j], &x, &y)) continue;
        const int corner = j;
        const real dx = x - quad[corner][0];
        const real dy = y - quad[corner][1];
        if (dx * dx + dy * dy > max_shift * max_shift) continue;
        refined[corner][0] = x;
        refined[corner][1] = y;
        moved = R_FMAX(moved, R_SQRT(dx * dx + dy * dy));
    }
    for (int i = 0; i < 4; ++i) {
        const int j = (i + 1) & 3;
        if (or_hypot(refined[j][0] - refined[i][0], refined[j][1] - refined[i][1]) < RL(3.0)) {
            return RL(0.0);
        }
    }
    real area2 = RL(0.0);
    for (int i = 0; i < 4; ++i) {
        const int j = (i + 1) & 3;
        area2 += refined[i][0] * refined[j][1] - refined[j][0] * refined[i][1];
    }
    if (area2 <= RL(0.0)) return RL(0.0);
    for (int i = 0; i < 4; ++i) {
        quad[i][0] = refined[i][0];
        quad[i][1] = refined[i][1];
    }
    return moved;
}

/* gpu_detector.cu:1081-1241 (k_fit_quads) for one candidate.  Returns 1 when
 * a quad was produced; hull_out gets the unrefined hull quad. */
static int fit_quad(const or_frame* f, const uint8_t* gray, int stride, uint32_t root,
                    real out[4][2], real hull_out[4][2], uint32_t* boundary, or_fit_trace* tr) {
    const int x0 = (int)f->bbox[root * 4 + 0];
    const int y0 = (int)f->bbox[root * 4 + 1];
    const int x1 = (int)f->bbox[root * 4 + 2];
    const int y1 = (int)f->bbox[root * 4 + 3];
    const int rows = y1 - y0 + 1;
    const int kept = imin(rows, kMaxBoundaryRows);
    const int points = 2 * kept;
    for (int i = 0; i < 2 * kMaxBoundaryRows; ++i) boundary[i] = 0xffffffffu;
    collect_row_extremes(f, root, x0, x1, y0, rows, kept, boundary);

    /* gpu_detector.cu:1132-1170: four warps each take a contiguous chunk of
     * the boundary; the first strictly larger projection wins in each chunk,
     * then across chunks. */
    uint32_t extreme[kFitDirections];
    const int chunk = (points + kFitWarps - 1) / kFitWarps;
    for (int dir = 0; dir < kFitDirections; ++dir) {
        real ux, uy;
        hull_direction(dir, &ux, &uy);
        real best = -R_MAX;
        uint32_t point = 0xffffffffu;
        for (int part = 0; part < kFitWarps; ++part) {
            const int begin = part * chunk;
            const int end = imin(points, begin + chunk);
            real best_dot = -R_MAX;
            uint32_t best_point = 0xffffffffu;
            for (int i = begin; i < end; ++i) {
                const uint32_t p = boundary[i];
                if (p == 0xffffffffu) continue;
                const real d = (real)(p & 0xffffu) * ux + (real)(p >> 16) * uy;
                if (d > best_dot) {
                    best_dot = d;
                    best_point = p;
                }
            }
            if (best_dot > best) {
                best = best_dot;
                point = best_point;
            }
        }
        extreme[dir] = point;
    }
    /* gpu_detector.cu:1172-1180: ordered hull without repeats. */
    uint32_t hull[kFitDirections];
    int hc = 0;
    for (int i = 0; i < kFitDirections; ++i) {
        const uint32_t p = extreme[i];
        if (p == 0xffffffffu) continue;
        if (hc > 0 && hull[hc - 1] == p) continue;
        hull[hc++] = p;
    }
    while (hc > 1 && hull[0] == hull[hc - 1]) --hc;

    real quad[4][2];
    if (tr) memset(tr, 0, sizeof *tr);
    if (!hull_to_quad(hull, hc, quad)) return 0;
    if (tr) tr->fitted = 1;
    for (int i = 0; i < 4; ++i) {
        hull_out[i][0] = quad[i][0];
        hull_out[i][1] = quad[i][1];
    }
    if (f->refine_edges) {
        /* gpu_detector.cu:1200-1233 */
        for (int iteration = 0; iteration < kFitIterations; ++iteration) {
            real lines[4][4];
            int have[4];
            for (int e = 0; e < 4; ++e) {
                int32_t* best = tr ? tr->best[iteration][e] : NULL;
                if (best) for (int k = 0; k < 16; ++k) best[k] = -1;
                have[e] = fit_edge_line((const real(*)[2])quad, e, gray, f->width, f->height,
                                        stride, f->dim, iteration == 0, lines[e], best);
                if (tr) tr->have[iteration][e] = have[e];
            }
            real before[4][2];
            memcpy(before, quad, sizeof before);
            const real moved = apply_refined_lines(quad, lines, have);
            if (tr) {
                tr->passes = iteration + 1;
                for (int j = 0; j < 4; ++j) {
                    if (quad[j][0] != before[j][0] || quad[j][1] != before[j][1]) tr->refined[iteration] |= 1 << j;
                }
            }
            if (moved < RL(0.15)) break;
        }
    }
    for (int i = 0; i < 4; ++i) {
        out[i][0] = quad[i][0];
        out[i][1] = quad[i][1];
    }
    return 1;
}

/* ------------------------------------------------------------------------ */
/* (e) decode                                                                */
/* ------------------------------------------------------------------------ */

/* gpu_detector.cu:1249-1272 (homography_from_unit_quad). */
static int homography_from_unit_quad(const real quad[4][2], real h[9]) {
    const real x0 = quad[0][0], y0 = quad[0][1];
    const real x1 = quad[1][0], y1 = quad[1][1];
    const real x2 = quad[2][0], y2 = quad[2][1];
    const real x3 = quad[3][0], y3 = quad[3][1];
    const real dx1 = x1 - x2, dx2 = x3 - x2, dx3 = x0 - x1 + x2 - x3;
    const real dy1 = y1 - y2, dy2 = y3 - y2, dy3 = y0 - y1 + y2 - y3;
    const real det = dx1 * dy2 - dx2 * dy1;
    if (R_FABS(det) < RL(1e-9)) return 0;
    const real g = (dx3 * dy2 - dx2 * dy3) / det;
    const real k = (dx1 * dy3 - dx3 * dy1) / det;
    h[0] = x1 - x0 + g * x1;
    h[1] = x3 - x0 + k * x3;
    h[2] = x0;
    h[3] = y1 - y0 + g * y1;
    h[4] = y3 - y0 + k * y3;
    h[5] = y0;
    h[6] = g;
    h[7] = k;
    h[8] = RL(1.0);
    return 1;
}

#ifndef OR_FP64
/* Precision policy of the decode homography (no counterpart in the
 * reference, which solves it in FP32 always): FP32 rounding in the sample
 * positions x = (h0 u + h1 v + h2) / (g u + k v + 1) grows with the
 * determinant's cancellation (|dx1 dy2| + |dx2 dy1|) / |det| and with
 * (1 + |g| + |k|) / min |g u + k v + 1| over the quad's corners.  When their
 * product exceeds 16 (0.2 % of the quads of heavily tilted scenes; 1 for a
 * square seen head-on) the homography and the sample positions are
 * computed in FP64 instead.  The measure itself is FP32, so the kernels and
 * this file take the same decision. */
static int decode_needs_fp64(const float q[4][2]) {
    const float dx1 = q[1][0] - q[2][0], dx2 = q[3][0] - q[2][0], dx3 = q[0][0] - q[1][0] + q[2][0] - q[3][0];
    const float dy1 = q[1][1] - q[2][1], dy2 = q[3][1] - q[2][1], dy3 = q[0][1] - q[1][1] + q[2][1] - q[3][1];
    const float det = dx1 * dy2 - dx2 * dy1;
    const float g = (dx3 * dy2 - dx2 * dy3) / det;
    const float k = (dx1 * dy3 - dx3 * dy1) / det;
    const float kappa = (fabsf(dx1 * dy2) + fabsf(dx2 * dy1)) / fabsf(det);
    const float den = fminf(1.0f, fminf(fabsf(1.0f + g), fminf(fabsf(1.0f + g + k), fabsf(1.0f + k))));
    const float amplification = kappa * (1.0f + fabsf(g) + fabsf(k)) / den;
    return !(amplification <= 16.0f);
}

/* homography_from_unit_quad in FP64 from FP32 corners. */
static void homography_from_unit_quad_f64(const float q[4][2], double h[9]) {
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

static long g_decode_fp64_count; /* quads that took the FP64 path (tests) */
#endif

/* gpu_detector.cu:1274-1289 (match_code). */
static int match_code(const uint64_t* codes, int count, uint64_t code, int max_hamming,
                      int* out_hamming) {
    int best = max_hamming + 1;
    int best_id = -1;
    for (int i = 0; i < count; ++i) {
        const int d = popc64(codes[i] ^ code);
        if (d < best) {
            best = d;
            best_id = i;
            if (d == 0) break;
        }
    }
    if (best_id < 0) return -1;
    *out_hamming = best;
    return best_id;
}

/* gpu_detector.cu:1291-1430 (decode_quad) for one rotation. */
static int decode_quad(const or_frame* f, const uint8_t* gray, int stride,
                       const real base_quad[4][2], int rot, int* out_id, int* out_hamming,
                       real* out_margin, real out_quad[4][2]) {
    const int dim = f->dim;
    real raw[64], diff[64], sharp[64];
    real corners[4][2];
    for (int k = 0; k < 4; ++k) {
        const int src = (rot + k) & 3;
        corners[k][0] = base_quad[src][0];
        corners[k][1] = base_quad[src][1];
    }
    real h[9];
    if (!homography_from_unit_quad((const real(*)[2])corners, h)) return 0;
#ifndef OR_FP64
    double h64[9];
    const int wide = decode_needs_fp64((const float(*)[2])corners);
    if (wide) {
        homography_from_unit_quad_f64((const float(*)[2])corners, h64);
        __atomic_fetch_add(&g_decode_fp64_count, 1, __ATOMIC_RELAXED);
    }
#endif
    for (int cy = 0; cy < dim; ++cy) {
        for (int cx = 0; cx < dim; ++cx) {
            const real u = ((real)cx + RL(0.5)) / (real)dim;
            const real v = ((real)cy + RL(0.5)) / (real)dim;
            real x, y;
#ifndef OR_FP64
            if (wide) {
                const double den = h64[6] * (double)u + h64[7] * (double)v + 1.0;
                if (fabs(den) < 1e-9) return 0;
                x = (float)((h64[0] * (double)u + h64[1] * (double)v + h64[2]) / den);
                y = (float)((h64[3] * (double)u + h64[4] * (double)v + h64[5]) / den);
            } else
#endif
            {
                const real den = h[6] * u + h[7] * v + RL(1.0);
                if (R_FABS(den) < RL(1e-9)) return 0;
                x = (h[0] * u + h[1] * v + h[2]) / den;
                y = (h[3] * u + h[4] * v + h[5]) / den;
            }
            const real value =
                sample_bilinear(gray, f->width, f->height, stride, x - (real)0, y - (real)0);
            if (value < RL(0.0)) return 0;
            raw[cy * dim + cx] = value;
        }
    }
    real mn = RL(255.0), mx = RL(0.0);
    for (int i = 0; i < dim * dim; ++i) {
        mn = R_FMIN(mn, raw[i]);
        mx = R_FMAX(mx, raw[i]);
    }
    if (mx - mn < RL(10.0)) return 0;
    const real thresh = RL(0.5) * (mn + mx);
    for (int cy = 0; cy < dim; ++cy) {
        for (int cx = 0; cx < dim; ++cx) {
            if (cx == 0 || cy == 0 || cx == dim - 1 || cy == dim - 1) {
                if (raw[cy * dim + cx] >= thresh) return 0;
            }
        }
    }
    for (int i = 0; i < dim * dim; ++i) diff[i] = raw[i] - thresh;
    for (int i = 0; i < dim * dim; ++i) sharp[i] = diff[i];
    const real sharpening = (real)f->sharpening;
    for (int cy = 1; cy <= dim - 2; ++cy) {
        for (int cx = 1; cx <= dim - 2; ++cx) {
            const int i = cy * dim + cx;
            const real c = diff[i];
            const real lap = RL(4.0) * c - diff[i - dim] - diff[i + dim] - diff[i - 1] - diff[i + 1];
            sharp[i] = c + sharpening * lap;
        }
    }
    uint64_t code = 0;
    real white_sum = RL(0.0), black_sum = RL(0.0);
    int white_count = 0, black_count = 0;
    for (int cy = 1; cy <= dim - 2; ++cy) {
        for (int cx = 1; cx <= dim - 2; ++cx) {
            const real d = sharp[cy * dim + cx];
            code = (code << 1) | (d > RL(0.0) ? 1ull : 0ull);
            if (d > RL(0.0)) {
                white_sum += d;
                ++white_count;
            } else {
                black_sum -= d;
                ++black_count;
            }
        }
    }
    int hamming = 0;
    const int id = match_code(f->codes, f->code_count, code, f->hamming, &hamming);
    if (id < 0) return 0;
    const real white_avg = white_sum / (real)imax(white_count, 1);
    const real black_avg = black_sum / (real)imax(black_count, 1);
    const real margin = R_FMIN(white_avg, black_avg);
    if (margin < (real)f->min_margin) return 0;
    const real score = (real)(4 - hamming) * RL(1000.0) + margin;
    if (!(score > RL(-1.0))) return 0;
    *out_id = id;
    *out_hamming = hamming;
    *out_margin = margin;
    for (int k = 0; k < 4; ++k) {
        out_quad[k][0] = corners[k][0];
        out_quad[k][1] = corners[k][1];
    }
    return 1;
}

/* gpu_detector.cu:1438-1537 (k_decode_quads) for one candidate: four
 * rotations of the refined quad, then of the hull quad only if none of the
 * refined ones decoded; highest score wins, first on ties. */
static int decode_candidate(const or_frame* f, const uint8_t* gray, int stride,
                            const real quad[4][2], const real hull[4][2], dec_result* out,
                            int32_t* variant) {
    dec_result results[8];
    for (int lane = 0; lane < 8; ++lane) {
        const real(*source)[2] = (lane >> 2) == 0 ? quad : hull;
        results[lane].id = -1;
        int id = 0, hamming = 0;
        real margin = RL(0.0);
        real corners[4][2];
        if (decode_quad(f, gray, stride, source, lane & 3, &id, &hamming, &margin, corners)) {
            results[lane].id = id;
            results[lane].hamming = hamming;
            results[lane].margin = margin;
            memcpy(results[lane].corners, corners, sizeof corners);
        }
    }
    int best = -1;
    real best_score = RL(-1.0);
    for (int variant = 0; variant < 2 && best < 0; ++variant) {
        for (int rotation = 0; rotation < 4; ++rotation) {
            const int r = variant * 4 + rotation;
            if (results[r].id < 0) continue;
            const real score = (real)(4 - results[r].hamming) * RL(1000.0) + results[r].margin;
            if (score > best_score) {
                best_score = score;
                best = r;
            }
        }
    }
    if (variant) *variant = best;
    if (best < 0) return 0;
    *out = results[best];
    return 1;
}

/* gpu_detector.cu:1739-1774 (fill_homography), FP64 on the host. */
static void fill_homography_both(const float corners[8], float H[9], double H64[9]) {
    const double x0 = corners[0], y0 = corners[1];
    const double x1 = corners[2], y1 = corners[3];
    const double x2 = corners[4], y2 = corners[5];
    const double x3 = corners[6], y3 = corners[7];
    const double dx1 = x1 - x2, dx2 = x3 - x2;
    const double dx3 = x0 - x1 + x2 - x3;
    const double dy1 = y1 - y2, dy2 = y3 - y2;
    const double dy3 = y0 - y1 + y2 - y3;
    double g = 0.0, h = 0.0;
    const double det = dx1 * dy2 - dx2 * dy1;
    if (fabs(det) > 1e-12) {
        g = (dx3 * dy2 - dx2 * dy3) / det;
        h = (dx1 * dy3 - dx3 * dy1) / det;
    }
    const double a = x1 - x0 + g * x1;
    const double b = x3 - x0 + h * x3;
    const double d = y1 - y0 + g * y1;
    const double e = y3 - y0 + h * y3;
    const double w = g + h + 2.0;
    const double inv = (fabs(w) > 1e-12) ? 1.0 / w : 0.0;
    const double v[9] = {a * inv,
                         b * inv,
                         (a + b + 2.0 * x0) * inv,
                         d * inv,
                         e * inv,
                         (d + e + 2.0 * y0) * inv,
                         g * inv,
                         h * inv,
                         w * inv};
    for (int i = 0; i < 9; ++i) {
        if (H) H[i] = (float)v[i];
        if (H64) H64[i] = v[i];
    }
}

void or_fill_homography(const float corners[8], float out[9], double out64[9]) {
    fill_homography_both(corners, out, out64);
}

/* gpu_detector.cu:1542-1601 (k_finalize): de-duplicate same-id detections
 * whose centres are within 0.35 of the mean side, keeping the higher margin,
 * then a stable insertion sort by id. */
typedef struct fin_entry {
    or_detection d;
    real corners[8];
    real margin;
} fin_entry;

static void finalize(or_frame* f, const dec_result* list, int count) {
    const int n = imin(count, f->max_detections);
    fin_entry* out = (fin_entry*)calloc((size_t)f->max_detections, sizeof(fin_entry));
    int kept = 0;
    for (int i = 0; i < n; ++i) {
        fin_entry e;
        memset(&e, 0, sizeof e);
        e.d.id = list[i].id;
        e.d.hamming = list[i].hamming;
        e.d.decision_margin = (float)list[i].margin;
        e.margin = list[i].margin;
        for (int k = 0; k < 4; ++k) {
            e.corners[k * 2] = list[i].corners[k][0];
            e.corners[k * 2 + 1] = list[i].corners[k][1];
            e.d.corners[k * 2] = (float)list[i].corners[k][0];
            e.d.corners[k * 2 + 1] = (float)list[i].corners[k][1];
        }
        e.d.reserved = 0;
        real side = RL(0.0), cx = RL(0.0), cy = RL(0.0);
        for (int k = 0; k < 4; ++k) {
            const int j = (k + 1) & 3;
            side += or_hypot(e.corners[j * 2] - e.corners[k * 2],
                             e.corners[j * 2 + 1] - e.corners[k * 2 + 1]);
            cx += e.corners[k * 2];
            cy += e.corners[k * 2 + 1];
        }
        side *= RL(0.25);
        cx *= RL(0.25);
        cy *= RL(0.25);
        int duplicate = -1;
        for (int k = 0; k < kept; ++k) {
            if (out[k].d.id != e.d.id) continue;
            real ocx = RL(0.0), ocy = RL(0.0);
            for (int c = 0; c < 4; ++c) {
                ocx += out[k].corners[c * 2];
                ocy += out[k].corners[c * 2 + 1];
            }
            ocx *= RL(0.25);
            ocy *= RL(0.25);
            const real limit = RL(0.35) * side;
            if (R_FABS(ocx - cx) < limit && R_FABS(ocy - cy) < limit) {
                duplicate = k;
                break;
            }
        }
        if (duplicate >= 0) {
            if (e.margin > out[duplicate].margin) out[duplicate] = e;
        } else if (kept < f->max_detections) {
            out[kept++] = e;
        }
    }
    for (int i = 1; i < kept; ++i) {
        const fin_entry value = out[i];
        int j = i - 1;
        while (j >= 0 && out[j].d.id > value.d.id) {
            out[j + 1] = out[j];
            --j;
        }
        out[j + 1] = value;
    }
    for (int i = 0; i < kept; ++i) {
        f->detections[i] = out[i].d;
        fill_homography_both(out[i].d.corners, f->detections[i].homography, f->det_h64 + i * 9);
        for (int k = 0; k < 8; ++k) f->det_corners[i * 8 + k] = (double)out[i].corners[k];
    }
    f->counts[1] = (uint32_t)kept;
    f->n_detections = kept;
    free(out);
}

/* ------------------------------------------------------------------------ */
/* Pipeline                                                                  */
/* ------------------------------------------------------------------------ */

int32_t or_run(or_frame* f, const uint8_t* gray, int32_t stride) {
    if (!f || !gray || stride < f->width) return -1;
    memset(f->counts, 0, sizeof f->counts);
    tile_stats(f, gray, stride);
    tile_reference(f);
    tile_threshold(f);
    build_row_roi(f);
    binarize(f, gray, stride);
    find_runs(f);
    merge_runs(f);
    region_stats(f);
    collect_candidates(f);

    const int mc = f->max_candidates;
    const double nan = (double)NAN;
    for (int i = 0; i < mc * 8; ++i) {
        f->quads64[i] = nan;
        f->hull64[i] = nan;
        f->dec_corners[i] = nan;
    }
    for (int i = 0; i < mc; ++i) {
        f->dec_id[i] = -1;
        f->dec_hamming[i] = 0;
        f->dec_margin[i] = 0.0;
    }
    uint32_t* boundary = (uint32_t*)malloc(2 * kMaxBoundaryRows * sizeof(uint32_t));
    dec_result* list = (dec_result*)calloc((size_t)(f->n_candidates + 1), sizeof(dec_result));
    int decoded = 0;
    uint32_t quads = 0;
    for (int ci = 0; ci < f->n_candidates; ++ci) {
        real quad[4][2], hull[4][2];
        or_fit_trace* tr = f->traces + ci;
        tr->decode_variant = -1;
        if (!fit_quad(f, gray, stride, f->candidates[ci], quad, hull, boundary, tr)) continue;
        ++quads;
        for (int i = 0; i < 4; ++i) {
            f->quads64[ci * 8 + i * 2] = (double)quad[i][0];
            f->quads64[ci * 8 + i * 2 + 1] = (double)quad[i][1];
            f->hull64[ci * 8 + i * 2] = (double)hull[i][0];
            f->hull64[ci * 8 + i * 2 + 1] = (double)hull[i][1];
        }
        dec_result r;
        if (decode_candidate(f, gray, stride, (const real(*)[2])quad, (const real(*)[2])hull,
                             &r, &tr->decode_variant)) {
            f->dec_id[ci] = r.id;
            f->dec_hamming[ci] = r.hamming;
            f->dec_margin[ci] = (double)r.margin;
            for (int k = 0; k < 4; ++k) {
                f->dec_corners[ci * 8 + k * 2] = (double)r.corners[k][0];
                f->dec_corners[ci * 8 + k * 2 + 1] = (double)r.corners[k][1];
            }
            list[decoded++] = r;
        }
    }
    f->counts[3] = quads;
    /* Detections beyond max_detections are dropped at emission
     * (gpu_detector.cu:1525-1526); counts[1] then holds the kept count. */
    finalize(f, list, decoded);
    free(list);
    free(boundary);
    return f->n_detections;
}

void or_get_view(const or_frame* f, or_view* v) {
    memset(v, 0, sizeof *v);
    v->width = f->width;
    v->height = f->height;
    v->tile_w = f->tile_w;
    v->tile_h = f->tile_h;
    v->words = f->words;
    v->row_cap = f->row_cap;
    v->max_candidates = f->max_candidates;
    v->max_detections = f->max_detections;
    v->max_hamming = f->hamming;
    v->dim = f->dim;
    v->code_count = f->code_count;
    v->min_cluster_pixels = f->min_cluster_pixels;
    v->min_tag_pixels = f->min_tag_pixels;
    v->refine_edges = f->refine_edges;
    v->min_decision_margin = f->min_margin;
    v->decode_sharpening = f->sharpening;
    v->tile_min = f->tile_min;
    v->tile_max = f->tile_max;
    v->tile_mean = f->tile_mean;
    v->tile_thresh = f->tile_thresh;
    v->tile_active = f->tile_active;
    v->span_hist = f->span_hist;
    v->mean_hist = f->mean_hist;
    v->ref[0] = f->ref[0];
    v->ref[1] = f->ref[1];
    v->row_active = f->row_active;
    v->row_start = f->row_start;
    v->row_end = f->row_end;
    v->bits = f->bits;
    v->row_count = f->row_count;
    v->runs = f->runs;
    v->parent = f->parent;
    v->area = f->area;
    v->bbox = f->bbox;
    memcpy(v->counts, f->counts, sizeof v->counts);
    v->candidates = f->candidates;
    v->n_candidates = f->n_candidates;
    v->quads = f->quads64;
    v->hull_quads = f->hull64;
    v->dec_id = f->dec_id;
    v->dec_hamming = f->dec_hamming;
    v->dec_margin = f->dec_margin;
    v->dec_corners = f->dec_corners;
    v->detections = f->detections;
    v->n_detections = f->n_detections;
    v->det_corners = f->det_corners;
    v->det_homography = f->det_h64;
}

int32_t or_hull_direction(int32_t dir, double* ux, double* uy) {
    if (dir < 0 || dir >= kFitDirections) return -1;
    real x, y;
    hull_direction(dir, &x, &y);
    *ux = (double)x;
    *uy = (double)y;
    return 0;
}

void or_axis_normal(double cxx, double cxy, double cyy, double* nx, double* ny) {
    real x, y;
    axis_normal((real)cxx, (real)cxy, (real)cyy, &x, &y);
    *nx = (double)x;
    *ny = (double)y;
}

int32_t or_match_code(int32_t family, uint64_t code, int32_t max_hamming, int32_t* hamming) {
    int32_t count = 0;
    const uint64_t* codes = or_family_codes(family, &count, NULL);
    if (!codes) return -1;
    int h = 0;
    const int id = match_code(codes, count, code, max_hamming, &h);
    if (hamming) *hamming = h;
    return id;
}

/* One edge fit of a frame created by or_create (its size and tag family),
 * on `gray`: quad corners in pixels, edge e runs from corner e to e + 1. */
int32_t or_fit_edge(const or_frame* f, const uint8_t* gray, int32_t stride, const double quad8[8],
                    int32_t edge, int32_t first_pass, double line_out[4]) {
    real quad[4][2];
    for (int i = 0; i < 4; ++i) {
        quad[i][0] = (real)quad8[2 * i];
        quad[i][1] = (real)quad8[2 * i + 1];
    }
    real line[4];
    const int ok = fit_edge_line((const real(*)[2])quad, edge & 3, gray, f->width, f->height, stride, f->dim,
                                 first_pass != 0, line, NULL);
    for (int k = 0; k < 4 && ok; ++k) line_out[k] = (double)line[k];
    return ok;
}

/* decode_quad of one rotation with the frame's family and thresholds;
 * returns 1 with id, hamming and margin when the quad decodes. */
int32_t or_decode_rotation(const or_frame* f, const uint8_t* gray, int32_t stride, const double quad8[8],
                           int32_t rot, int32_t* id, int32_t* hamming, double* margin) {
    real quad[4][2], corners[4][2];
    for (int i = 0; i < 4; ++i) {
        quad[i][0] = (real)quad8[2 * i];
        quad[i][1] = (real)quad8[2 * i + 1];
    }
    int i_id = -1, i_ham = 0;
    real m = RL(0.0);
    const int ok = decode_quad(f, gray, stride, (const real(*)[2])quad, rot & 3, &i_id, &i_ham, &m, corners);
    if (ok) {
        *id = i_id;
        *hamming = i_ham;
        *margin = (double)m;
    }
    return ok;
}

/* hull_to_quad on packed points x | y << 16. */
int32_t or_hull_to_quad(const uint32_t* hull, int32_t count, double quad_out[8]) {
    real quad[4][2];
    const int ok = hull_to_quad(hull, count, quad);
    for (int i = 0; i < 4 && ok; ++i) {
        quad_out[2 * i] = (double)quad[i][0];
        quad_out[2 * i + 1] = (double)quad[i][1];
    }
    return ok;
}

int32_t or_get_fit_trace(const or_frame* f, int32_t candidate, or_fit_trace* out) {
    if (!f || candidate < 0 || candidate >= f->n_candidates) return 0;
    *out = f->traces[candidate];
    return 1;
}

/* Decode quads that took the FP64 homography since the last reset (0 in the
 * FP64 build, where every decode is FP64). */
int64_t or_decode_fp64_count(int32_t reset) {
#ifdef OR_FP64
    (void)reset;
    return 0;
#else
    return reset ? __atomic_exchange_n(&g_decode_fp64_count, 0, __ATOMIC_RELAXED)
                 : __atomic_load_n(&g_decode_fp64_count, __ATOMIC_RELAXED);
#endif
}
