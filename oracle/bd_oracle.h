/*
 * Scalar CPU oracle of the bdcudatag detection pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * detector is compared against; it is loaded by tests/, by the smoke check and
 * by the CPU-baseline leg of bench.py, and by nothing in the product path.
 *
 * It restates the reference kernels of detector/src/gpu_detector.cu one stage
 * at a time, in pipeline order, and keeps every intermediate buffer so that
 * each stage can be compared on its own.  bd_oracle.c cites, per function, the
 * reference lines it follows.
 *
 * The same source is compiled twice: with OR_REAL=float (the parity oracle)
 * and with OR_REAL=double (the FP64 twin used for the precision report).  Only
 * the geometric stages (quad fit, edge refinement, decode sampling) change
 * precision; threshold, binarization, components and candidate filtering are
 * integer or FP32 in both builds.
 */
#ifndef BD_ORACLE_H
#define BD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Raw configuration, same meaning as bd_config (clamping happens inside). */
typedef struct or_config {
    int32_t family; /* 0 = tag36h11, 1 = tag16h5 */
    int32_t width;
    int32_t height;
    int32_t max_hamming;
    int32_t min_cluster_pixels;
    int32_t min_tag_pixels;
    int32_t max_detections;
    int32_t max_candidates;
    int32_t refine_edges;
    float min_decision_margin;
    float decode_sharpening;
} or_config;

/* Same 84-byte layout as bd_detection. */
typedef struct or_detection {
    int32_t id;
    int32_t hamming;
    float decision_margin;
    float corners[8];
    int32_t reserved;
    float homography[9];
} or_detection;

/* Read-only view of one frame's per-stage results. */
typedef struct or_view {
    /* resolved sizes and parameters */
    int32_t width, height, tile_w, tile_h, words, row_cap;
    int32_t max_candidates, max_detections, max_hamming, dim, code_count;
    int32_t min_cluster_pixels, min_tag_pixels, refine_edges;
    float min_decision_margin, decode_sharpening;
    /* (a) threshold and ROI */
    const uint8_t* tile_min;
    const uint8_t* tile_max;
    const uint8_t* tile_mean;
    const uint8_t* tile_thresh;
    const uint8_t* tile_active;
    const uint32_t* span_hist; /* 256 */
    const uint32_t* mean_hist; /* 64 */
    uint32_t ref[2];
    const uint8_t* row_active;  /* tile_h */
    const int32_t* row_start;   /* tile_h */
    const int32_t* row_end;     /* tile_h */
    /* (b) bit plane and components (reference 32-px fragment layout) */
    const uint32_t* bits;       /* height * words, bit x&31 of word x>>5, 1 = white */
    const uint32_t* row_count;  /* height */
    const uint32_t* runs;       /* height * row_cap, x0 | x1 << 16 */
    const uint32_t* parent;     /* height * row_cap, fully compressed roots */
    const uint32_t* area;       /* height * row_cap, valid at roots */
    const uint32_t* bbox;       /* height * row_cap * 4, valid at roots */
    uint32_t counts[5];         /* candidates, detections, runs, quads, roi tile rows */
    /* (c) ordered candidates */
    const uint32_t* candidates; /* root slots, n_candidates */
    int32_t n_candidates;
    /* (d) quads per candidate slot, NaN when the fit failed; widened to double */
    const double* quads;        /* max_candidates * 8 */
    const double* hull_quads;   /* max_candidates * 8 */
    /* (e) per-candidate decode result (before de-duplication) */
    const int32_t* dec_id;      /* max_candidates, -1 = none */
    const int32_t* dec_hamming;
    const double* dec_margin;
    const double* dec_corners;  /* max_candidates * 8 */
    /* final output */
    const or_detection* detections;
    int32_t n_detections;
    const double* det_corners;     /* n_detections * 8, at the build's precision */
    const double* det_homography;  /* n_detections * 9, FP64 before the float cast */
} or_view;

typedef struct or_frame or_frame;

/* The discrete choices of one candidate's quad fit and decode, for tests
 * that compare builds of different precision. */
typedef struct or_fit_trace {
    int32_t fitted;          /* the hull gave a quad */
    int32_t passes;          /* refinement passes run */
    int32_t refined[2];      /* per pass: bit j set when corner j moved */
    int32_t have[2][4];      /* per pass and edge: a line was fitted */
    int32_t best[2][4][16];  /* per pass, edge and sample: probe kept, -1 none */
    int32_t decode_variant;  /* 0-3 refined quad at that rotation, 4-7 hull, -1 none */
} or_fit_trace;

/* Build precision: 32 or 64. */
int32_t or_real_bits(void);
/* 1 for the build that restates the reference's transcendental calls and
 * raw edge moments literally (OR_LITERAL), else 0. */
int32_t or_literal(void);

or_frame* or_create(const or_config* config);
void or_destroy(or_frame* frame);
/* Runs the whole pipeline; returns the detection count (>= 0) or -1. */
int32_t or_run(or_frame* frame, const uint8_t* gray, int32_t stride);
void or_get_view(const or_frame* frame, or_view* out);

/* Pieces exposed for focused tests. */
void or_fill_homography(const float corners[8], float out[9], double out64[9]);
int32_t or_hull_direction(int32_t dir, double* ux, double* uy);
/* Normal (-sin t, cos t) of the major axis t = atan2(2cxy, cxx - cyy) / 2. */
void or_axis_normal(double cxx, double cxy, double cyy, double* nx, double* ny);
int32_t or_match_code(int32_t family, uint64_t code, int32_t max_hamming, int32_t* hamming);
const uint64_t* or_family_codes(int32_t family, int32_t* count, int32_t* bits_per_side);
/* One edge-line fit (line = point x, y and unit normal), one decode
 * rotation, and the hull reduction, for closed-form tests. */
int32_t or_fit_edge(const or_frame* frame, const uint8_t* gray, int32_t stride, const double quad[8],
                    int32_t edge, int32_t first_pass, double line[4]);
int32_t or_decode_rotation(const or_frame* frame, const uint8_t* gray, int32_t stride, const double quad[8],
                           int32_t rot, int32_t* id, int32_t* hamming, double* margin);
int32_t or_hull_to_quad(const uint32_t* hull, int32_t count, double quad[8]);
/* The trace of candidate `candidate` (its index in the ordered list) of the
 * last or_run; 0 when out of range. */
int32_t or_get_fit_trace(const or_frame* frame, int32_t candidate, or_fit_trace* out);
/* Decode quads (one per rotation tried) that took the FP64 homography since
 * the last call with reset != 0 (process-wide count, for tests). */
int64_t or_decode_fp64_count(int32_t reset);

#ifdef __cplusplus
}
#endif

#endif /* BD_ORACLE_H */
