/*
 * bdcudatag C ABI -- MI355X (gfx950) implementation.
 *
 * Drop-in replacement for the reference C ABI
 * (detector/include/bdcudatag/bdcudatag.h:1-157): the same 14 symbols, the
 * same status codes and the same 40/84/40-byte structure layouts, so the Java
 * FFM wrapper (java-api/.../internal/NativeLibrary.java:104-132) and Python
 * ctypes callers (detector/bench/compare_detectors.py:133-156) bind it
 * unchanged.  A detector handles one tag family and one frame size; all
 * device memory is allocated by bd_detector_create and detect calls never
 * allocate.
 *
 * Version 3 adds batch entry points (many frames of uneven size per call) and
 * a per-stage introspection call used by the parity tests.  Every version 2
 * symbol and layout is unchanged.
 */
#ifndef BDCUDATAG_BDCUDATAG_H
#define BDCUDATAG_BDCUDATAG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BD_API __attribute__((visibility("default")))

#define BD_API_VERSION 3u
#define BD_MAX_WIDTH 8192
#define BD_MAX_HEIGHT 16384

typedef struct bd_detector bd_detector;

/* reference bdcudatag.h:34-37 */
typedef enum bd_family {
    BD_FAMILY_TAG36H11 = 0,
    BD_FAMILY_TAG16H5 = 1
} bd_family;

/* reference bdcudatag.h:39-46.  BD_ERROR_CUDA keeps its name and value; here
 * it reports a HIP runtime failure. */
typedef enum bd_status {
    BD_OK = 0,
    BD_ERROR_INVALID_ARGUMENT = -1,
    BD_ERROR_CUDA = -2,
    BD_ERROR_OUT_OF_MEMORY = -3,
    BD_ERROR_NOT_SUPPORTED = -4,
    BD_ERROR_UNKNOWN = -5
} bd_status;

/*
 * Detector configuration, reference bdcudatag.h:48-66.  Initialise with
 * bd_config_init().  max_hamming = -1 selects the family default (2 for
 * tag36h11, 1 for tag16h5, at most 3); max_candidates = 0 selects
 * clamp(next_pow2(w*h/4096)*3/2, 96, 1024); max_detections is clamped to
 * [1, 256].  use_cuda_graph keeps its name: it selects hipGraph replay.
 */
typedef struct bd_config {
    uint32_t struct_size;
    int32_t max_hamming;
    int32_t min_cluster_pixels;
    int32_t min_tag_pixels;
    int32_t max_detections;
    int32_t max_candidates;
    int32_t refine_edges;
    int32_t use_cuda_graph;
    float min_decision_margin;
    float decode_sharpening;
} bd_config;

/*
 * One detected tag, reference bdcudatag.h:68-89.  corners are TL, TR, BR, BL
 * in the tag's canonical orientation, in pixel-index coordinates.
 * homography maps the tag square (x, y) in [-1, 1]^2 (x right, y down, first
 * corner at (-1, -1)) to pixels, row-major, normalised so homography[8] == 1:
 *     s * (px, py, 1)^T = homography * (x, y, 1)^T
 */
typedef struct bd_detection {
    int32_t id;
    int32_t hamming;
    float decision_margin;
    float corners[8];
    int32_t reserved;
    float homography[9];
} bd_detection;

/* reference bdcudatag.h:91-101.  last_gpu_ms / average_gpu_ms are device
 * times from HIP events around the frame (EMA 0.9 / 0.1). */
typedef struct bd_stats {
    uint64_t frames;
    float last_gpu_ms;
    float average_gpu_ms;
    float last_total_ms;
    int32_t runs;
    int32_t candidates;
    int32_t detections;
    int32_t quads;
    int32_t roi_rows;
} bd_stats;

/* reference bdcudatag.h:103-112 */
BD_API uint32_t bd_api_version(void);
BD_API const char* bd_version_string(void);
BD_API int32_t bd_family_count(void);
BD_API const char* bd_family_name(int32_t family);
BD_API int32_t bd_family_bits_per_side(int32_t family);
BD_API int32_t bd_family_min_hamming(int32_t family);
BD_API int32_t bd_family_default_max_hamming(int32_t family);

/* reference bdcudatag.h:112, defaults of detector/src/detector.cpp:97-110 */
BD_API void bd_config_init(bd_config* config);

/*
 * reference bdcudatag.h:114-121.  Returns NULL on failure and writes a
 * bd_status to *status_out: BD_ERROR_INVALID_ARGUMENT for a bad family,
 * dimensions outside 16..8192 x 16..16384 or struct_size in 1..39,
 * BD_ERROR_CUDA when the HIP runtime fails.  bd_last_error() has the message.
 */
BD_API bd_detector* bd_detector_create(int32_t family, int32_t width, int32_t height,
                                       const bd_config* config, int32_t* status_out);
/* reference bdcudatag.h:122 */
BD_API void bd_detector_destroy(bd_detector* detector);

/*
 * reference bdcudatag.h:124-131.  8-bit grayscale frame in host memory with
 * row stride >= width.  Returns min(found, max_detections) and writes
 * min(that, out_capacity) records, or a negative bd_status.
 */
BD_API int32_t bd_detector_detect_host(bd_detector* detector, const uint8_t* gray,
                                       int32_t stride, bd_detection* out,
                                       int32_t out_capacity);

/*
 * reference bdcudatag.h:133-139: same, reading a HIP device pointer in place.
 * A device frame must span less than 4 GiB (stride * (height - 1) + width);
 * larger spans return BD_ERROR_INVALID_ARGUMENT.
 */
BD_API int32_t bd_detector_detect_device(bd_detector* detector, const uint8_t* gray_device,
                                         int32_t stride, bd_detection* out,
                                         int32_t out_capacity);

/* reference bdcudatag.h:141 */
BD_API void bd_detector_get_stats(bd_detector* detector, bd_stats* out);

/* reference bdcudatag.h:143-144: thread-local message of the last failure. */
BD_API const char* bd_last_error(void);

/* ---------------------------------------------------------------------- */
/* Version 3 additions (no counterpart in the reference ABI).              */
/* ---------------------------------------------------------------------- */

/* One frame of a batch: width and height must equal the size the batch
 * was created with for that position; stride >= width. */
typedef struct bd_frame {
    const uint8_t* gray;
    int32_t width;
    int32_t height;
    int32_t stride;
    int32_t reserved;
} bd_frame;

typedef struct bd_batch bd_batch;

/*
 * A batch detector for `count` frames whose sizes are fixed at creation
 * (they may differ from each other).  Every frame runs the same pipeline as
 * bd_detector_detect_host, all frames in one hipGraph launch.
 */
BD_API bd_batch* bd_batch_create(int32_t family, int32_t count, const int32_t* widths,
                                 const int32_t* heights, const bd_config* config,
                                 int32_t* status_out);
BD_API void bd_batch_destroy(bd_batch* batch);

/*
 * Detects tags in `count` host frames (count equals the creation count).
 * Frame i's results go to out + i * out_stride (at most out_stride records)
 * and its per-frame return value (as bd_detector_detect_host) to counts[i].
 * Returns the total number of records written or a negative bd_status.
 */
BD_API int32_t bd_batch_detect_host(bd_batch* batch, const bd_frame* frames, int32_t count,
                                    bd_detection* out, int32_t out_stride, int32_t* counts);
/* Same with device-resident frames (bd_frame.gray are HIP device pointers). */
BD_API int32_t bd_batch_detect_device(bd_batch* batch, const bd_frame* frames, int32_t count,
                                      bd_detection* out, int32_t out_stride, int32_t* counts);
/* Stats of frame `index` of the last batch call (frames counts batch calls). */
BD_API void bd_batch_get_stats(bd_batch* batch, int32_t index, bd_stats* out);

/*
 * Introspection for parity tests: copies one stage buffer of frame `index`
 * of the last completed call into `dst` (at most `capacity` bytes) and
 * returns the buffer's full size in bytes, or a negative bd_status.
 * Stages are the BD_STAGE_* values below.  For a single-frame detector pass
 * the handle through bd_detector_batch().
 */
enum {
    BD_STAGE_TILE_MIN = 0,     /* u8[tile_h * tile_w] */
    BD_STAGE_TILE_MAX = 1,     /* u8[tile_h * tile_w] */
    BD_STAGE_TILE_MEAN = 2,    /* u8[tile_h * tile_w] */
    BD_STAGE_HIST = 3,         /* u32[256 span + 64 mean] of the last frame */
    BD_STAGE_SCALARS = 4,      /* u32[16], see bd_internal.h */
    BD_STAGE_ROW_ROI = 5,      /* i32[4 * tile_h]: active, start, end, band joined locally */
    BD_STAGE_BITS = 6,         /* u64[height * ceil(width/64)], bit x&63, 1 = white
                                  (written only under stage capture) */
    BD_STAGE_ROW_COUNT = 7,    /* u32[height]: black runs per row (maximal; 32-px fragments in row 0) */
    BD_STAGE_RUNS = 8,         /* u32[height * row_cap]: x0 | x1 << 16, bit 31 = band root
                                  (rows inside a band only under stage capture) */
    BD_STAGE_LABELS = 9,       /* u32[height * row_cap]: a run's band root slot; a band root's
                                  is its component root (so two hops name the root) */
    BD_STAGE_AREA = 10,        /* u32[height * row_cap], valid at roots */
    BD_STAGE_BBOX = 11,        /* u32[height * row_cap * 4]: x0, y0, x1, y1 at roots */
    BD_STAGE_CANDIDATES = 12,  /* u32[max_candidates]: root slots, largest first */
    BD_STAGE_QUADS = 13,       /* f32[max_candidates * 8] */
    BD_STAGE_HULL_QUADS = 14,  /* f32[max_candidates * 8] */
    BD_STAGE_DECODE = 15,      /* bd_detection[max_candidates] per candidate, id -1 = none */
    BD_STAGE_TILE_THRESH = 16  /* u8[tile_h * tile_w], pooled thresholds (written only
                                  under stage capture) */
};
BD_API int64_t bd_batch_debug_copy(bd_batch* batch, int32_t index, int32_t stage, void* dst,
                                   int64_t capacity);
/* Stage capture (no counterpart in the reference; for stage-by-stage
 * comparisons): with on != 0, later calls also write the binary image, the
 * pooled thresholds and the runs and labels of the rows inside each 8-row
 * band, which no later kernel reads (the pipeline keeps each band's first
 * and last rows and its band roots).  Detections are the same either way.
 * Waits for the batch's streams. */
BD_API int32_t bd_batch_set_capture(bd_batch* batch, int32_t on);
/* Geometry of frame `index`: out[0..7] = width, height, tile_w, tile_h,
 * words64, row_cap, max_candidates, max_detections. */
BD_API int32_t bd_batch_frame_info(bd_batch* batch, int32_t index, int32_t out[8]);
/* The batch engine behind a single-frame detector (owned by the detector). */
BD_API bd_batch* bd_detector_batch(bd_detector* detector);

/*
 * Benchmark hooks on the frames staged by the last host call.
 * bd_batch_time_stages: runs the kernels `reps` times with HIP events between
 * consecutive kernels on the detector's stream; ms[0..8] = mean time per
 * kernel (tile_stats, binarize_runs, merge, stats, collect, select, fit,
 * decode, finalize), ms[9] = mean whole pipeline.
 * bd_batch_replay: replays the captured graph `reps` times back to back
 * without host synchronisation and returns the total device time in *ms;
 * device_path = 0 replays the host path (H2D of the staged frames, kernels,
 * D2H), 1 the device path on the frames of the last device call.
 */
BD_API int32_t bd_batch_time_stages(bd_batch* batch, int32_t reps, float* ms);
BD_API int32_t bd_batch_replay(bd_batch* batch, int32_t reps, int32_t device_path, float* ms);

/*
 * Launch parameters of a batch (no counterpart in the reference).  They
 * change speed only: every setting gives bit-identical detections.
 */
typedef struct bd_tune {
    int32_t fit_grid;    /* persistent quad-fit workgroups, 64..16384 (default: as many per CU as one CU holds) */
    int32_t decode_grid; /* persistent decode workgroups, 64..16384 (default: as many as the
                            device's CUs hold at once) */
    int32_t head_ranks;  /* each frame's largest candidates scheduled first, 0..4 (default 2) */
    int32_t streams;     /* 2: a batch of two or more frames runs as two concurrent parts (the
                            first about 4 Mpx larger, at most 1/16 of the batch, as its graph
                            starts first), on two streams; 1: as one launch sequence
                            (default 2; 0 in bd_batch_set_tune selects the default; a one-frame
                            batch reports 1) */
    int32_t decode_waves; /* waves per decode workgroup: 4 (two variants per wave), 8 (one variant
                             per wave) or 16 (two candidates per workgroup); default 8, 0 in
                             bd_batch_set_tune selects the default */
    int32_t chunk_px;     /* pixels a binarization thread packs per item: 64 or 128 (default 64;
                             0 in bd_batch_set_tune selects the default) */
    int32_t head_area;    /* candidates of at least this many pixels are also scheduled first
                             (at most 32 per frame), >= 1; INT32_MAX turns this off (default
                             1000; 0 in bd_batch_set_tune selects the default) */
    int32_t band_rows;    /* pixel rows per binarization workgroup: 8 (one band) or 16 (two bands
                             in turn), an execution tiling only (default 8; 0 in bd_batch_set_tune
                             selects the default) */
} bd_tune;

BD_API int32_t bd_batch_get_tune(bd_batch* batch, bd_tune* out);
/* Waits for the batch's stream and re-captures its graphs. */
BD_API int32_t bd_batch_set_tune(bd_batch* batch, const bd_tune* tune);
/*
 * Times `reps` graph replays of the frames of the last detect call on this
 * batch (host frames stay staged; device frames must still be valid) for a
 * small set of launch parameters, one parameter at a time keeping the best
 * so far, and leaves the fastest set in place.  Writes it to *best and the
 * mean pipeline time of that set to *best_ms.
 */
BD_API int32_t bd_batch_autotune(bd_batch* batch, int32_t reps, bd_tune* best, float* best_ms);

/* ---------------------------------------------------------------------- */
/* Serving: continuous batching over a preallocated arena (version 3).    */
/* ---------------------------------------------------------------------- */

typedef struct bd_server bd_server;

/* One frame-size class of a server: `lanes` engines (one HIP stream each)
 * of `lane_frames` positions; at most `queue_depth` admitted requests wait
 * for a free lane, later submissions are dropped. */
typedef struct bd_server_class {
    int32_t width;
    int32_t height;
    int32_t lanes;
    int32_t lane_frames;
    int32_t queue_depth;
    int32_t reserved;
} bd_server_class;

typedef struct bd_server_result {
    int64_t request;      /* id returned by bd_server_submit */
    int32_t class_index;
    int32_t count;        /* detections found (as bd_detector_detect_device returns) */
    int32_t written;      /* records copied to the detection buffer by bd_server_poll */
    int32_t batch_frames; /* frames in the launch that carried the request */
    double latency_ms;    /* admission -> results on the host */
} bd_server_result;

typedef struct bd_stream_spec {
    int32_t class_index;
    int32_t frame_count;          /* device frames cycled by the stream */
    double fps;                   /* Poisson arrival rate */
    const uint8_t* const* frames; /* HIP device pointers */
    int32_t stride;
    int32_t reserved;
} bd_stream_spec;

typedef struct bd_serve_report {
    int64_t offered;
    int64_t completed;
    int64_t dropped;
    int64_t launches;
    double seconds;
    double fps;
    double mpix_s;
    double p50_ms;
    double p99_ms;
    double p999_ms;
    double max_ms;
    double mean_ms;
    double mean_batch;
} bd_serve_report;

BD_API bd_server* bd_server_create(int32_t family, int32_t nclasses,
                                   const bd_server_class* classes, const bd_config* config,
                                   int32_t* status_out);
/* The same server with admission on the device: each class's requests are
 * published into a ring of queue_depth slots in pinned host memory (0: 256;
 * beyond it they are dropped), and every lane launch starts with a kernel
 * that waits on the device, at most admit_wait_us per launch or until the
 * server stops, for published requests and claims up to lane_frames of
 * them; the launch's last kernel publishes each request's records, count
 * and completion flag into the ring.  The host only re-launches each lane's
 * captured graph and timestamps completions.  At most 16 lanes in all (keep
 * them within the process's hardware queues, 4 by default, which the streams
 * of every other detector and batch the process holds share: a lane whose
 * stream shares a queue waits behind another stream's work, another lane's
 * admission wait included).  submit/poll/wait_idle/simulate and destroy work
 * as for bd_server_create. */
BD_API bd_server* bd_server_create_resident(int32_t family, int32_t nclasses,
                                            const bd_server_class* classes,
                                            const bd_config* config, double admit_wait_us,
                                            int32_t* status_out);
BD_API void bd_server_destroy(bd_server* server);
/* Non-blocking admission of a device-resident frame; returns the request id
 * (>= 0) or BD_ERROR_OUT_OF_MEMORY when the class queue is full (dropped). */
BD_API int64_t bd_server_submit(bd_server* server, int32_t class_index,
                                const uint8_t* gray_device, int32_t stride);
/* Non-blocking: moves up to max_results completed requests out; request i's
 * detections go to dets + i * dets_per_result. */
BD_API int32_t bd_server_poll(bd_server* server, bd_server_result* results, int32_t max_results,
                              bd_detection* dets, int32_t dets_per_result);
/* Waits until every admitted request has completed (BD_OK) or timeout. */
BD_API int32_t bd_server_wait_idle(bd_server* server, int32_t timeout_ms);
/* Drives the server with Poisson arrivals from `nstreams` simulated
 * cameras for `seconds`; latency is measured from each scheduled arrival
 * to its results on the host. */
BD_API int32_t bd_server_simulate(bd_server* server, const bd_stream_spec* streams,
                                  int32_t nstreams, double seconds, uint64_t seed,
                                  bd_serve_report* report);
BD_API const char* bd_server_last_error(void);

/*
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, serving loops that must not call
 * hipMalloc, harnesses that forbid it.  One width x height frame at
 * gray_device (a HIP device pointer) is detected with the kernels of
 * bd_detector_detect_device, queued on `stream` (a hipStream_t; NULL is the
 * null stream), and the call returns at once: BD_OK when queued, else a
 * negative bd_status with bd_last_error() set.  When the stream reaches it,
 * *count_device (device memory) holds min(found, max_detections) and
 * out_device (device memory) its first min(count, out_capacity) records,
 * homographies included.  Nothing is allocated: the working memory is one
 * static 64 MiB device array per GPU, held until the process ends, which
 * fits frames up to 1920x1200; larger frames are rejected with
 * BD_ERROR_INVALID_ARGUMENT before any GPU work.  Calls that overlap on one
 * GPU share that array and race: serialise them (one stream per GPU).
 */
BD_API int32_t bd_detect_device_noalloc(int32_t family, int32_t width, int32_t height,
                                        const bd_config* config, const uint8_t* gray_device,
                                        int32_t stride, bd_detection* out_device,
                                        int32_t out_capacity, int32_t* count_device, void* stream);

#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
static_assert(sizeof(bd_config) == 40, "bd_config layout changed");
static_assert(sizeof(bd_detection) == 84, "bd_detection layout changed");
static_assert(sizeof(bd_stats) == 40, "bd_stats layout changed");
#endif

#endif /* BDCUDATAG_BDCUDATAG_H */
