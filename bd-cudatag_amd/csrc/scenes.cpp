// Synthetic frame generator: deterministic scenes with exact ground truth.
//
// Host-only C++ used by the tests and the benchmark to make inputs; it is not
// part of the detector library.  Two families of scenes:
//
//  * bds_bench_scene: the reference's own benchmark scenes
//    (detector/bench/scenes.hpp:153-206, seed 6324): eight tag36h11 tags on a
//    jittered 4x2 grid, 2x2 supersampling, uniform noise of -3..4 DN, decoys
//    on scenes with index % 8 == 3.  It reproduces that generator's pixels so
//    a frame is identical on both sides.
//  * bds_scene: the benchmark workloads of this project, each a seeded frame
//    with a textured background, tags rendered through a real 3D pose
//    (perspective tilt), optional decoys and clutter, blur and Gaussian noise.
//
// Every tag record carries its pose (tag square [-1,1]^2 -> continuous pixel
// coordinates) and its ground-truth corners in pixel-index coordinates
// (continuous minus 0.5, as detector/bench/scenes.hpp:40-58 defines them).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "tag_codes.inc"

namespace {

struct Family {
    const uint64_t* codes;
    int count;
    int bits;
};

Family family_of(int f) {
    if (f == 1) return {bd_codes_tag16h5, 30, 4};
    return {bd_codes_tag36h11, 587, 6};
}

struct Mat3 {
    double v[9];
    void project(double x, double y, double* ox, double* oy) const {
        const double w = v[6] * x + v[7] * y + v[8];
        *ox = (v[0] * x + v[1] * y + v[2]) / w;
        *oy = (v[3] * x + v[4] * y + v[5]) / w;
    }
    Mat3 inverse() const {
        const double a = v[0], b = v[1], c = v[2], d = v[3], e = v[4], f = v[5], g = v[6],
                     h = v[7], i = v[8];
        const double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
        const double k = 1.0 / det;
        Mat3 o{};
        o.v[0] = (e * i - f * h) * k;
        o.v[1] = (c * h - b * i) * k;
        o.v[2] = (b * f - c * e) * k;
        o.v[3] = (f * g - d * i) * k;
        o.v[4] = (a * i - c * g) * k;
        o.v[5] = (c * d - a * f) * k;
        o.v[6] = (d * h - e * g) * k;
        o.v[7] = (b * g - a * h) * k;
        o.v[8] = (a * e - b * d) * k;
        return o;
    }
};

// Pose of the reference test fixture (detector/tests/test_support.hpp:46-65):
// the tag square [-1,1]^2 centred at (cx, cy), side size_px, rotated by angle,
// with an optional projective term.
Mat3 make_pose(double cx, double cy, double size_px, double angle, double perspective) {
    const double c = std::cos(angle), s = std::sin(angle), h = 0.5 * size_px;
    Mat3 m{};
    m.v[0] = c * h;
    m.v[1] = -s * h;
    m.v[2] = cx;
    m.v[3] = s * h;
    m.v[4] = c * h;
    m.v[5] = cy;
    m.v[6] = perspective * c;
    m.v[7] = -perspective * s;
    m.v[8] = 1.0;
    return m;
}

int popcount64(uint64_t v) { return __builtin_popcountll(v); }

// Rotates a row-major bits x bits code by 90 degrees
// (detector/tests/test_support.hpp:76-88).
uint64_t rotate_code(uint64_t code, int bits) {
    const int n = bits * bits;
    uint64_t out = 0;
    for (int r = 0; r < bits; ++r) {
        for (int c = 0; c < bits; ++c) {
            const uint64_t bit = (code >> (n - 1 - (r * bits + c))) & 1ull;
            const int nr = c, nc = bits - 1 - r;
            out |= bit << (n - 1 - (nr * bits + nc));
        }
    }
    return out;
}

// Cell colour of tag coordinates (tx, ty) in [-1,1]^2: 0 = black, 1 = white.
inline int tag_cell(const Family& fam, uint64_t code, double tx, double ty, bool flip_one_bit) {
    const int dim = fam.bits + 2;
    const int cx = std::min(dim - 1, std::max(0, static_cast<int>((tx + 1.0) * 0.5 * dim)));
    const int cy = std::min(dim - 1, std::max(0, static_cast<int>((ty + 1.0) * 0.5 * dim)));
    if (cx == 0 || cy == 0 || cx == dim - 1 || cy == dim - 1) return 0;
    const int pos = (cy - 1) * fam.bits + (cx - 1);
    uint64_t cell = (code >> (fam.bits * fam.bits - 1 - pos)) & 1ull;
    if (flip_one_bit && pos == 0) cell ^= 1ull;
    return cell ? 1 : 0;
}

// detector/tests/test_support.hpp:93-137 (render_tag): white image, the tag
// square drawn through the inverse pose with ss x ss supersampling.
void render_tag_ref(const Family& fam, uint64_t code, int width, int height, const Mat3& pose,
                    int supersample, bool flip_one_bit, uint8_t* image) {
    const Mat3 inv = pose.inverse();
    const int ss = std::max(1, supersample);
    for (int py = 0; py < height; ++py) {
        for (int px = 0; px < width; ++px) {
            int accum = 0;
            for (int sy = 0; sy < ss; ++sy) {
                for (int sx = 0; sx < ss; ++sx) {
                    const double ix = px + (sx + 0.5) / ss;
                    const double iy = py + (sy + 0.5) / ss;
                    double tx, ty;
                    inv.project(ix, iy, &tx, &ty);
                    if (tx < -1.0 || tx > 1.0 || ty < -1.0 || ty > 1.0) {
                        accum += 255;
                        continue;
                    }
                    accum += tag_cell(fam, code, tx, ty, flip_one_bit) ? 255 : 0;
                }
            }
            image[static_cast<size_t>(py) * width + px] = static_cast<uint8_t>(accum / (ss * ss));
        }
    }
}

void square_corners(const Mat3& pose, double out[8]) {
    static const double kUnit[4][2] = {{-1.0, -1.0}, {1.0, -1.0}, {1.0, 1.0}, {-1.0, 1.0}};
    for (int i = 0; i < 4; ++i) {
        pose.project(kUnit[i][0], kUnit[i][1], &out[2 * i], &out[2 * i + 1]);
        out[2 * i] -= 0.5;
        out[2 * i + 1] -= 0.5;
    }
}

// detector/bench/scenes.hpp:64-78 (mix32, next_unit).
uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
double next_unit(uint32_t& state) {
    state = mix32(state + 0x9e3779b9u);
    return (state >> 8) * (1.0 / 16777216.0);
}

// detector/bench/scenes.hpp:92-126 (render_tag_into): renders a patch around
// the tag and composites it with min().
void render_tag_into_ref(uint8_t* image, int width, int height, const Family& fam, uint64_t code,
                         const Mat3& pose) {
    double c[8];
    square_corners(pose, c);
    double min_x = c[0], max_x = c[0], min_y = c[1], max_y = c[1];
    for (int i = 1; i < 4; ++i) {
        min_x = std::min(min_x, c[2 * i]);
        max_x = std::max(max_x, c[2 * i]);
        min_y = std::min(min_y, c[2 * i + 1]);
        max_y = std::max(max_y, c[2 * i + 1]);
    }
    constexpr int kMargin = 4;
    const int x0 = std::max(0, static_cast<int>(std::floor(min_x)) - kMargin);
    const int y0 = std::max(0, static_cast<int>(std::floor(min_y)) - kMargin);
    const int x1 = std::min(width - 1, static_cast<int>(std::ceil(max_x)) + kMargin);
    const int y1 = std::min(height - 1, static_cast<int>(std::ceil(max_y)) + kMargin);
    if (x1 < x0 || y1 < y0) return;
    const int rw = x1 - x0 + 1, rh = y1 - y0 + 1;
    Mat3 shifted = pose;
    shifted.v[2] -= x0;
    shifted.v[5] -= y0;
    std::vector<uint8_t> patch(static_cast<size_t>(rw) * rh);
    render_tag_ref(fam, code, rw, rh, shifted, 2, false, patch.data());
    for (int y = 0; y < rh; ++y) {
        uint8_t* dst = image + static_cast<size_t>(y0 + y) * width + x0;
        const uint8_t* src = patch.data() + static_cast<size_t>(y) * rw;
        for (int x = 0; x < rw; ++x) dst[x] = std::min(dst[x], src[x]);
    }
}

// detector/bench/scenes.hpp:131-148 (random_non_tag_code).  The reference
// draws both halves from mix32 of the same, never-advanced state, so its
// retry loop either returns the first draw or spins without side effects;
// compiled with optimisation it returns the first draw (the reference's own
// exporter, built from its sources, does).  The port returns that draw.
uint64_t random_non_tag_code_ref(uint32_t& rng, const Family& fam) {
    const int cells = fam.bits * fam.bits;
    const uint64_t mask = cells >= 64 ? ~0ull : ((1ull << cells) - 1ull);
    return ((static_cast<uint64_t>(mix32(rng)) << 32) ^ mix32(rng)) & mask;
}

// ---------------------------------------------------------------------------
// Project workloads
// ---------------------------------------------------------------------------

uint64_t splitmix64(uint64_t& x) {
    x += 0x9e3779b97f4a7c15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s;
    Rng(uint64_t config, uint64_t frame, uint64_t stream) {
        s = (config << 56) ^ (stream << 48) ^ frame ^ 0x5851f42d4c957f2dull;
        splitmix64(s);
    }
    double unit() { return (splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double a, double b) { return a + (b - a) * unit(); }
    int range(int a, int b) { return a + static_cast<int>(unit() * (b - a + 1)); }  // [a, b]
    uint64_t bits() { return splitmix64(s); }
};

struct WorkTag {
    int family;
    int id;
    int decoy;
    uint64_t code;
    Mat3 pose;
    double cx, cy, radius;  // bounding circle of the quiet-zone square
    int black, white;
};

// Tag square -> image homography through a pinhole camera: the tag plane is
// rolled by `roll`, tilted by `tilt` about the in-plane axis at angle `axis`,
// and placed so its centre projects to (u0, v0) with apparent side `side`.
Mat3 camera_pose(double u0, double v0, double side, double roll, double tilt, double axis,
                 double focal) {
    const double ca = std::cos(axis), sa = std::sin(axis);
    const double ct = std::cos(tilt), st = std::sin(tilt);
    // Rotation about the unit axis (ca, sa, 0) by tilt (Rodrigues).
    double T[9] = {ct + ca * ca * (1 - ct), ca * sa * (1 - ct), sa * st,
                   sa * ca * (1 - ct),      ct + sa * sa * (1 - ct), -ca * st,
                   -sa * st,                ca * st,                 ct};
    const double cr = std::cos(roll), sr = std::sin(roll);
    const double Z[9] = {cr, -sr, 0, sr, cr, 0, 0, 0, 1};
    double R[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += T[i * 3 + k] * Z[k * 3 + j];
            R[i * 3 + j] = s;
        }
    // Camera centred on the tag: principal point at (u0, v0), depth = focal
    // so one world unit is one pixel at the tag centre.
    const double h = 0.5 * side;
    Mat3 m{};
    // columns: K * [r1*h, r2*h, t], K = [[f,0,u0],[0,f,v0],[0,0,1]], t = (0,0,f)
    for (int row = 0; row < 3; ++row) {
        const double r1 = R[row * 3 + 0] * h, r2 = R[row * 3 + 1] * h;
        const double tz = row == 2 ? focal : 0.0;
        if (row < 2) {
            const double pp = row == 0 ? u0 : v0;
            m.v[row * 3 + 0] = focal * r1 + pp * R[6] * h;
            m.v[row * 3 + 1] = focal * r2 + pp * R[7] * h;
            m.v[row * 3 + 2] = pp * focal;
        } else {
            m.v[6] = r1;
            m.v[7] = r2;
            m.v[8] = tz;
        }
    }
    for (int i = 0; i < 9; ++i) m.v[i] /= focal;
    return m;
}

struct SceneSpec {
    int min_tags, max_tags;
    double min_side, max_side;
    double max_tilt_deg;
    int family_mode;  // 0 = tag36h11, 1 = tag16h5, 2 = mixed
    bool decoys;      // two decoys on seeds % 4 == 0
    int rectangles;   // black rectangles
    int specks;       // glyph specks of 4..400 px^2
    double blur_sigma;  // 0 = 3x3 box blur
    double noise_sigma;
    bool mid_background;  // background between the cell levels
    int supersample;
};

SceneSpec spec_for(int config) {
    SceneSpec s{};
    s.supersample = 4;
    s.max_tilt_deg = 40.0;
    s.noise_sigma = 3.0;
    switch (config) {
        case 1:  // dense single frames
            s.min_tags = s.max_tags = 64;
            s.min_side = 24;
            s.max_side = 300;
            s.family_mode = 2;
            s.rectangles = 200;
            s.specks = 400;
            break;
        case 2:  // batching: 0..32 tags
            s.min_tags = 0;
            s.max_tags = 32;
            s.min_side = 60;
            s.max_side = 200;
            s.decoys = true;
            break;
        case 3:  // precision: heavy degradation
            s.min_tags = s.max_tags = 16;
            s.min_side = 20;
            s.max_side = 80;
            s.max_tilt_deg = 60.0;
            s.blur_sigma = 1.5;
            s.noise_sigma = 8.0;
            s.mid_background = true;
            break;
        case 4:  // serving: 0..16 tags
            s.min_tags = 0;
            s.max_tags = 16;
            s.min_side = 60;
            s.max_side = 200;
            s.decoys = true;
            break;
        default:  // 0: eight tags
            s.min_tags = s.max_tags = 8;
            s.min_side = 60;
            s.max_side = 200;
            s.decoys = true;
            break;
    }
    return s;
}

bool non_tag_code(Rng& rng, const Family& fam, uint64_t* out) {
    const int cells = fam.bits * fam.bits;
    const uint64_t mask = (1ull << cells) - 1ull;
    for (int attempt = 0; attempt < 64; ++attempt) {
        const uint64_t code = rng.bits() & mask;
        int best = 1000;
        for (int i = 0; i < fam.count && best >= 8; ++i) {
            uint64_t r = fam.codes[i];
            for (int k = 0; k < 4; ++k) {
                best = std::min(best, popcount64(code ^ r));
                r = rotate_code(r, fam.bits);
            }
        }
        if (best >= 8) {
            *out = code;
            return true;
        }
    }
    return false;
}

// Renders a tag with its one-cell white quiet zone over the background by
// supersampling; subsamples outside the quiet zone keep the background.
void render_work_tag(float* img, int W, int H, const WorkTag& t, int ss) {
    const Family fam = family_of(t.family);
    const int dim = fam.bits + 2;
    const double q = 1.0 + 2.0 / dim;  // quiet zone edge in tag units
    // bounding box of the quiet-zone square
    double xs[4], ys[4];
    const double cq[4][2] = {{-q, -q}, {q, -q}, {q, q}, {-q, q}};
    for (int i = 0; i < 4; ++i) t.pose.project(cq[i][0], cq[i][1], &xs[i], &ys[i]);
    const int x0 = std::max(0, static_cast<int>(std::floor(*std::min_element(xs, xs + 4))) - 1);
    const int x1 = std::min(W - 1, static_cast<int>(std::ceil(*std::max_element(xs, xs + 4))) + 1);
    const int y0 = std::max(0, static_cast<int>(std::floor(*std::min_element(ys, ys + 4))) - 1);
    const int y1 = std::min(H - 1, static_cast<int>(std::ceil(*std::max_element(ys, ys + 4))) + 1);
    const Mat3 inv = t.pose.inverse();
    const double inv_n = 1.0 / (ss * ss);
    for (int py = y0; py <= y1; ++py) {
        for (int px = x0; px <= x1; ++px) {
            const float bg = img[static_cast<size_t>(py) * W + px];
            double acc = 0.0;
            bool touched = false;
            for (int sy = 0; sy < ss; ++sy) {
                for (int sx = 0; sx < ss; ++sx) {
                    double tx, ty;
                    inv.project(px + (sx + 0.5) / ss, py + (sy + 0.5) / ss, &tx, &ty);
                    if (tx < -q || tx > q || ty < -q || ty > q) {
                        acc += bg;
                        continue;
                    }
                    touched = true;
                    if (tx < -1.0 || tx > 1.0 || ty < -1.0 || ty > 1.0) {
                        acc += t.white;
                        continue;
                    }
                    acc += tag_cell(fam, t.code, tx, ty, false) ? t.white : t.black;
                }
            }
            if (touched) img[static_cast<size_t>(py) * W + px] = static_cast<float>(acc * inv_n);
        }
    }
}

void box_blur3(std::vector<float>& img, int W, int H) {
    std::vector<float> tmp(img.size());
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const int a = std::max(0, x - 1), b = std::min(W - 1, x + 1);
            const float* r = &img[static_cast<size_t>(y) * W];
            tmp[static_cast<size_t>(y) * W + x] = (r[a] + r[x] + r[b]) * (1.0f / 3.0f);
        }
    for (int y = 0; y < H; ++y) {
        const int a = std::max(0, y - 1), b = std::min(H - 1, y + 1);
        for (int x = 0; x < W; ++x)
            img[static_cast<size_t>(y) * W + x] =
                (tmp[static_cast<size_t>(a) * W + x] + tmp[static_cast<size_t>(y) * W + x] +
                 tmp[static_cast<size_t>(b) * W + x]) *
                (1.0f / 3.0f);
    }
}

void gaussian_blur(std::vector<float>& img, int W, int H, double sigma) {
    const int r = static_cast<int>(std::ceil(3.0 * sigma));
    std::vector<float> k(2 * r + 1);
    double s = 0;
    for (int i = -r; i <= r; ++i) s += (k[i + r] = static_cast<float>(std::exp(-0.5 * i * i / (sigma * sigma))));
    for (auto& v : k) v = static_cast<float>(v / s);
    std::vector<float> tmp(img.size());
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float acc = 0;
            for (int i = -r; i <= r; ++i) {
                const int xx = std::min(W - 1, std::max(0, x + i));
                acc += k[i + r] * img[static_cast<size_t>(y) * W + xx];
            }
            tmp[static_cast<size_t>(y) * W + x] = acc;
        }
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float acc = 0;
            for (int i = -r; i <= r; ++i) {
                const int yy = std::min(H - 1, std::max(0, y + i));
                acc += k[i + r] * tmp[static_cast<size_t>(yy) * W + x];
            }
            img[static_cast<size_t>(y) * W + x] = acc;
        }
}

}  // namespace

extern "C" {

typedef struct bds_tag {
    int32_t family;
    int32_t id;      // -1 for decoys
    int32_t decoy;
    int32_t reserved;
    double pose[9];     // tag square [-1,1]^2 -> continuous pixel coordinates
    double corners[8];  // black-square corners TL, TR, BR, BL, pixel-index coordinates
} bds_tag;

__attribute__((visibility("default"))) void bds_make_pose(double cx, double cy, double size,
                                                         double angle, double perspective,
                                                         double out[9]) {
    const Mat3 m = make_pose(cx, cy, size, angle, perspective);
    std::memcpy(out, m.v, sizeof m.v);
}

// detector/tests/test_support.hpp:93-137 on a white frame.
__attribute__((visibility("default"))) int32_t bds_render_tag(
    int32_t family, int32_t id, int32_t width, int32_t height, const double pose[9],
    int32_t supersample, int32_t use_code, uint64_t code, int32_t flip_one_bit, uint8_t* out) {
    const Family fam = family_of(family);
    if (id < 0 || id >= fam.count || width <= 0 || height <= 0 || !out) return -1;
    Mat3 m{};
    std::memcpy(m.v, pose, sizeof m.v);
    render_tag_ref(fam, use_code ? code : fam.codes[id], width, height, m, supersample,
                   flip_one_bit != 0, out);
    return 0;
}

__attribute__((visibility("default"))) uint64_t bds_rotate_code(uint64_t code, int32_t bits) {
    return rotate_code(code, bits);
}

// One scene of detector/bench/scenes.hpp:153-206 (make_scenes), by index.
// Writes width*height pixels and the 8 ground-truth tags; returns the number
// of tags.
__attribute__((visibility("default"))) int32_t bds_bench_scene(int32_t width, int32_t height,
                                                              int32_t s, uint32_t seed,
                                                              uint8_t* image, bds_tag* tags) {
    static constexpr int kIds[8] = {0, 3, 7, 11, 17, 42, 101, 586};
    static constexpr double kScale[8] = {1.00, 0.85, 0.70, 1.12, 0.95, 0.60, 1.05, 0.30};
    static constexpr double kPi = 3.14159265358979323846;
    const Family fam = family_of(0);
    const double base_size = width / 9.0;
    uint32_t rng = seed + 0x9e3779b9u * static_cast<uint32_t>(s + 1);
    std::fill(image, image + static_cast<size_t>(width) * height, 255);
    for (int t = 0; t < 8; ++t) {
        const int col = t % 4, row = t / 4;
        const double cx = (col + 0.5) / 4.0 * width + (next_unit(rng) - 0.5) * 0.05 * width;
        const double cy = (row + 0.5) / 2.0 * height + (next_unit(rng) - 0.5) * 0.06 * height;
        const double size = base_size * kScale[t] * (0.92 + 0.16 * next_unit(rng));
        const double angle = (next_unit(rng) - 0.5) * (0.8 * kPi / 2.0);
        const double perspective = (next_unit(rng) - 0.5) * 1.2e-3;
        const Mat3 pose = make_pose(cx, cy, size, angle, perspective);
        if (tags) {
            tags[t].family = 0;
            tags[t].id = kIds[t];
            tags[t].decoy = 0;
            tags[t].reserved = 0;
            std::memcpy(tags[t].pose, pose.v, sizeof pose.v);
            square_corners(pose, tags[t].corners);
        }
        render_tag_into_ref(image, width, height, fam, fam.codes[kIds[t]], pose);
    }
    if (s % 8 == 3) {
        const double dx[2] = {width * 0.30, width * 0.70};
        const double dy[2] = {height * 0.50, height * 0.50};
        const double ds[2] = {base_size * 0.50, base_size * 0.40};
        for (int k = 0; k < 2; ++k) {
            const Mat3 pose = make_pose(dx[k], dy[k], ds[k], next_unit(rng) * 0.6, 0.0);
            render_tag_into_ref(image, width, height, fam, random_non_tag_code_ref(rng, fam), pose);
        }
    }
    // detector/bench/scenes.hpp:80-89 (add_noise): delta = ((state >> 28) & 7) - 3.
    uint32_t state = (seed + static_cast<uint32_t>(s)) | 1u;
    const size_t n = static_cast<size_t>(width) * height;
    for (size_t i = 0; i < n; ++i) {
        state = mix32(state);
        const int v = static_cast<int>(image[i]) + static_cast<int>((state >> 28) & 7u) - 3;
        image[i] = static_cast<uint8_t>(v < 0 ? 0 : (v > 255 ? 255 : v));
    }
    return 8;
}

// One frame of workload `config` (0 eight tags; 1 dense mixed families with
// clutter; 2 batching, 0..32 tags; 3 heavy degradation; 4 serving, 0..16 tags)
// at width x height from `seed`.  Writes up to max_tags records (decoys
// included, id -1) and returns the number written.
__attribute__((visibility("default"))) int32_t bds_scene(int32_t config, int32_t width,
                                                        int32_t height, uint64_t seed,
                                                        uint8_t* image, bds_tag* tags,
                                                        int32_t max_tags) {
    if (width < 16 || height < 16 || !image) return -1;
    const SceneSpec spec = spec_for(config);
    Rng rng(static_cast<uint64_t>(config), seed, 0);
    const int W = width, H = height;
    std::vector<float> img(static_cast<size_t>(W) * H);

    // Background: uniform grey with a low-frequency linear gradient of +-20.
    int black = 0, white = 255;
    double g0 = rng.uniform(90.0, 200.0);
    if (spec.mid_background) {
        black = rng.range(15, 60);
        white = rng.range(190, 240);
        g0 = black + (white - black) * rng.uniform(0.35, 0.65);
    }
    const double phi = rng.uniform(0.0, 2.0 * M_PI);
    const double half_diag = 0.5 * std::hypot(W, H);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const double u = ((x - 0.5 * W) * std::cos(phi) + (y - 0.5 * H) * std::sin(phi)) / half_diag;
            img[static_cast<size_t>(y) * W + x] = static_cast<float>(g0 + 20.0 * u);
        }

    // Tags and decoys, placed without overlap of their quiet-zone circles.
    std::vector<WorkTag> placed;
    const int ntags = spec.min_tags == spec.max_tags ? spec.min_tags : rng.range(spec.min_tags, spec.max_tags);
    const int ndecoys = (spec.decoys && (seed % 4) == 0) ? 2 : 0;
    const double focal = 1.2 * std::max(W, H);
    for (int t = 0; t < ntags + ndecoys; ++t) {
        WorkTag wt{};
        wt.decoy = t >= ntags;
        wt.family = spec.family_mode == 2 ? rng.range(0, 1) : spec.family_mode;
        const Family fam = family_of(wt.family);
        if (wt.decoy) {
            wt.id = -1;
            if (!non_tag_code(rng, fam, &wt.code)) continue;
        } else {
            wt.id = rng.range(0, fam.count - 1);
            wt.code = fam.codes[wt.id];
        }
        wt.black = black;
        wt.white = white;
        double side = rng.uniform(spec.min_side, spec.max_side);
        bool ok = false;
        for (int attempt = 0; attempt < 400 && !ok; ++attempt) {
            if (attempt > 0 && attempt % 100 == 0) side = std::max(spec.min_side, side * 0.8);
            const double roll = rng.uniform(-M_PI, M_PI);
            const double tilt = rng.uniform(0.0, spec.max_tilt_deg) * M_PI / 180.0;
            const double axis = rng.uniform(0.0, 2.0 * M_PI);
            const double dim = fam.bits + 2;
            const double reach = 0.5 * side * (1.0 + 2.0 / dim) * 1.5 + 4.0;
            if (2 * reach >= std::min(W, H)) continue;
            const double u0 = rng.uniform(reach, W - reach);
            const double v0 = rng.uniform(reach, H - reach);
            const Mat3 pose = camera_pose(u0, v0, side, roll, tilt, axis, focal);
            // bounding circle of the projected quiet-zone square
            const double q = 1.0 + 2.0 / dim;
            double r = 0;
            bool inside = true;
            const double cq[4][2] = {{-q, -q}, {q, -q}, {q, q}, {-q, q}};
            for (int i = 0; i < 4; ++i) {
                double px, py;
                pose.project(cq[i][0], cq[i][1], &px, &py);
                r = std::max(r, std::hypot(px - u0, py - v0));
                if (px < 2 || py < 2 || px > W - 3 || py > H - 3) inside = false;
            }
            if (!inside) continue;
            bool clear = true;
            for (const WorkTag& o : placed) {
                if (std::hypot(o.cx - u0, o.cy - v0) < o.radius + r + 3.0) {
                    clear = false;
                    break;
                }
            }
            if (!clear) continue;
            wt.pose = pose;
            wt.cx = u0;
            wt.cy = v0;
            wt.radius = r;
            ok = true;
        }
        if (ok) placed.push_back(wt);
    }

    // Clutter: black rectangles and glyph specks, kept off the tags.
    auto off_tags = [&](double x, double y, double rad) {
        for (const WorkTag& o : placed)
            if (std::hypot(o.cx - x, o.cy - y) < o.radius + rad + 2.0) return false;
        return true;
    };
    for (int i = 0; i < spec.rectangles; ++i) {
        const double w = rng.uniform(6.0, 90.0), h = rng.uniform(6.0, 90.0);
        const double x = rng.uniform(0.0, W - 1.0), y = rng.uniform(0.0, H - 1.0);
        const double a = rng.uniform(0.0, M_PI);
        const float level = static_cast<float>(rng.uniform(0.0, 60.0));
        const double rad = 0.5 * std::hypot(w, h);
        if (!off_tags(x, y, rad)) continue;
        const double ca = std::cos(a), sa = std::sin(a);
        const int bx0 = std::max(0, static_cast<int>(x - rad)), bx1 = std::min(W - 1, static_cast<int>(x + rad));
        const int by0 = std::max(0, static_cast<int>(y - rad)), by1 = std::min(H - 1, static_cast<int>(y + rad));
        for (int py = by0; py <= by1; ++py)
            for (int px = bx0; px <= bx1; ++px) {
                const double dx = px + 0.5 - x, dy = py + 0.5 - y;
                const double lx = dx * ca + dy * sa, ly = -dx * sa + dy * ca;
                if (std::fabs(lx) <= 0.5 * w && std::fabs(ly) <= 0.5 * h)
                    img[static_cast<size_t>(py) * W + px] = level;
            }
    }
    for (int i = 0; i < spec.specks; ++i) {
        const double area = rng.uniform(4.0, 400.0);
        const double aspect = rng.uniform(0.3, 1.0);
        const double rx = std::sqrt(area / (M_PI * aspect)), ry = rx * aspect;
        const double x = rng.uniform(0.0, W - 1.0), y = rng.uniform(0.0, H - 1.0);
        const double a = rng.uniform(0.0, M_PI);
        const float level = static_cast<float>(rng.uniform(0.0, 70.0));
        if (!off_tags(x, y, rx)) continue;
        const double ca = std::cos(a), sa = std::sin(a);
        const int bx0 = std::max(0, static_cast<int>(x - rx - 1)), bx1 = std::min(W - 1, static_cast<int>(x + rx + 1));
        const int by0 = std::max(0, static_cast<int>(y - rx - 1)), by1 = std::min(H - 1, static_cast<int>(y + rx + 1));
        for (int py = by0; py <= by1; ++py)
            for (int px = bx0; px <= bx1; ++px) {
                const double dx = px + 0.5 - x, dy = py + 0.5 - y;
                const double lx = dx * ca + dy * sa, ly = -dx * sa + dy * ca;
                if ((lx * lx) / (rx * rx) + (ly * ly) / (ry * ry) <= 1.0)
                    img[static_cast<size_t>(py) * W + px] = level;
            }
    }

    for (const WorkTag& t : placed) render_work_tag(img.data(), W, H, t, spec.supersample);

    if (spec.blur_sigma > 0.0) {
        gaussian_blur(img, W, H, spec.blur_sigma);
    } else {
        box_blur3(img, W, H);
    }
    // Gaussian noise by Box-Muller on an independent stream, then round/clamp.
    Rng noise(static_cast<uint64_t>(config), seed, 1);
    const size_t n = img.size();
    for (size_t i = 0; i < n; i += 2) {
        const double u1 = std::max(noise.unit(), 1e-300), u2 = noise.unit();
        const double mag = spec.noise_sigma * std::sqrt(-2.0 * std::log(u1));
        const double z0 = mag * std::cos(2.0 * M_PI * u2), z1 = mag * std::sin(2.0 * M_PI * u2);
        for (int k = 0; k < 2 && i + k < n; ++k) {
            const double v = std::nearbyint(img[i + k] + (k ? z1 : z0));
            image[i + k] = static_cast<uint8_t>(v < 0 ? 0 : (v > 255 ? 255 : v));
        }
    }

    int written = 0;
    for (const WorkTag& t : placed) {
        if (!tags || written >= max_tags) break;
        bds_tag& o = tags[written++];
        o.family = t.family;
        o.id = t.id;
        o.decoy = t.decoy;
        o.reserved = 0;
        std::memcpy(o.pose, t.pose.v, sizeof t.pose.v);
        square_corners(t.pose, o.corners);
    }
    return tags ? written : static_cast<int32_t>(placed.size());
}

// Generates frames [first, first + count) of one workload into consecutive
// buffers of width*height bytes, on up to `threads` host threads.
__attribute__((visibility("default"))) int32_t bds_scene_batch(int32_t config, int32_t width,
                                                              int32_t height, uint64_t first,
                                                              int32_t count, uint8_t* images,
                                                              int32_t threads) {
    if (count <= 0 || !images) return -1;
    const int nt = std::max(1, std::min(threads, count));
    std::vector<std::thread> pool;
    const size_t frame = static_cast<size_t>(width) * height;
    for (int t = 0; t < nt; ++t) {
        pool.emplace_back([=] {
            for (int i = t; i < count; i += nt)
                bds_scene(config, width, height, first + static_cast<uint64_t>(i),
                          images + frame * static_cast<size_t>(i), nullptr, 0);
        });
    }
    for (auto& th : pool) th.join();
    return count;
}

}  // extern "C"
