// The C++ Detector API's observable behaviour (reference
// detector/src/detector.cpp:227-380 and detector/tests/test_detector.cpp):
// exceptions and their messages, the output-capacity clamp, move semantics
// and stats.  Run by tests/test_cpp_api.py.
//
//   test_detector_api cpu             checks that need no GPU
//   test_detector_api gpu FRAME.pgm   detection on a 640x480 frame holding
//                                      tag36h11 id 11
//
// Prints one "ok <check>" line per passed check; exits 1 at the first
// failure.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include <hip/hip_runtime.h>

#include "bdcudatag/bdcudatag.h"
#include "bdcudatag/detector.hpp"

using bdcudatag::Detection;
using bdcudatag::Detector;
using bdcudatag::DetectorConfig;
using bdcudatag::Family;

namespace {

[[noreturn]] void fail(const std::string& what) {
    std::printf("FAIL %s\n", what.c_str());
    std::fflush(stdout);
    std::exit(1);
}

void ok(const char* what) {
    std::printf("ok %s\n", what);
    std::fflush(stdout);
}

// Runs f, which must throw E whose message contains `needle`.
template <class E>
void expect_throw(const char* what, const std::function<void()>& f, const char* needle) {
    try {
        f();
    } catch (const E& e) {
        if (std::string(e.what()).find(needle) == std::string::npos)
            fail(std::string(what) + ": message '" + e.what() + "' lacks '" + needle + "'");
        ok(what);
        return;
    } catch (const std::exception& e) {
        fail(std::string(what) + ": wrong exception type: " + e.what());
    }
    fail(std::string(what) + ": no exception");
}

std::vector<uint8_t> read_pgm(const char* path, int* w, int* h) {
    std::ifstream in(path, std::ios::binary);
    std::string magic;
    int maxval = 0;
    in >> magic >> *w >> *h >> maxval;
    in.get();
    if (magic != "P5" || maxval != 255) fail("pgm header");
    std::vector<uint8_t> img(static_cast<size_t>(*w) * *h);
    in.read(reinterpret_cast<char*>(img.data()), static_cast<std::streamsize>(img.size()));
    if (!in) fail("pgm data");
    return img;
}

int cpu_checks() {
    expect_throw<std::runtime_error>("unknown family throws", [] { Detector d(static_cast<Family>(7), 640, 480); },
                                     "unknown tag family");
    if (std::string(bd_last_error()) != "unknown tag family") fail("last error after C++ failure");
    ok("last error set by the C++ API");
    expect_throw<std::runtime_error>("width below 16 throws", [] { Detector d(Family::Tag36h11, 8, 480); },
                                     "image dimensions out of range");
    expect_throw<std::runtime_error>("height above 16384 throws",
                                     [] { Detector d(Family::Tag16h5, 640, 20000); }, "image dimensions");
    return 0;
}

int gpu_checks(const char* pgm) {
    int w = 0, h = 0;
    const std::vector<uint8_t> img = read_pgm(pgm, &w, &h);
    DetectorConfig cfg;
    cfg.max_detections = 32;
    cfg.min_decision_margin = 2.0f;
    Detector det(Family::Tag36h11, w, h, cfg);
    if (det.width() != w || det.height() != h || det.family() != Family::Tag36h11) fail("accessors");
    ok("accessors");

    std::vector<Detection> found = det.detect(img.data(), w);
    if (found.size() != 1 || found[0].id != 11 || found[0].hamming != 0) fail("host detect id 11");
    ok("host detect finds id 11");
    if (found[0].homography[8] != 1.0f) fail("homography[8] == 1");
    ok("homography normalised");

    // Device input gives the same records.
    uint8_t* dev = nullptr;
    if (hipMalloc(&dev, img.size()) != hipSuccess) fail("hipMalloc");
    if (hipMemcpy(dev, img.data(), img.size(), hipMemcpyHostToDevice) != hipSuccess) fail("hipMemcpy");
    Detection out[4];
    const int n = det.detectDevice(dev, w, out, 4);
    if (n != 1 || out[0].id != 11 || out[0].corners != found[0].corners || out[0].homography != found[0].homography)
        fail("device detect equals host detect");
    ok("device detect equals host detect");

    // Capacity: 0 writes nothing and is valid with a null buffer; negative
    // or a null buffer with room throws invalid_argument.
    if (det.detect(img.data(), w, nullptr, 0) != 0) fail("capacity 0");
    ok("capacity 0 writes nothing");
    expect_throw<std::invalid_argument>("negative capacity throws",
                                        [&] { det.detect(img.data(), w, out, -1); }, "invalid detection output");
    expect_throw<std::invalid_argument>("null output throws",
                                        [&] { det.detectDevice(dev, w, nullptr, 2); }, "invalid detection output");
    // A capacity above max_detections is clamped to it.
    std::vector<Detection> many(100);
    if (det.detect(img.data(), w, many.data(), 100) != 1) fail("capacity clamp");
    ok("capacity above max_detections is clamped");
    // Invalid frames are reported as runtime errors with the C ABI's text.
    expect_throw<std::runtime_error>("short stride throws", [&] { det.detect(img.data(), w - 1, out, 4); },
                                     "stride is smaller than the image width");
    expect_throw<std::runtime_error>("null image throws", [&] { det.detect(nullptr, w, out, 4); },
                                     "null image pointer");
    expect_throw<std::runtime_error>("null device image throws", [&] { det.detectDevice(nullptr, w, out, 4); },
                                     "null device image pointer");

    bdcudatag::Stats st = det.stats();
    if (st.frames != 4 || st.detections != 1 || st.candidates < 1 || !(st.last_gpu_ms > 0.0f))
        fail("stats after four frames");
    ok("stats count the frames run");

    // Move construction and assignment transfer the engine; the source is
    // left moved-from.
    Detector moved(std::move(det));
    if (moved.detect(img.data(), w).size() != 1) fail("moved-to detector detects");
    ok("move construction transfers the engine");
    expect_throw<std::runtime_error>("moved-from detect throws", [&] { det.detect(img.data(), w); },
                                     "moved from");
    if (det.stats().frames != 0) fail("moved-from stats are empty");
    ok("moved-from stats are empty");
    Detector other(Family::Tag16h5, 320, 240);
    other = std::move(moved);
    if (other.width() != w || other.family() != Family::Tag36h11) fail("move assignment takes the size");
    if (other.detect(img.data(), w).size() != 1) fail("move-assigned detector detects");
    ok("move assignment transfers the engine");
    expect_throw<std::runtime_error>("moved-from detectDevice throws",
                                     [&] { moved.detectDevice(dev, w, out, 4); }, "moved from");
    (void)hipFree(dev);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc >= 2 && std::strcmp(argv[1], "cpu") == 0) return cpu_checks();
    if (argc >= 3 && std::strcmp(argv[1], "gpu") == 0) return gpu_checks(argv[2]);
    std::fprintf(stderr, "usage: %s cpu | gpu FRAME.pgm\n", argv[0]);
    return 2;
}
