// Test program for include/supra_bf.hpp (the RxBeamformerHip engine shim).
//
//   shim_test host  — builds the 128-element linear geometry with
//                     supra_geometry_build, hands it to the shim through
//                     stand-ins of SUPRA's RxBeamformerParameters and
//                     ScanlineRxParameters3D (same member and accessor names),
//                     and checks that the shim's conversion returns every
//                     field unchanged.  No GPU needed.
//   shim_test gpu   — also beamforms int16 and float RF through the shim and
//                     through supra_das_run directly and checks the images
//                     are bit-identical.
// Exit status 0 on success.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "supra_bf.hpp"

namespace stand_in {
// The member and accessor names of SUPRA's types (USImageProperties.h:43-95,
// RxBeamformerParameters.h:69-79, vec.h), holding the values the shim reads.
template <typename T> struct vec2T { T x, y; };
struct vec { double x, y, z; };
struct ScanlineRxParameters3D {
    struct TransmitParameters {
        vec2T<uint16_t> firstActiveElementIndex, lastActiveElementIndex;
        uint16_t txScanlineIdx;
        double initialDelay;
    };
    vec position, direction;
    double txWeights[4];
    TransmitParameters txParameters[4];
    vec2T<double> maxElementDistance;
};
struct RxBeamformerParameters {
    size_t numRx;
    vec2T<size_t> layout;
    double c;
    std::vector<float> depths, ex, ey;
    std::vector<ScanlineRxParameters3D> scanlines;
    std::vector<int32_t> map;
    size_t getNumRxScanlines() const { return numRx; }
    vec2T<size_t> getRxScanlineLayout() const { return layout; }
    double getSpeedOfSoundMMperS() const { return c; }
    const std::vector<float>& getRxDepths() const { return depths; }
    const std::vector<ScanlineRxParameters3D>& getRxScanlines() const { return scanlines; }
    const std::vector<float>& getRxElementXs() const { return ex; }
    const std::vector<float>& getRxElementYs() const { return ey; }
    bool getNonlinearElementToChannelMapping() const { return false; }
    const std::vector<int32_t>& getElementToChannelMap() const { return map; }
    size_t getRxNumDepths() const { return depths.size(); }
};
}  // namespace stand_in

static int fails = 0;
#define CHECK(cond, ...)                        \
    do {                                        \
        if (!(cond)) {                          \
            std::printf("FAIL: " __VA_ARGS__);  \
            std::printf("\n");                  \
            fails++;                            \
        }                                       \
    } while (0)

int main(int argc, char** argv) {
    const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
    supra_beamformer_config cfg{};
    cfg.scan_type = SUPRA_SCAN_LINEAR;
    cfg.transducer_type = SUPRA_TRANSDUCER_LINEAR;
    cfg.element_layout[0] = 128, cfg.element_layout[1] = 1;
    cfg.pitch[0] = cfg.pitch[1] = 0.3;
    cfg.num_scanlines[0] = 128, cfg.num_scanlines[1] = 1;
    cfg.rx_scanline_subdivision[0] = cfg.rx_scanline_subdivision[1] = 1;
    cfg.max_aperture[0] = 64, cfg.max_aperture[1] = 1;
    cfg.tx_max_aperture[0] = 64, cfg.tx_max_aperture[1] = 1;
    cfg.depth = 40.0;
    cfg.num_depths = 520;
    cfg.tx_focus_active = 1;
    cfg.tx_focus_depth = 20.0;
    cfg.speed_of_sound = 1540.0;
    uint32_t n_rx = 0, n_el = 0, layout[2];
    if (supra_geometry_sizes(&cfg, &n_rx, layout, &n_el) != SUPRA_OK) {
        std::printf("supra_geometry_sizes: %s\n", supra_last_error());
        return 2;
    }
    std::vector<supra_scanline_rx> scan(n_rx);
    std::vector<float> depths(cfg.num_depths), ex(n_el), ey(n_el);
    if (supra_geometry_build(&cfg, scan.data(), depths.data(), ex.data(), ey.data()) != SUPRA_OK) {
        std::printf("supra_geometry_build: %s\n", supra_last_error());
        return 2;
    }
    // the geometry as SUPRA would hold it
    stand_in::RxBeamformerParameters p;
    p.numRx = n_rx;
    p.layout = {layout[0], layout[1]};
    p.c = cfg.speed_of_sound * 1000.0;
    p.depths = depths;
    p.ex = ex;
    p.ey = ey;
    for (const supra_scanline_rx& s : scan) {
        stand_in::ScanlineRxParameters3D q{};
        q.position = {s.position[0], s.position[1], s.position[2]};
        q.direction = {s.direction[0], s.direction[1], s.direction[2]};
        for (int k = 0; k < 4; k++) {
            q.txWeights[k] = s.tx_weights[k];
            q.txParameters[k].firstActiveElementIndex = {s.tx[k].first_active_element[0], s.tx[k].first_active_element[1]};
            q.txParameters[k].lastActiveElementIndex = {s.tx[k].last_active_element[0], s.tx[k].last_active_element[1]};
            q.txParameters[k].txScanlineIdx = s.tx[k].tx_scanline_idx;
            q.txParameters[k].initialDelay = s.tx[k].initial_delay;
        }
        q.maxElementDistance = {s.max_element_distance[0], s.max_element_distance[1]};
        p.scanlines.push_back(q);
    }
    // host: the shim's conversion returns every field
    for (uint32_t i = 0; i < n_rx; i++) {
        const supra_scanline_rx c = supra_amd::to_c(p.scanlines[i]);
        const supra_scanline_rx& o = scan[i];
        bool same = std::memcmp(c.position, o.position, sizeof(o.position)) == 0 &&
                    std::memcmp(c.direction, o.direction, sizeof(o.direction)) == 0 &&
                    std::memcmp(c.tx_weights, o.tx_weights, sizeof(o.tx_weights)) == 0 &&
                    std::memcmp(c.max_element_distance, o.max_element_distance, sizeof(o.max_element_distance)) == 0;
        for (int k = 0; k < 4; k++)
            same = same && c.tx[k].first_active_element[0] == o.tx[k].first_active_element[0] &&
                   c.tx[k].first_active_element[1] == o.tx[k].first_active_element[1] &&
                   c.tx[k].last_active_element[0] == o.tx[k].last_active_element[0] &&
                   c.tx[k].last_active_element[1] == o.tx[k].last_active_element[1] &&
                   c.tx[k].tx_scanline_idx == o.tx[k].tx_scanline_idx &&
                   std::memcmp(&c.tx[k].initial_delay, &o.tx[k].initial_delay, sizeof(double)) == 0;
        CHECK(same, "scanline %u differs after the shim's conversion", i);
    }
    std::printf("host: %u scanlines converted, %d failures\n", n_rx, fails);
    if (!gpu) return fails ? 1 : 0;

    // GPU: the shim against supra_das_run on the same context geometry
    supra_amd::RxBeamformerHip shim(p, cfg.element_layout[0], cfg.element_layout[1], cfg.depth);
    supra_rx_geometry g{};
    g.num_rx_scanlines = n_rx;
    g.rx_scanline_layout[0] = layout[0], g.rx_scanline_layout[1] = layout[1];
    g.num_depths = cfg.num_depths;
    g.num_elements = n_el;
    g.element_layout[0] = cfg.element_layout[0], g.element_layout[1] = cfg.element_layout[1];
    g.scanlines = scan.data();
    g.depths = depths.data();
    g.element_x = ex.data();
    g.element_y = ey.data();
    g.speed_of_sound_mm_per_s = p.c;
    g.depth = cfg.depth;
    supra_bf_ctx* ctx = nullptr;
    if (supra_bf_create(&g, &ctx) != SUPRA_OK) {
        std::printf("supra_bf_create: %s\n", supra_last_error());
        return 2;
    }
    const uint32_t n_ch = 64, n_s = 2100, n_tx = n_rx;
    const size_t n_rf = static_cast<size_t>(n_tx) * n_ch * n_s, n_img = static_cast<size_t>(n_rx) * cfg.num_depths;
    std::vector<int16_t> rf16(n_rf);
    std::vector<float> rf32(n_rf);
    for (size_t i = 0; i < n_rf; i++) {
        const double v = 8000.0 * std::sin(0.37 * static_cast<double>(i % n_s) + 0.011 * static_cast<double>(i / n_s));
        rf16[i] = static_cast<int16_t>(v);
        rf32[i] = static_cast<float>(v);
    }
    int16_t* d_rf16 = nullptr;
    float *d_rf32 = nullptr, *d_a = nullptr, *d_b = nullptr;
    bool ok = hipMalloc(&d_rf16, n_rf * sizeof(int16_t)) == hipSuccess && hipMalloc(&d_rf32, n_rf * sizeof(float)) == hipSuccess &&
              hipMalloc(&d_a, n_img * sizeof(float)) == hipSuccess && hipMalloc(&d_b, n_img * sizeof(float)) == hipSuccess &&
              hipMemcpy(d_rf16, rf16.data(), n_rf * sizeof(int16_t), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d_rf32, rf32.data(), n_rf * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        std::printf("device buffers failed\n");
        return 2;
    }
    std::vector<float> a(n_img), b(n_img);
    for (int pass = 0; pass < 2; pass++) {
        int st1, st2;
        if (pass == 0) {
            st1 = shim.performRxBeamforming<int16_t, float>(supra_amd::RxBeamformerHip::DelayAndSum, d_rf16, n_tx, n_ch,
                                                            n_s, 40e6, 1.0, 1.54e6, SUPRA_WINDOW_HANN, 0.0f, false, 0, d_a);
            st2 = supra_das_run(ctx, SUPRA_BF_DELAY_AND_SUM, d_rf16, SUPRA_INT16, n_tx, n_ch, n_s, 40e6, 1.0, 1.54e6,
                                SUPRA_WINDOW_HANN, 0.0f, 0, 0, d_b, SUPRA_FLOAT32, nullptr);
        } else {
            st1 = shim.performRxBeamforming<float, float>(supra_amd::RxBeamformerHip::CoherenceFactorDelayAndSum, d_rf32,
                                                          n_tx, n_ch, n_s, 40e6, 1.5, 1.50e6, SUPRA_WINDOW_HAMMING,
                                                          0.0f, false, 3, d_a);
            st2 = supra_das_run(ctx, SUPRA_BF_COHERENCE_FACTOR, d_rf32, SUPRA_FLOAT32, n_tx, n_ch, n_s, 40e6, 1.5,
                                1.50e6, SUPRA_WINDOW_HAMMING, 0.0f, 0, 3, d_b, SUPRA_FLOAT32, nullptr);
        }
        CHECK(st1 == SUPRA_OK && st2 == SUPRA_OK, "pass %d: status %d / %d: %s", pass, st1, st2, supra_last_error());
        ok = hipDeviceSynchronize() == hipSuccess &&
             hipMemcpy(a.data(), d_a, n_img * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess &&
             hipMemcpy(b.data(), d_b, n_img * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess;
        CHECK(ok, "pass %d: device copy", pass);
        size_t nonzero = 0;
        for (size_t i = 0; i < n_img; i++) nonzero += a[i] != 0.0f;
        CHECK(std::memcmp(a.data(), b.data(), n_img * sizeof(float)) == 0, "pass %d: shim and direct call differ", pass);
        CHECK(nonzero > n_img / 2, "pass %d: image mostly zero (%zu of %zu non-zero)", pass, nonzero, n_img);
        std::printf("gpu pass %d: %zu of %zu samples non-zero, shim == supra_das_run\n", pass, nonzero, n_img);
    }
    (void)hipFree(d_rf16);
    (void)hipFree(d_rf32);
    (void)hipFree(d_a);
    (void)hipFree(d_b);
    supra_bf_destroy(ctx);
    std::printf("%d failures\n", fails);
    return fails ? 1 : 0;
}
