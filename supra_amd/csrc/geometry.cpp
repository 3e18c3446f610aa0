// Host-side geometry producer: transducer element centres, transmit events and
// the receive scanline table (ScanlineRxParameters3D) that the beamformers read.
//
// Follows the reference's Beamformer / USTransducer / RxBeamformerParameters
// (src/SupraLib/Beamformer/Beamformer.cpp:409-1185, USTransducer.cpp:179-258,
// RxBeamformerParameters.cpp:38-88).  Every double expression is evaluated in
// the reference's operand order so that aperture rounding (std::round of
// x.5 values) and focus delays come out bit-identical; this is host code that
// runs once per imaging setting and is never timed.
#include "supra_bf.h"
#include "status.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace supra_amd {
namespace {

struct D2 { double x, y; };
struct D3 { double x, y, z; };
struct S2 { size_t x, y; };

inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 scale(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline D3 mul(D3 a, D3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline double dot3(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm3(D3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline D3 unit3(D3 a) { double n = norm3(a); return {a.x / n, a.y / n, a.z / n}; }

// Machine-epsilon guard of the reference's slerp (utilities/utility.h:43).
constexpr double kSlerpEps = 2.2204e-16;

// Spherical interpolation between unit vectors (vec.h:422-445).
D3 slerp(D3 a, D3 b, double t) {
    double omega = std::acos(dot3(a, b));
    if (omega < kSlerpEps) return a;
    if (t < kSlerpEps) return a;
    if (1.0 - t < kSlerpEps) return b;
    double so = std::sin(omega);
    return scale(a, std::sin((1.0 - t) * omega) / so) + scale(b, std::sin(t * omega) / so);
}

struct Aperture { S2 begin, end; };   // inclusive element index corners

// Beamformer::computeAperture (Beamformer.cpp:1147-1185): centre the aperture
// on a fractional element position, shift it inside the array, round half away.
Aperture place_aperture(S2 layout, S2 size, D2 rel) {
    double bx = rel.x - static_cast<double>(size.x - 1) / 2;
    double by = rel.y - static_cast<double>(size.y - 1) / 2;
    double ex = rel.x + static_cast<double>(size.x - 1) / 2;
    double ey = rel.y + static_cast<double>(size.y - 1) / 2;
    if (std::round(bx) < 0) { bx = 0; ex = static_cast<double>(size.x - 1); }
    else if (std::round(ex) > static_cast<double>(layout.x - 1)) {
        ex = static_cast<double>(layout.x - 1); bx = ex - static_cast<double>(size.x - 1);
    }
    if (std::round(by) < 0) { by = 0; ey = static_cast<double>(size.y - 1); }
    else if (std::round(ey) > static_cast<double>(layout.y - 1)) {
        ey = static_cast<double>(layout.y - 1); by = ey - static_cast<double>(size.y - 1);
    }
    Aperture a;
    a.begin = {static_cast<size_t>(std::round(bx)), static_cast<size_t>(std::round(by))};
    a.end = {static_cast<size_t>(std::round(ex)), static_cast<size_t>(std::round(ey))};
    return a;
}

// The parts of a transmit event that the receive side consumes
// (ScanlineTxParameters3D, Beamformer.h:28-49).
struct TxEvent {
    Aperture active;
    D3 position;
    D3 direction;
    double max_delay;   // [s] -> TransmitParameters::initialDelay
};

class Producer {
public:
    explicit Producer(const supra_beamformer_config& c) : cfg_(c) {
        layout_ = {cfg_.element_layout[0], cfg_.element_layout[1]};
        if (cfg_.transducer_type == SUPRA_TRANSDUCER_LINEAR) layout_.y = 1;
        num_elements_ = layout_.x * layout_.y;
        c_mm_ = cfg_.speed_of_sound * 1000.0;
        steer_ = {cfg_.steering_angle_deg[0] * M_PI / 180.0, cfg_.steering_angle_deg[1] * M_PI / 180.0};
        sector_ = {cfg_.sector_angle_deg[0] * M_PI / 180.0, cfg_.sector_angle_deg[1] * M_PI / 180.0};
        n_tx_ = {cfg_.num_scanlines[0], cfg_.num_scanlines[1]};
        subdiv_ = {cfg_.rx_scanline_subdivision[0], cfg_.rx_scanline_subdivision[1]};
        if (cfg_.scan_type == SUPRA_SCAN_PLANEWAVE) n_rx_ = subdiv_;
        else n_rx_ = {n_tx_.x + (subdiv_.x - 1) * (n_tx_.x - 1), n_tx_.y + (subdiv_.y - 1) * (n_tx_.y - 1)};
        rx_aperture_ = {cfg_.max_aperture[0], cfg_.max_aperture[1]};
        tx_aperture_ = {cfg_.tx_max_aperture[0], cfg_.tx_max_aperture[1]};
    }

    int validate() const {
        if (cfg_.scan_type < SUPRA_SCAN_LINEAR || cfg_.scan_type > SUPRA_SCAN_PLANEWAVE)
            return fail(SUPRA_ERR_INVALID_ARG, "geometry: unknown scan_type %d", cfg_.scan_type);
        if (cfg_.transducer_type != SUPRA_TRANSDUCER_LINEAR && cfg_.transducer_type != SUPRA_TRANSDUCER_PLANAR)
            return fail(SUPRA_ERR_INVALID_ARG, "geometry: unknown transducer_type %d", cfg_.transducer_type);
        if (layout_.x < 2 || layout_.y < 1)
            return fail(SUPRA_ERR_INVALID_ARG, "geometry: element layout must have >= 2 elements in x");
        if (n_tx_.x < 2 || n_tx_.y < 1 || subdiv_.x < 1 || subdiv_.y < 1)
            return fail(SUPRA_ERR_INVALID_ARG, "geometry: need >= 2 scanlines in x and subdivision >= 1");
        if (cfg_.scan_type == SUPRA_SCAN_BIPHASED && n_tx_.y < 2)
            return fail(SUPRA_ERR_INVALID_ARG, "geometry: biphased scans need >= 2 scanlines in y");
        if (rx_aperture_.x < 1 || rx_aperture_.y < 1 || rx_aperture_.x > layout_.x || rx_aperture_.y > layout_.y ||
            tx_aperture_.x < 1 || tx_aperture_.y < 1 || tx_aperture_.x > layout_.x || tx_aperture_.y > layout_.y)
            return fail(SUPRA_ERR_INVALID_ARG, "geometry: apertures must lie within the element layout");
        if (!(cfg_.depth > 0) || cfg_.num_depths < 2 || !(cfg_.speed_of_sound > 0) || cfg_.tx_focus_depth < 0)
            return fail(SUPRA_ERR_INVALID_ARG, "geometry: depth, num_depths >= 2, speed of sound must be positive");
        if (!(cfg_.pitch[0] > 0) || (cfg_.transducer_type == SUPRA_TRANSDUCER_PLANAR && !(cfg_.pitch[1] > 0)))
            return fail(SUPRA_ERR_INVALID_ARG, "geometry: pitch must be positive");
        if (cfg_.scan_type == SUPRA_SCAN_PLANEWAVE && cfg_.transducer_type != SUPRA_TRANSDUCER_LINEAR)
            return fail(SUPRA_ERR_UNSUPPORTED, "geometry: plane waves need a linear array (Beamformer.cpp:540-545)");
        if (cfg_.scan_type == SUPRA_SCAN_LINEAR && cfg_.transducer_type != SUPRA_TRANSDUCER_LINEAR)
            return fail(SUPRA_ERR_UNSUPPORTED, "geometry: linear scans need a linear array (Beamformer.cpp:429-434)");
        if (cfg_.scan_type == SUPRA_SCAN_PHASED && rx_aperture_.x != num_elements_)
            return fail(SUPRA_ERR_UNSUPPORTED, "geometry: phased scans use the full aperture (Beamformer.cpp:466-467)");
        size_t total = n_rx_.x * n_rx_.y * (cfg_.scan_type == SUPRA_SCAN_PLANEWAVE ? n_tx_.x : 1);
        if (total == 0 || total > (1u << 24) || n_tx_.x * n_tx_.y >= 65535)
            return fail(SUPRA_ERR_TOO_LARGE, "geometry: scanline count out of range");
        return SUPRA_OK;
    }

    void layout_rx(uint32_t* n, uint32_t lay[2]) const {
        if (cfg_.scan_type == SUPRA_SCAN_PLANEWAVE) {
            lay[0] = static_cast<uint32_t>(n_rx_.x * n_tx_.x);
            lay[1] = static_cast<uint32_t>(n_rx_.y * n_tx_.y);
        } else {
            lay[0] = static_cast<uint32_t>(n_rx_.x);
            lay[1] = static_cast<uint32_t>(n_rx_.y);
        }
        *n = lay[0] * lay[1];
    }
    uint32_t num_elements() const { return static_cast<uint32_t>(num_elements_); }

    void build(supra_scanline_rx* out, float* depths, float* ex, float* ey) {
        element_centres();
        transmit_events();
        receive_scanlines(out);
        // RxBeamformerParameters.cpp:61-64, 76-81
        for (uint32_t z = 0; z < cfg_.num_depths; z++)
            depths[z] = static_cast<float>(z * cfg_.depth / (cfg_.num_depths - 1));
        for (size_t e = 0; e < num_elements_; e++) {
            ex[e] = static_cast<float>(centres_[e].x);
            ey[e] = static_cast<float>(centres_[e].y);
        }
    }

private:
    // USTransducer::computeInternals (USTransducer.cpp:189-243) with uniform pitch.
    void element_centres() {
        centres_.assign(num_elements_, D3{0, 0, 0});
        if (cfg_.transducer_type == SUPRA_TRANSDUCER_LINEAR) {
            double width = 0.0;
            for (size_t i = 0; i + 1 < num_elements_; i++) width += cfg_.pitch[0];
            double x = -width / 2;
            for (size_t i = 0; i < num_elements_; i++) {
                centres_[i] = {x, 0, 0};
                if (i < num_elements_ - 1) x += cfg_.pitch[0];
            }
        } else {
            double wx = 0.0, wy = 0.0;
            for (size_t i = 0; i + 1 < layout_.x; i++) wx += cfg_.pitch[0];
            for (size_t i = 0; i + 1 < layout_.y; i++) wy += cfg_.pitch[1];
            double y = -wy / 2;
            for (size_t iy = 0; iy < layout_.y; iy++) {
                double x = -wx / 2;
                for (size_t ix = 0; ix < layout_.x; ix++) {
                    centres_[ix + layout_.x * iy] = {x, y, 0};
                    if (ix < layout_.x - 1) x += cfg_.pitch[0];
                }
                if (iy < layout_.y - 1) y += cfg_.pitch[1];
            }
        }
    }

    const D3& centre(size_t ix, size_t iy) const { return centres_[ix + iy * layout_.x]; }

    // Beamformer::getTxScanline3D (Beamformer.cpp:840-1066), reduced to the
    // fields the receive path reads: apertures, focus-delay maximum, position
    // and direction (with the reference's y/z swap, :1061-1063).
    TxEvent tx_event(Aperture active, Aperture txa, D2 start, D2 angle) const {
        D3 start3{start.x, start.y, 0};
        struct Corner { D2 rel; S2 idx; };
        std::vector<Corner> corners;
        if (txa.begin.x == txa.end.x && txa.begin.y == txa.end.y) {
            corners.push_back({{0.5, 0.5}, txa.begin});
        } else if (txa.begin.x == txa.end.x) {
            corners.push_back({{0.5, 0}, txa.begin});
            corners.push_back({{0.5, 1}, txa.end});
        } else if (txa.begin.y == txa.end.y) {
            corners.push_back({{0, 0.5}, txa.begin});
            corners.push_back({{1, 0.5}, txa.end});
        } else {
            corners.push_back({{0, 0}, {txa.begin.x, txa.begin.y}});
            corners.push_back({{0, 1}, {txa.begin.x, txa.end.y}});
            corners.push_back({{1, 0}, {txa.end.x, txa.begin.y}});
            corners.push_back({{1, 1}, {txa.end.x, txa.end.y}});
        }
        D3 dir = unit3({std::tan(angle.x), std::tan(angle.y), 1});
        D3 perp_x = unit3(D3{1, 0, 0} - scale(dir, dot3(D3{1, 0, 0}, dir)));
        D3 perp_y = unit3(D3{0, 1, 0} - scale(dir, dot3(D3{0, 1, 0}, dir)));

        double max_delay = 0.0;
        if (cfg_.tx_focus_active) {
            double max_transit = 0;
            for (const Corner& c : corners) {
                D3 off = scale(perp_x, (c.rel.x - 0.5) * cfg_.tx_focus_width) +
                         scale(perp_y, (c.rel.y - 0.5) * cfg_.tx_focus_width);
                D3 to_focus = start3 + scale(dir, cfg_.tx_focus_depth) + off - centre(c.idx.x, c.idx.y);
                max_transit = std::max(max_transit, norm3(to_focus) / c_mm_);
            }
            size_t span_x = std::max<size_t>(txa.end.x - txa.begin.x, 1);
            size_t span_y = std::max<size_t>(txa.end.y - txa.begin.y, 1);
            for (size_t ix = txa.begin.x; ix <= txa.end.x; ix++) {
                for (size_t iy = txa.begin.y; iy <= txa.end.y; iy++) {
                    D2 rel{static_cast<double>(ix - txa.begin.x) / static_cast<double>(span_x),
                           static_cast<double>(iy - txa.begin.y) / static_cast<double>(span_y)};
                    if (txa.end.x == txa.begin.x) rel.x = 0.5;
                    if (txa.end.y == txa.begin.y) rel.y = 0.5;
                    D3 off = scale(perp_x, (rel.x - 0.5) * cfg_.tx_focus_width) +
                             scale(perp_y, (rel.y - 0.5) * cfg_.tx_focus_width);
                    D3 to_focus = start3 + scale(dir, cfg_.tx_focus_depth) + off - centre(ix, iy);
                    double delay = max_transit - norm3(to_focus) / c_mm_;
                    max_delay = std::max(max_delay, delay);
                }
            }
        } else {
            double max_transit = 0;
            for (const Corner& c : corners) {
                D3 d = mul(start3 - centre(c.idx.x, c.idx.y), dir);
                max_transit = std::max(max_transit, (d.x + d.y + d.z + cfg_.tx_focus_depth) / c_mm_);
            }
            for (size_t ix = txa.begin.x; ix <= txa.end.x; ix++) {
                for (size_t iy = txa.begin.y; iy <= txa.end.y; iy++) {
                    D3 d = mul(start3 - centre(ix, iy), dir);
                    double delay = max_transit - (d.x + d.y + d.z + cfg_.tx_focus_depth) / c_mm_;
                    max_delay = std::max(max_delay, delay);
                }
            }
        }
        TxEvent ev;
        ev.active = active;
        ev.max_delay = max_delay;
        ev.position = {start.x, 0, start.y};
        ev.direction = {dir.x, dir.z, dir.y};
        return ev;
    }

    // Beamformer::computeTxParameters, transmit half (Beamformer.cpp:419-580).
    void transmit_events() {
        tx_.clear();
        if (cfg_.scan_type == SUPRA_SCAN_LINEAR) {
            const D3& first = centres_.front();
            const D3& last = centres_[num_elements_ - 1];
            for (size_t i = 0; i < n_tx_.x; i++) {
                double pos = first.x + static_cast<double>(i) / (n_tx_.x - 1) * (last.x - first.x);
                double rel = static_cast<double>(i) / (n_tx_.x - 1) * (num_elements_ - 1);
                Aperture act = place_aperture(layout_, rx_aperture_, {rel, 0});
                Aperture txa = place_aperture(layout_, tx_aperture_, {rel, 0});
                double ang = steer_.x - (sector_.x / 2) + (static_cast<double>(i) / (n_tx_.x - 1) * sector_.x);
                tx_.push_back(tx_event(act, txa, {pos, 0}, {ang, 0}));
            }
        } else if (cfg_.scan_type == SUPRA_SCAN_PHASED) {
            double start_x;
            if (num_elements_ % 2 == 1) start_x = centres_[(num_elements_ - 1) / 2].x;
            else start_x = (centres_[num_elements_ / 2 - 1].x + centres_[num_elements_ / 2].x) / 2;
            D2 rel{(static_cast<double>(num_elements_) - 1) / 2, 0};
            Aperture act = place_aperture(layout_, rx_aperture_, rel);
            Aperture txa = place_aperture(layout_, tx_aperture_, rel);
            for (size_t i = 0; i < n_tx_.x; i++) {
                // Integer quotient i / (n - 1), as written in Beamformer.cpp:496:
                // every event but the last is steered to steer - sector/2.
                double ang = steer_.x - (sector_.x / 2) + static_cast<double>(i / (n_tx_.x - 1)) * sector_.x;
                tx_.push_back(tx_event(act, txa, {start_x, 0}, {ang, 0}));
            }
        } else if (cfg_.scan_type == SUPRA_SCAN_BIPHASED) {
            for (size_t iy = 0; iy < n_tx_.y; iy++) {
                for (size_t ix = 0; ix < n_tx_.x; ix++) {
                    // (aperture - 1) / 2 is an integer quotient in Beamformer.cpp:521.
                    D2 rel{static_cast<double>(ix) / static_cast<double>(n_tx_.x - 1) *
                               static_cast<double>(layout_.x - rx_aperture_.x) +
                               static_cast<double>((rx_aperture_.x - 1) / 2),
                           static_cast<double>(iy) / static_cast<double>(n_tx_.y - 1) *
                               static_cast<double>(layout_.y - rx_aperture_.y) +
                               static_cast<double>((rx_aperture_.y - 1) / 2)};
                    Aperture act = place_aperture(layout_, rx_aperture_, rel);
                    Aperture txa = place_aperture(layout_, tx_aperture_, rel);
                    D2 start{(centre(act.begin.x, 0).x + centre(act.end.x, 0).x) / 2,
                             (centre(0, act.begin.y).y + centre(0, act.end.y).y) / 2};
                    D2 ang{steer_.x + (static_cast<double>(ix) / (n_tx_.x - 1) - 0.5) * sector_.x,
                           steer_.y + (static_cast<double>(iy) / (n_tx_.y - 1) - 0.5) * sector_.y};
                    tx_.push_back(tx_event(act, txa, start, ang));
                }
            }
        } else {  // plane waves (Beamformer.cpp:538-574)
            const D3& first = centres_.front();
            const D3& last = centres_[num_elements_ - 1];
            double pos = first.x + 0.5 * (last.x - first.x);
            double rel = 0.5 * (num_elements_ - 1);
            Aperture act = place_aperture(layout_, rx_aperture_, {rel, 0});
            Aperture txa = place_aperture(layout_, tx_aperture_, {rel, 0});
            for (size_t i = 0; i < n_tx_.x; i++) {
                double ang = steer_.x - (sector_.x / 2) + i * sector_.x / (n_tx_.x - 1);
                tx_.push_back(tx_event(act, txa, {pos, 0}, {ang, 0}));
            }
        }
    }

    // Beamformer::getRxScanline3DInterpolated (Beamformer.cpp:1068-1135).
    supra_scanline_rx rx_line(const size_t idx[4], const TxEvent* ev[4], D2 t) const {
        supra_scanline_rx rx;
        std::memset(&rx, 0, sizeof(rx));
        for (int k = 0; k < 4; k++) {
            rx.tx[k].first_active_element[0] = static_cast<uint16_t>(ev[k]->active.begin.x);
            rx.tx[k].first_active_element[1] = static_cast<uint16_t>(ev[k]->active.begin.y);
            rx.tx[k].last_active_element[0] = static_cast<uint16_t>(ev[k]->active.end.x);
            rx.tx[k].last_active_element[1] = static_cast<uint16_t>(ev[k]->active.end.y);
            rx.tx[k].tx_scanline_idx = static_cast<uint16_t>(idx[k]);
            rx.tx[k].initial_delay = ev[k]->max_delay;
        }
        rx.tx_weights[0] = (1 - t.x) * (1 - t.y);
        rx.tx_weights[1] = t.x * (1 - t.y);
        rx.tx_weights[2] = (1 - t.x) * t.y;
        rx.tx_weights[3] = t.x * t.y;
        D3 pos = scale(scale(ev[0]->position, 1 - t.x) + scale(ev[1]->position, t.x), 1 - t.y) +
                 scale(scale(ev[2]->position, 1 - t.x) + scale(ev[3]->position, t.x), t.y);
        D3 dir = slerp(slerp(ev[0]->direction, ev[1]->direction, t.x),
                       slerp(ev[2]->direction, ev[3]->direction, t.x), t.y);
        rx.position[0] = pos.x; rx.position[1] = pos.y; rx.position[2] = pos.z;
        rx.direction[0] = dir.x; rx.direction[1] = dir.y; rx.direction[2] = dir.z;
        double mx = 0, mz = 0;
        for (int k = 0; k < 4; k++) {
            const supra_tx_params& p = rx.tx[k];
            const size_t fx = p.first_active_element[0], fy = p.first_active_element[1];
            const size_t lx = p.last_active_element[0], ly = p.last_active_element[1];
            const D3* cs[4] = {&centre(fx, fy), &centre(fx, ly), &centre(lx, fy), &centre(lx, ly)};
            for (const D3* c : cs) {
                mx = std::max(mx, std::abs(c->x - pos.x));
                mz = std::max(mz, std::abs(c->y - pos.z));
            }
        }
        rx.max_element_distance[0] = mx;
        rx.max_element_distance[1] = mz;
        return rx;
    }

    supra_scanline_rx rx_between(size_t i1, size_t i2, size_t i3, size_t i4, D2 t) const {
        size_t idx[4] = {i1, i2, i3, i4};
        const TxEvent* ev[4] = {&tx_[i1], &tx_[i2], &tx_[i3], &tx_[i4]};
        return rx_line(idx, ev, t);
    }

    // Receive half of computeTxParameters (Beamformer.cpp:582-800); output is
    // flattened x fastest like RxBeamformerParameters.cpp:65-74.
    void receive_scanlines(supra_scanline_rx* out) const {
        const size_t nx = n_tx_.x, ny = n_tx_.y;
        if (cfg_.scan_type == SUPRA_SCAN_PLANEWAVE) {
            const D3& first = centres_.front();
            const D3& last = centres_[num_elements_ - 1];
            Aperture act = place_aperture(layout_, rx_aperture_, {0, 0});
            Aperture txa = place_aperture(layout_, tx_aperture_, {0, 0});
            const size_t nrx = n_rx_.x;
            for (size_t p = 0; p < nx; p++) {
                double ang = steer_.x - (sector_.x / 2) + p * sector_.x / (nx - 1);
                TxEvent a = tx_event(act, txa, {first.x, 0}, {ang, 0});
                TxEvent b = tx_event(act, txa, {last.x, 0}, {ang, 0});
                for (size_t r = 0; r < nrx; r++) {
                    D2 t{static_cast<double>(r) / static_cast<double>(n_rx_.x), 0.0 / static_cast<double>(n_rx_.y)};
                    size_t idx[4] = {0, 1, 0, 1};
                    const TxEvent* ev[4] = {&a, &b, &a, &b};
                    out[r + p * nrx] = rx_line(idx, ev, t);
                }
            }
            return;
        }
        const size_t rnx = n_rx_.x;
        auto put = [&](size_t x, size_t y, const supra_scanline_rx& s) { out[x + y * rnx] = s; };
        const double sx = static_cast<double>(subdiv_.x), sy = static_cast<double>(subdiv_.y);
        if (ny > 1) {
            for (size_t y = 0; y + 1 < ny; y++) {
                for (size_t x = 0; x + 1 < nx; x++) {
                    size_t i1 = x + y * nx, i2 = (x + 1) + y * nx, i3 = x + (y + 1) * nx, i4 = (x + 1) + (y + 1) * nx;
                    for (size_t ry = 0; ry < subdiv_.y; ry++)
                        for (size_t rx = 0; rx < subdiv_.x; rx++)
                            put(x * subdiv_.x + rx, y * subdiv_.y + ry,
                                rx_between(i1, i2, i3, i4, {static_cast<double>(rx) / sx, static_cast<double>(ry) / sy}));
                }
                size_t x = nx - 1;
                size_t i1 = x + y * nx, i3 = x + (y + 1) * nx;
                for (size_t ry = 0; ry < subdiv_.y; ry++)
                    put(x * subdiv_.x, y * subdiv_.y + ry,
                        rx_between(i1, i1, i3, i3, {0.0 / sx, static_cast<double>(ry) / sy}));
                put((nx - 1) * subdiv_.x, y * subdiv_.y, rx_between(i1, i1, i1, i1, {0, 0}));
            }
            size_t y = ny - 1;
            for (size_t x = 0; x + 1 < nx; x++) {
                size_t i1 = x + y * nx, i2 = (x + 1) + y * nx;
                for (size_t rx = 0; rx < subdiv_.x; rx++)
                    put(x * subdiv_.x + rx, y * subdiv_.y,
                        rx_between(i1, i2, i1, i2, {static_cast<double>(rx) / sx, 0.0 / sy}));
            }
            size_t il = (nx - 1) + (ny - 1) * nx;
            put((nx - 1) * subdiv_.x, (ny - 1) * subdiv_.y, rx_between(il, il, il, il, {0, 0}));
        } else {
            for (size_t x = 0; x + 1 < nx; x++) {
                for (size_t rx = 0; rx < subdiv_.x; rx++)
                    put(rx + x * subdiv_.x, 0,
                        rx_between(x, x + 1, x, x + 1, {static_cast<double>(rx) / sx, 0.0 / sy}));
            }
            put((nx - 1) * subdiv_.x, 0, rx_between(nx - 1, nx - 1, nx - 1, nx - 1, {0, 0}));
        }
    }

    supra_beamformer_config cfg_;
    S2 layout_{0, 0}, n_tx_{0, 0}, subdiv_{0, 0}, n_rx_{0, 0}, rx_aperture_{0, 0}, tx_aperture_{0, 0};
    size_t num_elements_ = 0;
    double c_mm_ = 0;
    D2 steer_{0, 0}, sector_{0, 0};
    std::vector<D3> centres_;
    std::vector<TxEvent> tx_;
};

}  // namespace
}  // namespace supra_amd

extern "C" int supra_geometry_sizes(const supra_beamformer_config* cfg, uint32_t* num_rx,
                                    uint32_t rx_layout[2], uint32_t* num_elements) {
    using namespace supra_amd;
    clear_error();
    if (!cfg || !num_rx || !rx_layout || !num_elements)
        return fail(SUPRA_ERR_INVALID_ARG, "geometry: null argument");
    Producer p(*cfg);
    int st = p.validate();
    if (st != SUPRA_OK) return st;
    p.layout_rx(num_rx, rx_layout);
    *num_elements = p.num_elements();
    return SUPRA_OK;
}

extern "C" int supra_geometry_build(const supra_beamformer_config* cfg, supra_scanline_rx* scanlines,
                                    float* depths, float* element_x, float* element_y) {
    using namespace supra_amd;
    clear_error();
    if (!cfg || !scanlines || !depths || !element_x || !element_y)
        return fail(SUPRA_ERR_INVALID_ARG, "geometry: null argument");
    Producer p(*cfg);
    int st = p.validate();
    if (st != SUPRA_OK) return st;
    p.build(scanlines, depths, element_x, element_y);
    return SUPRA_OK;
}
