// SUPRA's raw-data mock format: the metadata files that describe a recorded
// acquisition (receive geometry, sizes, sampling) and the int16 frames that
// go with them.
//
//   metadata, ".json": RxBeamformerParameters::writeMetaDataForMock /
//     readMetaDataForMockJson (src/SupraLib/Beamformer/RxBeamformerParameters.cpp:93-156, :311-450);
//   metadata, ".mock" (ASCII "rawDataMockMetadata v 1"): readMetaDataForMockAscii
//     (:187-309) with the scanline record of operator>> (USImageProperties.cpp:59-79);
//   frames: consecutive int16 [tx][channel][sample] blocks of one file, read
//     frame by frame by UltrasoundInterfaceRawDataMock (:49-92, :187-204).
//
// Host code only (no device memory): the caller uploads frames itself.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "status.hpp"
#include "supra_bf.h"

using namespace supra_amd;

namespace {

// ---- a small JSON reader (objects, arrays, numbers, strings, literals) -----

struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    double num = 0;
    bool b = false;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;

    const Json& operator[](const std::string& k) const {
        static const Json null_value;
        auto it = obj.find(k);
        return (kind == Object && it != obj.end()) ? it->second : null_value;
    }
    const Json& operator[](size_t i) const {
        static const Json null_value;
        return (kind == Array && i < arr.size()) ? arr[i] : null_value;
    }
    // jsoncpp's asDouble / asLargestUInt of a missing member are 0
    double as_double() const { return kind == Number ? num : (kind == Bool ? (b ? 1.0 : 0.0) : 0.0); }
    uint64_t as_uint() const { return kind == Number && num > 0 ? static_cast<uint64_t>(num) : 0; }
};

class JsonParser {
public:
    explicit JsonParser(const std::string& text) : s_(text) {}
    bool parse(Json& out, std::string& err) {
        try {
            value(out);
            ws();
            if (p_ != s_.size()) throw std::runtime_error("trailing characters");
            return true;
        } catch (const std::exception& e) {
            err = std::string(e.what()) + " at offset " + std::to_string(p_);
            return false;
        }
    }

private:
    const std::string& s_;
    size_t p_ = 0;
    void ws() {
        while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\n' || s_[p_] == '\r' || s_[p_] == '\t')) p_++;
    }
    char peek() {
        ws();
        if (p_ >= s_.size()) throw std::runtime_error("unexpected end");
        return s_[p_];
    }
    void expect(char c) {
        if (peek() != c) throw std::runtime_error(std::string("expected '") + c + "'");
        p_++;
    }
    void value(Json& v) {
        const char c = peek();
        if (c == '{') {
            v.kind = Json::Object;
            p_++;
            if (peek() == '}') { p_++; return; }
            for (;;) {
                Json key;
                string_value(key);
                expect(':');
                value(v.obj[key.str]);
                if (peek() == ',') { p_++; continue; }
                expect('}');
                return;
            }
        } else if (c == '[') {
            v.kind = Json::Array;
            p_++;
            if (peek() == ']') { p_++; return; }
            for (;;) {
                v.arr.emplace_back();
                value(v.arr.back());
                if (peek() == ',') { p_++; continue; }
                expect(']');
                return;
            }
        } else if (c == '"') {
            string_value(v);
        } else if (s_.compare(p_, 4, "true") == 0) {
            v.kind = Json::Bool, v.b = true, p_ += 4;
        } else if (s_.compare(p_, 5, "false") == 0) {
            v.kind = Json::Bool, v.b = false, p_ += 5;
        } else if (s_.compare(p_, 4, "null") == 0) {
            v.kind = Json::Null, p_ += 4;
        } else {
            const char* start = s_.c_str() + p_;
            char* end = nullptr;
            errno = 0;
            v.num = std::strtod(start, &end);
            if (end == start) throw std::runtime_error("bad value");
            v.kind = Json::Number;
            p_ += static_cast<size_t>(end - start);
        }
    }
    void string_value(Json& v) {
        expect('"');
        v.kind = Json::String;
        while (p_ < s_.size() && s_[p_] != '"') {
            if (s_[p_] == '\\' && p_ + 1 < s_.size()) p_++;
            v.str.push_back(s_[p_++]);
        }
        if (p_ >= s_.size()) throw std::runtime_error("unterminated string");
        p_++;
    }
};

bool ends_with(const std::string& s, const char* suffix) {
    const size_t n = std::strlen(suffix);
    return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

// The parsed metadata of either format.
struct MockFile {
    supra_mock_meta meta;
    std::vector<supra_scanline_rx> scanlines;
    std::vector<float> depths, ex, ey;
};

int read_ascii(const std::string& path, MockFile& m) {
    std::ifstream f(path);
    if (!f.good()) return fail(SUPRA_ERR_INVALID_ARG, "mock: cannot open %s", path.c_str());
    std::string tag, v;
    int version = 0;
    f >> tag >> v >> version;   // "rawDataMockMetadata v 1"
    supra_mock_meta& h = m.meta;
    std::memset(&h, 0, sizeof(h));
    uint64_t n_el, lx, ly, n_ch, n_s, n_tx, sx, sy, n_d;
    f >> n_el >> lx >> ly >> n_ch >> n_s >> n_tx >> sx >> sy >> h.depth >> h.sampling_frequency_hz >> n_d >>
        h.speed_of_sound_mm_per_s;
    if (!f || tag != "rawDataMockMetadata")
        return fail(SUPRA_ERR_INVALID_ARG, "mock: %s is not a rawDataMockMetadata file", path.c_str());
    if (n_el > (1u << 24) || n_ch > (1u << 24) || n_s > (1u << 30) || n_tx > (1u << 24) || sx * sy > (1u << 24) ||
        n_d > (1u << 24))
        return fail(SUPRA_ERR_INVALID_ARG, "mock: sizes out of range in %s", path.c_str());
    h.num_elements = static_cast<uint32_t>(n_el);
    h.element_layout[0] = static_cast<uint32_t>(lx);
    h.element_layout[1] = static_cast<uint32_t>(ly);
    h.num_channels = static_cast<uint32_t>(n_ch);
    h.num_samples = static_cast<uint32_t>(n_s);
    h.num_tx = static_cast<uint32_t>(n_tx);
    h.rx_scanline_layout[0] = static_cast<uint32_t>(sx);
    h.rx_scanline_layout[1] = static_cast<uint32_t>(sy);
    h.num_depths = static_cast<uint32_t>(n_d);
    m.scanlines.assign(sx * sy, supra_scanline_rx{});
    // scanlines y-outer, x-inner, stored in that (x fastest) order
    for (supra_scanline_rx& s : m.scanlines) {
        f >> s.position[0] >> s.position[1] >> s.position[2] >> s.direction[0] >> s.direction[1] >> s.direction[2] >>
            s.max_element_distance[0] >> s.max_element_distance[1];
        for (int k = 0; k < 4; k++) {
            f >> s.tx[k].first_active_element[0] >> s.tx[k].first_active_element[1] >>
                s.tx[k].last_active_element[0] >> s.tx[k].last_active_element[1] >> s.tx[k].tx_scanline_idx >>
                s.tx[k].initial_delay >> s.tx_weights[k];
        }
    }
    m.depths.resize(n_d);
    for (float& d : m.depths) f >> d;
    m.ex.resize(n_el);
    m.ey.resize(n_el);
    for (float& x : m.ex) f >> x;
    for (float& y : m.ey) f >> y;
    if (!f) return fail(SUPRA_ERR_INVALID_ARG, "mock: %s ends early", path.c_str());
    return SUPRA_OK;
}

int read_json(const std::string& path, MockFile& m) {
    std::ifstream f(path, std::ios::binary);
    if (!f.good()) return fail(SUPRA_ERR_INVALID_ARG, "mock: cannot open %s", path.c_str());
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    Json doc;
    std::string err;
    if (!JsonParser(text).parse(doc, err))
        return fail(SUPRA_ERR_INVALID_ARG, "mock: cannot parse %s: %s", path.c_str(), err.c_str());
    supra_mock_meta& h = m.meta;
    std::memset(&h, 0, sizeof(h));
    const uint64_t n_el = doc["numElements"].as_uint(), n_d = doc["rxNumDepths"].as_uint();
    const uint64_t sx = doc["scanlineLayout"]["x"].as_uint(), sy = doc["scanlineLayout"]["y"].as_uint();
    if (n_el > (1u << 24) || sx * sy > (1u << 24) || n_d > (1u << 24) || doc["numReceivedChannels"].as_uint() > (1u << 24) ||
        doc["numSamples"].as_uint() > (1u << 30) || doc["numTxScanlines"].as_uint() > (1u << 24))
        return fail(SUPRA_ERR_INVALID_ARG, "mock: sizes out of range in %s", path.c_str());
    h.num_elements = static_cast<uint32_t>(n_el);
    h.num_channels = static_cast<uint32_t>(doc["numReceivedChannels"].as_uint());
    h.num_samples = static_cast<uint32_t>(doc["numSamples"].as_uint());
    h.num_tx = static_cast<uint32_t>(doc["numTxScanlines"].as_uint());
    h.rx_scanline_layout[0] = static_cast<uint32_t>(sx);
    h.rx_scanline_layout[1] = static_cast<uint32_t>(sy);
    h.element_layout[0] = static_cast<uint32_t>(doc["elementLayout"]["x"].as_uint());
    h.element_layout[1] = static_cast<uint32_t>(doc["elementLayout"]["y"].as_uint());
    h.depth = doc["depth"].as_double();
    h.sampling_frequency_hz = doc["samplingFrequency"].as_double();
    h.num_depths = static_cast<uint32_t>(n_d);
    h.speed_of_sound_mm_per_s = doc["speedOfSoundMMperS"].as_double();
    m.scanlines.assign(sx * sy, supra_scanline_rx{});
    for (size_t i = 0; i < m.scanlines.size(); i++) {
        const Json& p = doc["rxScanlines"][i];
        supra_scanline_rx& s = m.scanlines[i];
        s.position[0] = p["position"]["x"].as_double();
        s.position[1] = p["position"]["y"].as_double();
        s.position[2] = p["position"]["z"].as_double();
        s.direction[0] = p["direction"]["x"].as_double();
        s.direction[1] = p["direction"]["y"].as_double();
        s.direction[2] = p["direction"]["z"].as_double();
        s.max_element_distance[0] = p["maxElementDistance"]["x"].as_double();
        s.max_element_distance[1] = p["maxElementDistance"]["y"].as_double();
        for (int k = 0; k < 4; k++) {
            const Json& t = p["txParameters"][k];
            s.tx[k].first_active_element[0] = static_cast<uint16_t>(t["firstActiveElementIndex"]["x"].as_uint());
            s.tx[k].first_active_element[1] = static_cast<uint16_t>(t["firstActiveElementIndex"]["y"].as_uint());
            s.tx[k].last_active_element[0] = static_cast<uint16_t>(t["lastActiveElementIndex"]["x"].as_uint());
            s.tx[k].last_active_element[1] = static_cast<uint16_t>(t["lastActiveElementIndex"]["y"].as_uint());
            s.tx[k].tx_scanline_idx = static_cast<uint16_t>(t["txScanlineIdx"].as_uint());
            s.tx[k].initial_delay = t["initialDelay"].as_double();
            s.tx_weights[k] = t["txWeights"].as_double();
        }
    }
    // not stored in the file: depth * idx / (N - 1), as the reader deduces it (:408-412)
    m.depths.resize(n_d);
    for (size_t i = 0; i < n_d; i++)
        m.depths[i] = static_cast<float>(h.depth * static_cast<double>(i) / static_cast<double>(n_d - 1));
    m.ex.resize(n_el);
    m.ey.resize(n_el);
    for (size_t i = 0; i < n_el; i++) {
        m.ex[i] = static_cast<float>(doc["rxElementPosition"][i]["x"].as_double());
        m.ey[i] = static_cast<float>(doc["rxElementPosition"][i]["y"].as_double());
    }
    return SUPRA_OK;
}

int read_any(const char* path, MockFile& m) {
    if (!path) return fail(SUPRA_ERR_INVALID_ARG, "mock: null path");
    const std::string p(path);
    if (ends_with(p, ".mock")) return read_ascii(p, m);
    if (ends_with(p, ".json")) return read_json(p, m);
    return fail(SUPRA_ERR_INVALID_ARG, "mock: unknown suffix of %s (.mock or .json)", path);
}

uint64_t frame_elements(const supra_mock_meta& m) {
    return static_cast<uint64_t>(m.num_channels) * m.num_samples * m.num_tx;
}

std::string num17(double v) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.17g", v);
    return buf;
}

}  // namespace

extern "C" int supra_mock_read_meta(const char* path, supra_mock_meta* meta) {
    clear_error();
    if (!meta) return fail(SUPRA_ERR_INVALID_ARG, "mock: null argument");
    MockFile m;
    int st = read_any(path, m);
    if (st == SUPRA_OK) *meta = m.meta;
    return st;
}

extern "C" int supra_mock_read_geometry(const char* path, supra_scanline_rx* scanlines, float* depths,
                                        float* element_x, float* element_y) {
    clear_error();
    if (!scanlines || !depths || !element_x || !element_y) return fail(SUPRA_ERR_INVALID_ARG, "mock: null argument");
    MockFile m;
    int st = read_any(path, m);
    if (st != SUPRA_OK) return st;
    std::copy(m.scanlines.begin(), m.scanlines.end(), scanlines);
    std::copy(m.depths.begin(), m.depths.end(), depths);
    std::copy(m.ex.begin(), m.ex.end(), element_x);
    std::copy(m.ey.begin(), m.ey.end(), element_y);
    return SUPRA_OK;
}

extern "C" int supra_mock_write_json(const char* path, const supra_mock_meta* meta, const supra_rx_geometry* g) {
    clear_error();
    if (!path || !meta || !g || !g->scanlines || !g->element_x || !g->element_y)
        return fail(SUPRA_ERR_INVALID_ARG, "mock: null argument");
    if (g->num_rx_scanlines != static_cast<uint64_t>(meta->rx_scanline_layout[0]) * meta->rx_scanline_layout[1] ||
        g->num_elements != meta->num_elements)
        return fail(SUPRA_ERR_INVALID_ARG, "mock: geometry and metadata sizes differ");
    std::ofstream f(path);
    if (!f.good()) return fail(SUPRA_ERR_INVALID_ARG, "mock: cannot create %s", path);
    const supra_mock_meta& h = *meta;
    f << "{\n";
    f << "    \"numElements\" : " << h.num_elements << ",\n";
    f << "    \"numReceivedChannels\" : " << h.num_channels << ",\n";
    f << "    \"numSamples\" : " << h.num_samples << ",\n";
    f << "    \"numTxScanlines\" : " << h.num_tx << ",\n";
    f << "    \"scanlineLayout\" : { \"x\" : " << h.rx_scanline_layout[0] << ", \"y\" : " << h.rx_scanline_layout[1]
      << " },\n";
    f << "    \"elementLayout\" : { \"x\" : " << h.element_layout[0] << ", \"y\" : " << h.element_layout[1] << " },\n";
    f << "    \"depth\" : " << num17(h.depth) << ",\n";
    f << "    \"samplingFrequency\" : " << num17(h.sampling_frequency_hz) << ",\n";
    f << "    \"rxNumDepths\" : " << h.num_depths << ",\n";
    f << "    \"speedOfSoundMMperS\" : " << num17(h.speed_of_sound_mm_per_s) << ",\n";
    f << "    \"rxScanlines\" : [\n";
    for (uint32_t i = 0; i < g->num_rx_scanlines; i++) {
        const supra_scanline_rx& s = g->scanlines[i];
        f << "        {\n";
        f << "            \"position\" : { \"x\" : " << num17(s.position[0]) << ", \"y\" : " << num17(s.position[1])
          << ", \"z\" : " << num17(s.position[2]) << " },\n";
        f << "            \"direction\" : { \"x\" : " << num17(s.direction[0]) << ", \"y\" : "
          << num17(s.direction[1]) << ", \"z\" : " << num17(s.direction[2]) << " },\n";
        f << "            \"maxElementDistance\" : { \"x\" : " << num17(s.max_element_distance[0]) << ", \"y\" : "
          << num17(s.max_element_distance[1]) << " },\n";
        f << "            \"txParameters\" : [\n";
        for (int k = 0; k < 4; k++) {
            const supra_tx_params& t = s.tx[k];
            f << "                { \"firstActiveElementIndex\" : { \"x\" : " << t.first_active_element[0]
              << ", \"y\" : " << t.first_active_element[1] << " }, \"lastActiveElementIndex\" : { \"x\" : "
              << t.last_active_element[0] << ", \"y\" : " << t.last_active_element[1]
              << " }, \"txScanlineIdx\" : " << t.tx_scanline_idx << ", \"initialDelay\" : "
              << num17(t.initial_delay) << ", \"txWeights\" : " << num17(s.tx_weights[k]) << " }"
              << (k < 3 ? "," : "") << "\n";
        }
        f << "            ]\n";
        f << "        }" << (i + 1 < g->num_rx_scanlines ? "," : "") << "\n";
    }
    f << "    ],\n";
    f << "    \"rxElementPosition\" : [\n";
    for (uint32_t i = 0; i < g->num_elements; i++)
        f << "        { \"x\" : " << num17(g->element_x[i]) << ", \"y\" : " << num17(g->element_y[i]) << " }"
          << (i + 1 < g->num_elements ? "," : "") << "\n";
    f << "    ]\n}\n";
    if (!f.good()) return fail(SUPRA_ERR_INVALID_ARG, "mock: write to %s failed", path);
    return SUPRA_OK;
}

extern "C" int supra_mock_write_ascii(const char* path, const supra_mock_meta* meta, const supra_rx_geometry* g) {
    clear_error();
    if (!path || !meta || !g || !g->scanlines || !g->depths || !g->element_x || !g->element_y)
        return fail(SUPRA_ERR_INVALID_ARG, "mock: null argument");
    if (g->num_rx_scanlines != static_cast<uint64_t>(meta->rx_scanline_layout[0]) * meta->rx_scanline_layout[1] ||
        g->num_elements != meta->num_elements || g->num_depths != meta->num_depths)
        return fail(SUPRA_ERR_INVALID_ARG, "mock: geometry and metadata sizes differ");
    std::ofstream f(path);
    if (!f.good()) return fail(SUPRA_ERR_INVALID_ARG, "mock: cannot create %s", path);
    const supra_mock_meta& h = *meta;
    f << "rawDataMockMetadata v 1\n";
    f << h.num_elements << " " << h.element_layout[0] << " " << h.element_layout[1] << " " << h.num_channels << " "
      << h.num_samples << " " << h.num_tx << " " << h.rx_scanline_layout[0] << " " << h.rx_scanline_layout[1] << " "
      << std::setprecision(17) << h.depth << " " << h.sampling_frequency_hz << " " << h.num_depths << " "
      << h.speed_of_sound_mm_per_s << "\n";
    // operator<< of ScanlineRxParameters3D (USImageProperties.cpp:36-57), here
    // with 17 significant digits so that doubles survive the round trip
    for (uint32_t i = 0; i < g->num_rx_scanlines; i++) {
        const supra_scanline_rx& s = g->scanlines[i];
        f << s.position[0] << " " << s.position[1] << " " << s.position[2] << " " << s.direction[0] << " "
          << s.direction[1] << " " << s.direction[2] << " " << s.max_element_distance[0] << " "
          << s.max_element_distance[1] << " ";
        for (int k = 0; k < 4; k++)
            f << s.tx[k].first_active_element[0] << " " << s.tx[k].first_active_element[1] << " "
              << s.tx[k].last_active_element[0] << " " << s.tx[k].last_active_element[1] << " "
              << s.tx[k].tx_scanline_idx << " " << s.tx[k].initial_delay << " " << s.tx_weights[k] << " ";
        f << "\n";
    }
    f << std::setprecision(9);
    for (uint32_t i = 0; i < g->num_depths; i++) f << g->depths[i] << (i + 1 < g->num_depths ? " " : "\n");
    for (uint32_t i = 0; i < g->num_elements; i++) f << g->element_x[i] << (i + 1 < g->num_elements ? " " : "\n");
    for (uint32_t i = 0; i < g->num_elements; i++) f << g->element_y[i] << (i + 1 < g->num_elements ? " " : "\n");
    if (!f.good()) return fail(SUPRA_ERR_INVALID_ARG, "mock: write to %s failed", path);
    return SUPRA_OK;
}

extern "C" int supra_mock_count_frames(const char* raw_path, const supra_mock_meta* meta, uint64_t* num_frames) {
    clear_error();
    if (!raw_path || !meta || !num_frames) return fail(SUPRA_ERR_INVALID_ARG, "mock: null argument");
    std::ifstream f(raw_path, std::ios::binary | std::ios::ate);
    if (!f.good()) return fail(SUPRA_ERR_INVALID_ARG, "mock: cannot open %s", raw_path);
    const uint64_t bytes = static_cast<uint64_t>(f.tellg());
    const uint64_t per = frame_elements(*meta) * sizeof(int16_t);
    if (per == 0) return fail(SUPRA_ERR_INVALID_ARG, "mock: empty frame size");
    *num_frames = bytes / per;   // the reference's sequence length (filesize / frame bytes)
    return SUPRA_OK;
}

extern "C" int supra_mock_read_frames(const char* raw_path, const supra_mock_meta* meta, uint64_t first_frame,
                                      uint32_t num_frames, int16_t* out) {
    clear_error();
    if (!raw_path || !meta || !out) return fail(SUPRA_ERR_INVALID_ARG, "mock: null argument");
    uint64_t have = 0;
    int st = supra_mock_count_frames(raw_path, meta, &have);
    if (st != SUPRA_OK) return st;
    if (first_frame + num_frames > have)
        return fail(SUPRA_ERR_INVALID_ARG, "mock: frames %llu..%llu requested, %s holds %llu",
                    static_cast<unsigned long long>(first_frame),
                    static_cast<unsigned long long>(first_frame + num_frames),
                    raw_path, static_cast<unsigned long long>(have));
    std::ifstream f(raw_path, std::ios::binary);
    const uint64_t per = frame_elements(*meta);
    f.seekg(static_cast<std::streamoff>(first_frame * per * sizeof(int16_t)));
    f.read(reinterpret_cast<char*>(out), static_cast<std::streamsize>(per * num_frames * sizeof(int16_t)));
    if (!f) return fail(SUPRA_ERR_INVALID_ARG, "mock: short read from %s", raw_path);
    return SUPRA_OK;
}
