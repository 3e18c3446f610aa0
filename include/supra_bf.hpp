/*
 * supra_bf.hpp — C++ engine shim over the C ABI of supra_bf.h, shaped like
 * SUPRA's RxBeamformerCuda (src/SupraLib/Beamformer/RxBeamformerCuda.h:28-80)
 * so that a SUPRA build can construct it where BeamformingNode constructs
 * RxBeamformerCuda (BeamformingNode.cpp:141-150).
 *
 * Header-only, no SUPRA headers needed: the geometry arrives through any
 * type with RxBeamformerParameters' accessors (RxBeamformerParameters.h:
 * 69-79) and scanlines with ScanlineRxParameters3D's members
 * (USImageProperties.h:43-95), so SUPRA's own classes plug in unchanged.
 * Buffers are device pointers; the call is asynchronous on `stream`
 * (a hipStream_t; nullptr: the null stream).
 */
#ifndef SUPRA_BF_HPP
#define SUPRA_BF_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "supra_bf.h"

namespace supra_amd {

// ScanlineRxParameters3D -> supra_scanline_rx, field by field.
template <class Scanline>
supra_scanline_rx to_c(const Scanline& s) {
    supra_scanline_rx c{};
    c.position[0] = s.position.x;
    c.position[1] = s.position.y;
    c.position[2] = s.position.z;
    c.direction[0] = s.direction.x;
    c.direction[1] = s.direction.y;
    c.direction[2] = s.direction.z;
    for (int k = 0; k < 4; k++) {
        c.tx_weights[k] = s.txWeights[k];
        const auto& t = s.txParameters[k];
        c.tx[k].first_active_element[0] = static_cast<uint16_t>(t.firstActiveElementIndex.x);
        c.tx[k].first_active_element[1] = static_cast<uint16_t>(t.firstActiveElementIndex.y);
        c.tx[k].last_active_element[0] = static_cast<uint16_t>(t.lastActiveElementIndex.x);
        c.tx[k].last_active_element[1] = static_cast<uint16_t>(t.lastActiveElementIndex.y);
        c.tx[k].tx_scanline_idx = static_cast<uint16_t>(t.txScanlineIdx);
        c.tx[k].initial_delay = t.initialDelay;
    }
    c.max_element_distance[0] = s.maxElementDistance.x;
    c.max_element_distance[1] = s.maxElementDistance.y;
    return c;
}

template <class T> struct element_type;
template <> struct element_type<int16_t> { static constexpr int32_t value = SUPRA_INT16; };
template <> struct element_type<float> { static constexpr int32_t value = SUPRA_FLOAT32; };

class RxBeamformerHip {
public:
    // RxBeamformerCuda::RxSampleBeamformer (RxBeamformerCuda.h:31-37)
    enum RxSampleBeamformer : int32_t {
        DelayAndSum = SUPRA_BF_DELAY_AND_SUM,
        DelayAndStdDev = SUPRA_BF_DELAY_AND_STD_DEV,
        CoherenceFactorDelayAndSum = SUPRA_BF_COHERENCE_FACTOR,
        TestSignal = SUPRA_BF_TEST_SIGNAL,
    };

    // RxBeamformerCuda(const RxBeamformerParameters&) (RxBeamformerCuda.cu:24-52).
    // Params: RxBeamformerParameters or a type with its accessors.  The
    // element layout is the transducer's (USTransducer::getElementLayout,
    // which RxBeamformerParameters does not hold); `depth` [mm] is only
    // needed by scan conversion.  Copies the geometry to the current device.
    template <class Params>
    RxBeamformerHip(const Params& p, uint32_t elementLayoutX, uint32_t elementLayoutY, double depth = 0.0) {
        const auto& lines = p.getRxScanlines();
        std::vector<supra_scanline_rx> scan;
        scan.reserve(lines.size());
        for (const auto& s : lines) scan.push_back(to_c(s));
        const auto layout = p.getRxScanlineLayout();
        supra_rx_geometry g{};
        g.num_rx_scanlines = static_cast<uint32_t>(p.getNumRxScanlines());
        g.rx_scanline_layout[0] = static_cast<uint32_t>(layout.x);
        g.rx_scanline_layout[1] = static_cast<uint32_t>(layout.y);
        g.num_depths = static_cast<uint32_t>(p.getRxNumDepths());
        g.num_elements = static_cast<uint32_t>(p.getRxElementXs().size());
        g.element_layout[0] = elementLayoutX;
        g.element_layout[1] = elementLayoutY;
        g.scanlines = scan.data();
        g.depths = p.getRxDepths().data();
        g.element_x = p.getRxElementXs().data();
        g.element_y = p.getRxElementYs().data();
        g.element_to_channel = p.getNonlinearElementToChannelMapping() ? p.getElementToChannelMap().data() : nullptr;
        g.speed_of_sound_mm_per_s = p.getSpeedOfSoundMMperS();
        g.depth = depth;
        if (scan.size() != g.num_rx_scanlines)
            throw std::invalid_argument("RxBeamformerHip: scanline count and layout disagree");
        if (supra_bf_create(&g, &m_ctx) != SUPRA_OK)
            throw std::runtime_error(std::string("supra_bf_create: ") + supra_last_error());
        m_numRx = g.num_rx_scanlines;
        m_numDepths = g.num_depths;
    }

    ~RxBeamformerHip() { supra_bf_destroy(m_ctx); }
    RxBeamformerHip(const RxBeamformerHip&) = delete;
    RxBeamformerHip& operator=(const RxBeamformerHip&) = delete;

    // RxBeamformerCuda::performRxBeamforming<ChannelDataType, ImageDataType>
    // (RxBeamformerCuda.h:43-52, RxBeamformerCuda.cu:539-663) on device
    // buffers: rf [numTx][numChannels][numSamples] (a USRawData payload),
    // out [numDepths][numRxScanlines] (a USImage payload).  Returns SUPRA_OK
    // or a SUPRA_ERR_* code (message: lastError()).
    template <typename ChannelDataType, typename ImageDataType>
    int performRxBeamforming(RxSampleBeamformer sampleBeamformer, const ChannelDataType* rf, uint32_t numTx,
                             uint32_t numChannels, uint32_t numSamples, double samplingFrequency, double fNumber,
                             double speedOfSoundMMperS, int32_t windowType, float windowParameter,
                             bool interpolateBetweenTransmits, int32_t additionalOffset, ImageDataType* out,
                             void* stream = nullptr) const {
        return supra_das_run(m_ctx, sampleBeamformer, rf, element_type<ChannelDataType>::value, numTx, numChannels,
                             numSamples, samplingFrequency, fNumber, speedOfSoundMMperS, windowType, windowParameter,
                             interpolateBetweenTransmits ? 1 : 0, additionalOffset, out,
                             element_type<ImageDataType>::value, stream);
    }

    uint32_t numRxScanlines() const { return m_numRx; }
    uint32_t numDepths() const { return m_numDepths; }
    supra_bf_ctx* handle() const { return m_ctx; }
    static const char* lastError() { return supra_last_error(); }

private:
    supra_bf_ctx* m_ctx = nullptr;
    uint32_t m_numRx = 0, m_numDepths = 0;
};

}  // namespace supra_amd

#endif /* SUPRA_BF_HPP */
