// Law nets declared `half` on the matrix cores (net_mfma.hip): a net's weights packed for the card once, and the
// forward pass of a batch of normalised fp16 input rows into fp32 outputs.
#pragma once

#include "field/Law.hpp"

#include <memory>

namespace fieldengine::field::hip {

constexpr int kHalfNetMaxLayers = 8;
constexpr int kHalfNetMaxWidth = 128; // inputs and every layer's width, padded to 16

struct HalfNetDesc {
    int layers = 0;
    int in_pad = 0;  // inputs padded to 16: the row stride of the input rows, in halves
    int outs = 0;    // the net's outputs: the row stride of the output rows, in floats
    int act = 0;     // Activation's order: tanh, relu, silu, linear, gelu
    float out_scale = 1.0f;
    int width[kHalfNetMaxLayers] = {};              // each layer's outputs padded to 16
    const void* w[kHalfNetMaxLayers] = {};          // W[width][previous width], fp16
    const float* b[kHalfNetMaxLayers] = {};         // b[width], fp32
};

struct HalfNetCard {
    HalfNetDesc desc;
    void* blob = nullptr;
    HalfNetCard() = default;
    HalfNetCard(const HalfNetCard&) = delete;
    HalfNetCard& operator=(const HalfNetCard&) = delete;
    ~HalfNetCard();
};

// Throws std::runtime_error when the net does not fit the path (more than kHalfNetMaxLayers layers, a width beyond
// kHalfNetMaxWidth) or the device copy fails.
std::unique_ptr<HalfNetCard> half_net_create(const NetDecl& net);
int half_net_in_pad(const NetDecl& net);
// New weights (same shape) into the card's copy, in place; synchronous.
void half_net_update(HalfNetCard& card, const NetDecl& net);
// out[row * outs + o] for n_rows rows of in_pad halves each, queued on the runtime's stream. 0 or a HIP status.
int half_net_run(const HalfNetCard& net, const void* rows, long long n_rows, float* out);

} // namespace fieldengine::field::hip
