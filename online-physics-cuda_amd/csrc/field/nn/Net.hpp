// A network declared by a law (`net NAME in 4 hidden 32 32 out 1 act tanh [half] [file PATH]`) and its weights.
//
// The members are the reference's (src/field/nn/Net.hpp:18-46) so a backend reads a net the same way under either
// header set: sizes in, hidden..., out; one activation on every hidden layer, the last layer linear; the output is
// net((x - in_shift) * in_scale) * out_scale; weights per layer l are W_l (sizes[l+1] rows of sizes[l], row-major)
// followed by b_l. `gelu` (the tanh form) is this build's addition to the activations.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace fieldengine::field {

enum class Activation { Tanh, Relu, Silu, Linear, Gelu };

const char* to_string(Activation a);
bool parse_activation(const std::string& text, Activation& out);

struct NetDecl {
    std::string name;
    std::string where; // "source:line" of its declaration
    std::string file;  // its weight file
    std::vector<int> sizes;
    Activation act = Activation::Tanh;
    bool half = false; // may run on the matrix cores (fp16 operands, fp32 accumulation)
    std::vector<float> in_shift;
    std::vector<float> in_scale;
    float out_scale = 1.0f;
    std::vector<float> weights;
    bool loaded = false; // read from its file (else initialised: an untrained net outputs exactly 0)

    [[nodiscard]] int inputs() const { return sizes.empty() ? 0 : sizes.front(); }
    [[nodiscard]] int outputs() const { return sizes.empty() ? 0 : sizes.back(); }
    [[nodiscard]] int layers() const { return sizes.size() < 2 ? 0 : static_cast<int>(sizes.size()) - 1; }
    [[nodiscard]] int widest() const;
    [[nodiscard]] std::size_t weight_offset(int layer) const; // where W_layer starts; b_layer follows it
    [[nodiscard]] std::size_t parameter_count() const;
};

// Untrained weights: hidden layers Glorot-uniform from a hash of the name and `seed` (no generator state: a law always
// builds the same net), the last layer zero so the net's output is exactly 0 until it is trained.
void init_net(NetDecl& net, std::uint64_t seed = 0);
// The reference's text format (sections `net`, `sizes`, `act`, `in_shift`, `in_scale`, `out_scale`, `weights`).
bool load_net(NetDecl& net, std::string& error);
bool save_net(const NetDecl& net, const std::string& path, std::string& error);

} // namespace fieldengine::field
