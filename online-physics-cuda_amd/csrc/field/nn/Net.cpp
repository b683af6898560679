// Net declarations' weights: sizes, untrained initialisation and the text weight file. See Net.hpp.
#include "field/nn/Net.hpp"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <numeric>
#include <utility>

namespace fieldengine::field {

namespace {

constexpr std::array<std::pair<Activation, const char*>, 5> kActivationNames{{
    {Activation::Tanh, "tanh"},
    {Activation::Relu, "relu"},
    {Activation::Silu, "silu"},
    {Activation::Linear, "linear"},
    {Activation::Gelu, "gelu"},
}};

// splitmix64's finaliser: a well-mixed 64-bit value from a counter.
std::uint64_t scramble(std::uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// A number token of the weight file: a decimal or exponent form, read whole; no words (inf, nan) and no hex.
bool read_number(const std::string& word, double& out)
{
    if (word.empty() || std::isalpha(static_cast<unsigned char>(word[0])) != 0) {
        return false;
    }
    char* end = nullptr;
    out = std::strtod(word.c_str(), &end);
    return end != word.c_str();
}

} // namespace

const char* to_string(Activation a)
{
    for (const auto& [act, name] : kActivationNames) {
        if (act == a) {
            return name;
        }
    }
    return "?";
}

bool parse_activation(const std::string& text, Activation& out)
{
    const auto it = std::find_if(kActivationNames.begin(), kActivationNames.end(),
        [&](const auto& entry) { return text == entry.second; });
    if (it == kActivationNames.end()) {
        return false;
    }
    out = it->first;
    return true;
}

int NetDecl::widest() const
{
    return sizes.empty() ? 0 : *std::max_element(sizes.begin(), sizes.end());
}

std::size_t NetDecl::weight_offset(int layer) const
{
    // Layer l holds sizes[l+1] rows of (sizes[l] weights + 1 bias).
    std::size_t at = 0;
    for (int l = 0; l < layer; ++l) {
        const auto in = static_cast<std::size_t>(sizes[static_cast<std::size_t>(l)]);
        const auto out = static_cast<std::size_t>(sizes[static_cast<std::size_t>(l) + 1]);
        at += (in + 1) * out;
    }
    return at;
}

std::size_t NetDecl::parameter_count() const
{
    return weight_offset(layers());
}

void init_net(NetDecl& net, std::uint64_t seed)
{
    net.weights.assign(net.parameter_count(), 0.0f);
    std::uint64_t key = 0xcbf29ce484222325ull ^ scramble(seed);
    for (const char c : net.name) {
        key = (key ^ static_cast<unsigned char>(c)) * 0x100000001b3ull;
    }
    for (int l = 0; l + 1 < net.layers(); ++l) { // the last layer stays zero
        const int fan_in = net.sizes[static_cast<std::size_t>(l)];
        const int fan_out = net.sizes[static_cast<std::size_t>(l) + 1];
        const double bound = std::sqrt(6.0 / static_cast<double>(fan_in + fan_out));
        float* w = net.weights.data() + net.weight_offset(l);
        for (int k = 0; k < fan_in * fan_out; ++k) {
            const std::uint64_t r = scramble(key + (static_cast<std::uint64_t>(l) << 40) + static_cast<std::uint64_t>(k));
            const double u = static_cast<double>(r >> 11) * 0x1.0p-53; // [0, 1)
            w[k] = static_cast<float>((2.0 * u - 1.0) * bound);
        }
    }
    net.in_shift.assign(static_cast<std::size_t>(net.inputs()), 0.0f);
    net.in_scale.assign(static_cast<std::size_t>(net.inputs()), 1.0f);
    net.out_scale = 1.0f;
    net.loaded = false;
}

bool load_net(NetDecl& net, std::string& error)
{
    std::ifstream file(net.file);
    if (!file) {
        error = "no weights at " + net.file;
        return false;
    }
    // Numbers go to the list of the last section word read; `net`, `act` and `out_scale` take one word each.
    std::vector<double> sizes, shift, scale, weights;
    std::vector<double>* into = nullptr;
    float out_scale = 1.0f;
    Activation act = net.act;
    std::string word;
    while (file >> word) {
        if (word == "net") {
            file >> word;
            into = nullptr;
        } else if (word == "act") {
            file >> word;
            if (!parse_activation(word, act)) {
                error = net.file + ": unknown activation '" + word + "'";
                return false;
            }
            into = nullptr;
        } else if (word == "out_scale") {
            file >> out_scale; // read as a float (a double rounded again could differ)
            into = nullptr;
        } else if (word == "sizes" || word == "in_shift" || word == "in_scale" || word == "weights") {
            into = word == "sizes" ? &sizes : word == "in_shift" ? &shift : word == "in_scale" ? &scale : &weights;
        } else {
            double v = 0.0;
            if (!read_number(word, v)) {
                error = net.file + ": unexpected '" + word + "'";
                return false;
            }
            if (into == nullptr) {
                error = net.file + ": a number outside any section";
                return false;
            }
            into->push_back(v);
        }
    }
    const bool same_sizes = sizes.size() == net.sizes.size() &&
        std::equal(sizes.begin(), sizes.end(), net.sizes.begin(), [](double a, int b) { return static_cast<int>(a) == b; });
    if (!same_sizes) {
        error = net.file + ": its sizes are not the ones the law declares for '" + net.name + "'";
        return false;
    }
    if (act != net.act) {
        error = net.file + ": trained with another activation than the law declares";
        return false;
    }
    if (weights.size() != net.parameter_count()) {
        error = net.file + ": " + std::to_string(weights.size()) + " weights, the net has " +
            std::to_string(net.parameter_count());
        return false;
    }
    const auto as_floats = [](const std::vector<double>& v, std::size_t n, float pad) {
        std::vector<float> out(n, pad);
        for (std::size_t i = 0; i < std::min(n, v.size()); ++i) {
            out[i] = static_cast<float>(v[i]);
        }
        return out;
    };
    net.weights = as_floats(weights, weights.size(), 0.0f);
    net.in_shift = as_floats(shift, static_cast<std::size_t>(net.inputs()), 0.0f);
    net.in_scale = as_floats(scale, static_cast<std::size_t>(net.inputs()), 1.0f);
    net.out_scale = out_scale;
    net.loaded = true;
    return true;
}

bool save_net(const NetDecl& net, const std::string& path, std::string& error)
{
    std::ofstream file(path);
    if (!file) {
        error = "cannot write " + path;
        return false;
    }
    const auto row = [&](const char* section, const auto& values) {
        file << section;
        for (const auto v : values) {
            file << ' ' << v;
        }
        file << '\n';
    };
    file << std::setprecision(9) << "net " << net.name << '\n';
    row("sizes", net.sizes);
    file << "act " << to_string(net.act) << '\n';
    row("in_shift", net.in_shift);
    row("in_scale", net.in_scale);
    file << "out_scale " << net.out_scale << "\nweights";
    for (std::size_t i = 0; i < net.weights.size(); ++i) {
        file << (i % 12 == 0 ? '\n' : ' ') << net.weights[i];
    }
    file << '\n';
    return static_cast<bool>(file);
}

} // namespace fieldengine::field
