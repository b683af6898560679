// The law language: fields, params, terms, derives and nets as text, compiled to expression trees.
//
// The front end the HIP backend consumes. Its types and grammar follow the reference's (src/field/Law.hpp:37-211,
// src/field/LawNode.hpp:15-67, src/field/Law.cpp:47-410,1662-2439) so that existing law files compile unchanged: the
// same keywords, functions, precedence (unary minus binds looser than ^, ^ is right-associative), refusals and
// derived reads/writes. Creature blocks (the demo's agents) and `show` lines (the renderer's) parse and are ignored:
// nothing on the lattice reads them.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <optional>
#include <string>
#include <string_view>
#include <vector>

namespace fieldengine::field {

// The op set, in the reference's order (src/field/LawNode.hpp:15-67): the IR's JSON names them by op_name().
enum class Op {
    Const,
    Param,
    Cell, // h, the voxel edge in metres
    Rim,  // 1 on the outermost cells of the lattice
    Field,
    Neg,
    Add,
    Sub,
    Mul,
    Div,
    Pow,
    Min,
    Max,
    Abs,
    Sqrt,
    Exp,
    Log,
    Step, // 1 where x > 0
    Pos,  // max(x, 0)
    Clamp,
    Lap,
    Gx,
    Gy,
    Gz,
    Diffuse,
    Adv,
    Distance,
    Poisson,
    MaxOf,
    MinOf,
    SumOf,
    Fall,
    Spread,
    Topple,
    Shift,
    Local,
    Mlp,
    Flux,
    LDiffuse,
    Dt,
    Tridiag,
    ShiftT,
    GxT,
    GyT,
    GzT,
};

struct Node;
using NodePtr = std::shared_ptr<Node>;

struct Node {
    Op op = Op::Const;
    double number = 0.0;
    int index = -1;     // Field, Param, Local, a net, Tridiag's axis
    int iterations = 0; // Poisson
    std::vector<NodePtr> args;
};

struct FieldDecl {
    std::string name;
    double init = 0.0;
    double lo = -1e30;
    double hi = 1e30;
    bool clamp = false;
    bool is_static = false;
    bool derived = false;
    std::string unit;
};

enum class Activation { Tanh, Relu, Silu, Linear, Gelu };
const char* to_string(Activation a);

// A network a law uses (src/field/nn/Net.hpp:26-46): sizes in -> hidden... -> out, one activation on every hidden layer,
// the last layer linear. y = net((x - in_shift) * in_scale) * out_scale. Weights: per layer l, W_l (sizes[l+1] x
// sizes[l], row-major) then b_l. An untrained net (no file) has a zero last layer: its output is exactly 0.
struct NetDecl {
    std::string name;
    std::string where;
    std::string file;
    std::vector<int> sizes;
    Activation act = Activation::Tanh;
    bool half = false;
    bool loaded = false;
    std::vector<float> in_shift;
    std::vector<float> in_scale;
    float out_scale = 1.0f;
    std::vector<float> weights;
    [[nodiscard]] int inputs() const { return sizes.empty() ? 0 : sizes.front(); }
    [[nodiscard]] int outputs() const { return sizes.empty() ? 0 : sizes.back(); }
    [[nodiscard]] int layers() const { return sizes.size() < 2 ? 0 : static_cast<int>(sizes.size()) - 1; }
    [[nodiscard]] int widest() const;
    [[nodiscard]] std::size_t weight_offset(int layer) const;
    [[nodiscard]] std::size_t parameter_count() const;
};

bool parse_activation(const std::string& text, Activation& out);

struct Equation {
    int target = -1;
    NodePtr rhs;
};

struct Term {
    std::string name;
    std::string where;
    std::vector<Equation> equations;
    std::vector<std::vector<int>> conserves;
    NodePtr max_dt;
    std::vector<Equation> learned;
    NodePtr learned_max_dt;
    std::vector<int> reads;
    std::vector<int> writes;
};

struct Derive {
    int target = -1;
    bool accumulate = false;
    std::string where;
    NodePtr rhs;
    std::vector<int> reads;
};

struct Law {
    std::vector<FieldDecl> fields;
    std::vector<std::string> param_names;
    std::vector<double> param_values;
    std::vector<Term> terms;
    std::vector<Derive> derives;
    std::vector<NetDecl> nets;

    [[nodiscard]] int field_id(std::string_view name) const;
    [[nodiscard]] int net_id(std::string_view name) const;
    [[nodiscard]] int param_id(std::string_view name) const;
};

struct LawSource {
    std::string name;
    std::string text;
};

struct CompileResult {
    std::optional<Law> law;
    std::vector<std::string> errors;
};

CompileResult compile_law(const std::vector<LawSource>& sources);
std::optional<LawSource> read_law_file(const std::string& path);

// Whether a node is one number whatever the fields hold (src/field/LawNode.hpp:69-101).
bool always_scalar(const Node& node);

// A scalar expression in double, as the reference's scalar path computes it; nullopt when it reads a field.
std::optional<double> evaluate_scalar(const Node& node, const std::vector<double>& params, double h, double dt);

// Loads a net's weights from its text file (the reference's .net format, src/field/nn/Net.cpp:267-369), or
// initialises it untrained: a hash-Glorot body with a zero last layer (src/field/nn/Net.cpp:220-265).
void init_net(NetDecl& net);
bool load_net(NetDecl& net, std::string& why);
bool save_net(const NetDecl& net, const std::string& path, std::string& why);

} // namespace fieldengine::field
