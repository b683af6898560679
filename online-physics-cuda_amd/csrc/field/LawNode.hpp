// A parsed law expression: the tree the front end builds and the IR lowers.
//
// The op set and its order are the law language's, as the reference numbers them (src/field/LawNode.hpp:15-67): the
// IR's JSON names ops by op_name() and tools compare it across compilers. A node carries what the language says (the
// op, a constant, the index of a field, param, net or axis, Poisson's sweep count) and its arguments; the reference's
// node also carries its CPU interpreter's buffers, which nothing here uses.
#pragma once

#include <memory>
#include <vector>

namespace fieldengine::field {

enum class Op {
    Const,
    Param,
    Cell, // h: the cell edge in metres
    Rim,  // 1 on the outermost layer of cells, else 0
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
    Step, // 1 where the argument is > 0
    Pos,  // the argument where it is > 0, else 0
    Clamp,
    Lap,
    Gx,
    Gy,
    Gz,
    Diffuse,  // div(k grad F) in flux form
    Adv,      // -(u . grad F), upwinded
    Distance, // signed distance in metres to where the argument crosses 0.5
    Poisson,  // phi with lap(phi) = source, by warm-started Jacobi sweeps
    MaxOf,
    MinOf,
    SumOf,
    Fall,
    Spread,
    Topple,
    Shift,    // the argument in the cell (dx, dy, dz) away, clamped at the edges
    Local,    // a creature's own state (read at a point; never on the lattice)
    Mlp,      // a net's first output, cell by cell (index: the net)
    Flux,     // a field moved across faces by a net's antisymmetric, limited flux
    LDiffuse, // diffusion through faces weighted by a net over four cells
    Dt,       // the step the term is taken for
    Tridiag,  // the solution of a tridiagonal system along the lines of one axis (index: the axis)
    ShiftT,   // shift's transpose
    GxT,      // gx's, gy's, gz's transposes
    GyT,
    GzT,
};

struct Node {
    Op op = Op::Const;
    double number = 0.0; // Const
    int index = -1;      // Field, Param, Local, a net, Tridiag's axis
    int iterations = 0;  // Poisson
    std::vector<std::shared_ptr<Node>> args;
};

// Whether a node is one number whatever the fields hold. A reduction is one number; every op whose value varies over
// the lattice by construction is not; the rest are one number when all their arguments are (a leaf with no arguments,
// such as rim, counts as one).
inline bool always_scalar(const Node& node)
{
    switch (node.op) {
    case Op::Const: case Op::Param: case Op::Cell: case Op::MaxOf: case Op::MinOf: case Op::SumOf:
        return true;
    case Op::Field: case Op::Distance: case Op::Poisson: case Op::Fall: case Op::Spread: case Op::Topple:
    case Op::Flux: case Op::LDiffuse: case Op::Tridiag: case Op::ShiftT: case Op::GxT: case Op::GyT: case Op::GzT:
        return false;
    default:
        for (const auto& arg : node.args) {
            if (!always_scalar(*arg)) {
                return false;
            }
        }
        return true;
    }
}

} // namespace fieldengine::field
