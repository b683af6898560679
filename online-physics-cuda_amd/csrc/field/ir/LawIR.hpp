// The law as a program: one SSA graph for the whole law and one Program per term, learned scheme and derive.
//
// The same shape as the reference's IR (src/field/ir/LawIR.hpp:29-73): instructions with a shape (one number or one per
// cell), a read radius and a cost; values found twice kept once (structural hashing across all programs); constants
// folded as the CPU's scalar path would compute them; stages (what is computed over the lattice before the gather
// kernel reads it: global ops, nets, tridiagonal solves and whatever costs more inlined than stored); and the in-place
// decision (a term whose equations read no target off its own cell stores in place, else it ping-pongs). The JSON is
// the reference's format (src/field/ir/LawIR.cpp:528-593), so the IR of either compiler can be read by the same tools.
#pragma once

#include "field/Law.hpp"

#include <cstdint>
#include <string>
#include <vector>

namespace fieldengine::field::ir {

enum class Shape : std::uint8_t { Scalar, Cell };

struct Inst {
    Op op = Op::Const;
    Shape shape = Shape::Scalar;
    double number = 0.0;
    int index = -1;
    int iterations = 0;
    int offset[3] = {0, 0, 0};
    std::vector<int> args;
    int radius = 0;
    bool global = false;
    bool net = false;
    int uses = 0;
    int cost = 1;
};

struct Equation {
    int target = -1;
    int value = -1;
};

struct Program {
    std::string name;
    std::vector<Equation> equations;
    std::vector<int> stages;
    std::vector<int> limit_stages;
    bool in_place = false;
    int max_dt = -1;
    std::vector<int> reads;
    std::vector<int> writes;
    std::vector<std::vector<int>> conserves;
    bool accumulate = false;
};

struct Module {
    std::vector<Inst> insts;
    std::vector<Program> terms;
    std::vector<Program> learned;
    std::vector<Program> derives;
    std::vector<char> materialized;
    int folded = 0;
    int shared = 0;
    [[nodiscard]] bool is_stage(int id) const;
};

// A stage argument is stored, not inlined, when one point of its reader would cost more than this.
inline constexpr int kInlineBudget = 48;

Module lower(const Law& law);
const char* op_name(Op op);
std::string to_json(const Module& m, const Law& law);
std::string describe(const Module& m, const Law& law);

} // namespace fieldengine::field::ir
