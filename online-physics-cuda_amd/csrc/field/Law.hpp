// A law as the front end compiles it: fields with their contracts, params, terms (equations evaluated on one
// pre-state, with their max_dt, conservation groups and learned schemes), derives and nets.
//
// The types are member for member the reference's (src/field/Law.hpp:37-197), so a backend written against either
// compiles against the other: the engine-facing boundary is the Law a backend is handed. Creature blocks and `show`
// lines belong to the reference's demo; this front end reads `show` lines into `visuals` and passes over creature
// blocks (their agents live outside the lattice), so `creatures` stays empty.
#pragma once

#include "field/Grid.hpp"
#include "field/LawNode.hpp"
#include "field/nn/Net.hpp"

#include <memory>
#include <optional>
#include <string>
#include <string_view>
#include <vector>

namespace fieldengine::field {

// One source text; several compile together (a base law, an add-on), and errors name the source and line.
struct LawSource {
    std::string name;
    std::string text;
};

struct FieldDecl {
    std::string name;
    double init = 0.0;
    double lo = -1e30;
    double hi = 1e30;
    bool clamp = false;     // out-of-range values are clamped rather than reported
    bool is_static = false; // no term may write it
    bool derived = false;   // set by a derive, never integrated
    std::string unit;
};

struct Equation {
    int target = -1;
    std::shared_ptr<Node> rhs;
};

struct Term {
    std::string name;
    std::string where; // "source:line"
    std::vector<Equation> equations;
    std::vector<std::vector<int>> conserves; // groups of fields whose summed total the term keeps
    std::shared_ptr<Node> max_dt;            // one number, or null
    std::vector<Equation> learned;           // a learned scheme run under guard in place of `equations`
    std::shared_ptr<Node> learned_max_dt;
    std::vector<int> reads;
    std::vector<int> writes;
};

struct Derive {
    int target = -1;
    bool accumulate = false; // `+=`: adds to what the derives before it set
    std::string where;
    std::shared_ptr<Node> rhs;
    std::vector<int> reads;
};

// The demo's creatures (the reference's src/field/Law.hpp:82-150): declared so the Law has the reference's members.
inline constexpr int kLocalHp = 0;
inline constexpr int kLocalEaten = 1;
inline constexpr int kLocalAge = 2;
inline constexpr int kFirstState = 3;

struct CreatureFlow {
    int field = -1;
    bool eat = false;
    std::shared_ptr<Node> rate;
};

struct CreatureEvent {
    std::string name;
    int state = -1;
    bool above = true;
    double line = 0.0;
    double weight = 0.5;
    int sets = -1;
    double to = 0.0;
};

struct CreatureDef {
    std::string name;
    std::string where;
    std::string skin;
    std::vector<CreatureEvent> events;
    double radius = 0.35;
    double speed = 2.0;
    double hp = 40.0;
    double wander = 0.4;
    double sense = 0.5;
    std::string fights;
    int skill = 0;
    bool riled = false;
    std::vector<std::string> prey;
    double heed = 0.05;
    std::vector<std::string> locals;
    std::vector<double> init;
    std::vector<double> lo;
    std::vector<double> hi;
    std::vector<std::shared_ptr<Node>> rates;
    std::shared_ptr<Node> seek;
    std::shared_ptr<Node> hurt;
    std::shared_ptr<Node> breed;
    float colour[3] = {0.63f, 0.49f, 0.31f};
    std::vector<CreatureFlow> flows;
    std::vector<int> reads;
    std::shared_ptr<Node> plan;
    std::vector<int> observe;
    int horizon = 12;
    int samples = 128;
    double plan_dt = 0.25;
};

// `show FIELD colour R G B above LO full HI [alpha A]`: how the reference's renderer draws a field.
struct FieldVisual {
    int field = -1;
    float colour[3] = {1.0f, 1.0f, 1.0f};
    double above = 0.0;
    double full = 1.0;
    double alpha = 0.6;
};

struct Law {
    std::vector<FieldDecl> fields; // ids are positions
    std::vector<std::string> param_names;
    std::vector<double> param_values;
    std::vector<Term> terms;
    std::vector<Derive> derives;
    std::vector<CreatureDef> creatures;
    std::vector<FieldVisual> visuals;
    std::vector<NetDecl> nets;

    [[nodiscard]] int field_id(std::string_view name) const;
    [[nodiscard]] int net_id(std::string_view name) const;
    [[nodiscard]] int param_id(std::string_view name) const;
    [[nodiscard]] int creature_id(std::string_view name) const;
};

struct CompileResult {
    std::optional<Law> law;
    std::vector<std::string> errors; // "source:line: what is wrong", in the order found
};

CompileResult compile_law(const std::vector<LawSource>& sources);
std::optional<LawSource> read_law_file(const std::string& path);
// A copy whose expression trees are new nodes (shared subtrees stay shared within the copy).
Law clone_law(const Law& law);

} // namespace fieldengine::field
