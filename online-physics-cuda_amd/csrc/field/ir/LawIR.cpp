// Lowering a law to its IR, the stage plan, and the IR's JSON and listing. See LawIR.hpp.
//
// What must come out is fixed by the IR's JSON, which tools compare across compilers byte for byte: the instruction
// numbering (arguments first, left to right, program by program: each term's equations then its max_dt, its learned
// scheme likewise, then the derives), the folding rules, the shapes and read radii, and the cost model that decides
// which values are stored as stages (src/field/ir/LawIR.cpp:69-526 states them). How it is computed here:
//
//   - every per-op fact (name, shape rule, read radius, points read, global, net, foldable) is one row of a table;
//   - lowering walks each expression tree with an explicit stack (no recursion depth limit on long laws) and interns
//     instructions in a hash table keyed by their fields, so a value met twice anywhere in the law is one instruction;
//   - rewrites (folding, identities) are a rule function applied before interning;
//   - what a program needs (its stages, the fields it reads, how far off its cell it reads each field) is computed by
//     sweeping the instruction list once from the highest id down, which visits every instruction after all of its
//     users because arguments always have lower ids: reachability and the longest read distance become one pass each
//     instead of a walk of every path.
#include "field/ir/LawIR.hpp"

#include <algorithm>
#include <array>
#include <bit>
#include <cmath>
#include <cstdio>
#include <limits>
#include <set>
#include <sstream>
#include <unordered_map>

namespace fieldengine::field::ir {

namespace {

enum class ShapeRule : unsigned char { Scalar, Cell, Args };

struct OpFacts {
    Op op;
    const char* name;
    ShapeRule shape;
    int radius;       // cells off its own cell it reads its arguments at (-1: its shift offsets say)
    int points;       // points of each argument one of its points reads
    bool global;      // needs its argument over the whole lattice first
    bool net;         // runs a network
    bool foldable;    // plain arithmetic: all-constant arguments make a constant
};

constexpr int kOpCount = static_cast<int>(Op::GzT) + 1;

constexpr std::array<OpFacts, kOpCount> kFacts{{
    {Op::Const, "const", ShapeRule::Scalar, 0, 1, false, false, false},
    {Op::Param, "param", ShapeRule::Scalar, 0, 1, false, false, false},
    {Op::Cell, "h", ShapeRule::Scalar, 0, 1, false, false, false},
    {Op::Rim, "rim", ShapeRule::Cell, 0, 1, false, false, false},
    {Op::Field, "field", ShapeRule::Cell, 0, 1, false, false, false},
    {Op::Neg, "neg", ShapeRule::Args, 0, 1, false, false, true},
    {Op::Add, "add", ShapeRule::Args, 0, 1, false, false, true},
    {Op::Sub, "sub", ShapeRule::Args, 0, 1, false, false, true},
    {Op::Mul, "mul", ShapeRule::Args, 0, 1, false, false, true},
    {Op::Div, "div", ShapeRule::Args, 0, 1, false, false, true},
    {Op::Pow, "pow", ShapeRule::Args, 0, 1, false, false, true},
    {Op::Min, "min", ShapeRule::Args, 0, 1, false, false, true},
    {Op::Max, "max", ShapeRule::Args, 0, 1, false, false, true},
    {Op::Abs, "abs", ShapeRule::Args, 0, 1, false, false, true},
    {Op::Sqrt, "sqrt", ShapeRule::Args, 0, 1, false, false, true},
    {Op::Exp, "exp", ShapeRule::Args, 0, 1, false, false, true},
    {Op::Log, "log", ShapeRule::Args, 0, 1, false, false, true},
    {Op::Step, "step", ShapeRule::Args, 0, 1, false, false, true},
    {Op::Pos, "pos", ShapeRule::Args, 0, 1, false, false, true},
    {Op::Clamp, "clamp", ShapeRule::Args, 0, 1, false, false, true},
    {Op::Lap, "lap", ShapeRule::Cell, 1, 7, false, false, false},
    {Op::Gx, "gx", ShapeRule::Cell, 1, 2, false, false, false},
    {Op::Gy, "gy", ShapeRule::Cell, 1, 2, false, false, false},
    {Op::Gz, "gz", ShapeRule::Cell, 1, 2, false, false, false},
    {Op::Diffuse, "diffuse", ShapeRule::Cell, 1, 7, false, false, false},
    {Op::Adv, "adv", ShapeRule::Cell, 1, 7, false, false, false},
    {Op::Distance, "distance", ShapeRule::Cell, 0, 1, true, false, false},
    {Op::Poisson, "poisson", ShapeRule::Cell, 0, 1, true, false, false},
    {Op::MaxOf, "maxof", ShapeRule::Scalar, 0, 1, true, false, false},
    {Op::MinOf, "minof", ShapeRule::Scalar, 0, 1, true, false, false},
    {Op::SumOf, "sumof", ShapeRule::Scalar, 0, 1, true, false, false},
    {Op::Fall, "fall", ShapeRule::Cell, 1, 3, false, false, false},
    {Op::Spread, "spread", ShapeRule::Cell, 1, 5, false, false, false},
    {Op::Topple, "topple", ShapeRule::Cell, 1, 13, false, false, false},
    {Op::Shift, "shift", ShapeRule::Args, -1, 1, false, false, false},
    {Op::Local, "local", ShapeRule::Scalar, 0, 1, false, false, false},
    {Op::Mlp, "mlp", ShapeRule::Args, 0, 1, false, true, false},
    {Op::Flux, "flux", ShapeRule::Cell, 1, 1, false, true, false},
    {Op::LDiffuse, "ldiffuse", ShapeRule::Cell, 2, 1, false, true, false},
    {Op::Dt, "dt", ShapeRule::Scalar, 0, 1, false, false, false},
    {Op::Tridiag, "tridiag", ShapeRule::Cell, 0, 1, true, false, false},
    {Op::ShiftT, "shift_t", ShapeRule::Cell, -1, 1, false, false, false},
    {Op::GxT, "gx_t", ShapeRule::Cell, 1, 1, false, false, false},
    {Op::GyT, "gy_t", ShapeRule::Cell, 1, 1, false, false, false},
    {Op::GzT, "gz_t", ShapeRule::Cell, 1, 1, false, false, false},
}};

constexpr bool table_in_op_order()
{
    for (int i = 0; i < kOpCount; ++i) {
        if (static_cast<int>(kFacts[static_cast<std::size_t>(i)].op) != i) {
            return false;
        }
    }
    return true;
}
static_assert(table_in_op_order(), "one row per op, in the op's order");

const OpFacts& facts(Op op) { return kFacts[static_cast<std::size_t>(op)]; }

// The value the reference CPU's scalar path gives an arithmetic op of numbers (in double).
double fold(Op op, const std::vector<double>& v)
{
    const double a = v[0];
    const double b = v.size() > 1 ? v[1] : 0.0;
    switch (op) {
    case Op::Neg: return -a;
    case Op::Abs: return std::fabs(a);
    case Op::Sqrt: return std::sqrt(0.0 < a ? a : 0.0);
    case Op::Exp: return std::exp(80.0 < a ? 80.0 : a);
    case Op::Log: return std::log(a < 1e-30 ? 1e-30 : a);
    case Op::Step: return a > 0.0 ? 1.0 : 0.0;
    case Op::Pos: return a < 0.0 ? 0.0 : a;
    case Op::Add: return a + b;
    case Op::Sub: return a - b;
    case Op::Mul: return a * b;
    case Op::Div: return b == 0.0 ? 0.0 : a / b;
    case Op::Pow: return std::pow(a, b);
    case Op::Min: return b < a ? b : a;
    case Op::Max: return a < b ? b : a;
    case Op::Clamp: {
        const double top = b < v[2] ? v[2] : b;
        return a < b ? b : (top < a ? top : a);
    }
    default: return 0.0;
    }
}

// An instruction's identity: two instructions with equal keys are one value.
struct Key {
    int op;
    std::uint64_t number;
    int index;
    int iterations;
    std::array<int, 3> offset;
    std::vector<int> args;
    bool operator==(const Key&) const = default;
};

struct KeyHash {
    std::size_t operator()(const Key& k) const
    {
        std::uint64_t h = 0x9e3779b97f4a7c15ull;
        const auto mix = [&](std::uint64_t v) { h = (h ^ v) * 0x100000001b3ull; h ^= h >> 29; };
        mix(static_cast<std::uint64_t>(k.op));
        mix(k.number);
        mix(static_cast<std::uint64_t>(static_cast<std::uint32_t>(k.index)));
        mix(static_cast<std::uint64_t>(k.iterations));
        for (const int o : k.offset) {
            mix(static_cast<std::uint64_t>(static_cast<std::uint32_t>(o)));
        }
        for (const int a : k.args) {
            mix(static_cast<std::uint64_t>(a));
        }
        return static_cast<std::size_t>(h);
    }
};

// A constant's bits as its identity; every NaN of one sign is one value.
std::uint64_t number_bits(double v)
{
    if (std::isnan(v)) {
        return std::signbit(v) ? 0xfff8000000000000ull : 0x7ff8000000000000ull;
    }
    return std::bit_cast<std::uint64_t>(v);
}

class Builder {
public:
    explicit Builder(Module& m) : m_(m) {}

    // The instruction computing `root`'s value: its arguments' first, left to right, depth first.
    int lower(const Node& root)
    {
        struct Frame {
            const Node* node;
            std::size_t next = 0;
            std::vector<int> args;
        };
        std::vector<Frame> stack;
        stack.push_back(Frame{&root});
        int last = -1;
        while (!stack.empty()) {
            Frame& f = stack.back();
            if (f.next == 0) {
                if (const auto seen = done_.find(f.node); seen != done_.end()) {
                    last = seen->second;
                    stack.pop_back();
                    if (!stack.empty()) {
                        stack.back().args.push_back(last);
                    }
                    continue;
                }
            }
            // A shift's offsets are part of the instruction, not arguments.
            const std::size_t operands = is_shift(f.node->op) ? 1 : f.node->args.size();
            if (f.next < operands) {
                const Node* child = f.node->args[f.next++].get();
                stack.push_back(Frame{child});
                continue;
            }
            last = finish(*f.node, std::move(f.args));
            done_.emplace(f.node, last);
            stack.pop_back();
            if (!stack.empty()) {
                stack.back().args.push_back(last);
            }
        }
        return last;
    }

private:
    static bool is_shift(Op op) { return op == Op::Shift || op == Op::ShiftT; }

    [[nodiscard]] const Inst& inst(int id) const { return m_.insts[static_cast<std::size_t>(id)]; }
    [[nodiscard]] bool number_is(int id, double v) const { return inst(id).op == Op::Const && inst(id).number == v; }
    [[nodiscard]] bool scalar(int id) const { return inst(id).shape == Shape::Scalar; }

    int finish(const Node& n, std::vector<int> args)
    {
        Inst in;
        in.op = n.op;
        in.number = n.number;
        in.index = n.index;
        in.iterations = n.iterations;
        in.args = std::move(args);
        if (is_shift(n.op)) {
            for (int a = 0; a < 3; ++a) {
                in.offset[a] = static_cast<int>(n.args[static_cast<std::size_t>(a) + 1]->number);
            }
        }
        const int replaced = rewrite(in);
        if (replaced >= 0) {
            ++m_.folded;
            return replaced;
        }
        return intern(std::move(in));
    }

    int number(double v)
    {
        Inst c;
        c.op = Op::Const;
        c.number = v;
        return intern(std::move(c));
    }

    // The instruction `in` reduces to without being computed, or -1 when it stands.
    int rewrite(const Inst& in)
    {
        const auto& a = in.args;
        if (facts(in.op).foldable &&
            std::all_of(a.begin(), a.end(), [&](int id) { return inst(id).op == Op::Const; })) {
            std::vector<double> v;
            for (const int id : a) {
                v.push_back(inst(id).number);
            }
            return number(fold(in.op, v));
        }
        switch (in.op) {
        case Op::Mul: // 0 * anything is 0 without looking at the anything (as the CPU computes it); 1 * x is x
            if (number_is(a[0], 0.0) || number_is(a[1], 0.0)) {
                return number(0.0);
            }
            return number_is(a[0], 1.0) ? a[1] : number_is(a[1], 1.0) ? a[0] : -1;
        case Op::Add:
            return number_is(a[0], 0.0) ? a[1] : number_is(a[1], 0.0) ? a[0] : -1;
        case Op::Sub:
            return number_is(a[1], 0.0) ? a[0] : -1;
        case Op::Div:
            if (number_is(a[1], 1.0)) {
                return a[0];
            }
            return number_is(a[0], 0.0) ? number(0.0) : -1;
        case Op::Neg:
            return inst(a[0]).op == Op::Neg ? inst(a[0]).args[0] : -1;
        case Op::Shift:
            return scalar(a[0]) || (in.offset[0] == 0 && in.offset[1] == 0 && in.offset[2] == 0) ? a[0] : -1;
        // What reads a field that is one number everywhere sees no difference, so no gradient and no flux.
        case Op::Lap: case Op::Gx: case Op::Gy: case Op::Gz: case Op::Adv: case Op::Fall: case Op::Spread: case Op::Topple:
            return scalar(a[0]) ? number(0.0) : -1;
        case Op::Diffuse: case Op::LDiffuse:
            return scalar(a[1]) ? number(0.0) : -1;
        default:
            return -1;
        }
    }

    int intern(Inst in)
    {
        const OpFacts& f = facts(in.op);
        switch (f.shape) {
        case ShapeRule::Scalar: in.shape = Shape::Scalar; break;
        case ShapeRule::Cell: in.shape = Shape::Cell; break;
        case ShapeRule::Args:
            in.shape = std::any_of(in.args.begin(), in.args.end(), [&](int id) { return !scalar(id); }) ? Shape::Cell
                                                                                                       : Shape::Scalar;
            break;
        }
        in.radius = f.radius >= 0
            ? f.radius
            : std::max({std::abs(in.offset[0]), std::abs(in.offset[1]), std::abs(in.offset[2])});
        in.global = f.global;
        in.net = f.net;
        Key key{static_cast<int>(in.op), number_bits(in.number), in.index, in.iterations,
            {in.offset[0], in.offset[1], in.offset[2]}, in.args};
        const int next = static_cast<int>(m_.insts.size());
        const auto [it, fresh] = ids_.try_emplace(std::move(key), next);
        if (!fresh) {
            ++m_.shared;
            return it->second;
        }
        m_.insts.push_back(std::move(in));
        return next;
    }

    Module& m_;
    std::unordered_map<const Node*, int> done_;
    std::unordered_map<Key, int, KeyHash> ids_;
};

// Decides which values are stored as stages and what a point of every instruction costs with everything else
// inlined, in one pass over the instructions in id order (a value's readers come after it). A cell value read by a
// stencil is stored once instead of inlined when one point of the stencil would cost more than kInlineBudget through
// it; whatever a tridiagonal solve reads (a field aside) is stored, since its lines are walked one cell at a time.
void plan(Module& m)
{
    const std::size_t n = m.insts.size();
    m.materialized.assign(n, 0);
    for (std::size_t id = 0; id < n; ++id) {
        Inst& in = m.insts[id];
        const long long per = facts(in.op).points;
        long long cost = 1;
        for (const int a : in.args) {
            Inst& src = m.insts[static_cast<std::size_t>(a)];
            const bool stored = m.is_stage(a);
            long long unit = stored ? 1 : src.cost;
            const bool storable = !stored && src.shape == Shape::Cell && src.op != Op::Field;
            if (storable && (in.op == Op::Tridiag || (in.radius > 0 && unit * per > kInlineBudget))) {
                m.materialized[static_cast<std::size_t>(a)] = 1;
                unit = 1;
            }
            cost += unit * per;
            ++src.uses;
        }
        in.cost = static_cast<int>(std::min<long long>(cost, 1 << 20));
    }
}

// Marks every instruction reachable from `roots` (through all arguments), sweeping ids downwards.
std::vector<char> reachable(const Module& m, const std::vector<int>& roots)
{
    std::vector<char> mark(m.insts.size(), 0);
    for (const int r : roots) {
        mark[static_cast<std::size_t>(r)] = 1;
    }
    for (std::size_t id = m.insts.size(); id-- > 0;) {
        if (mark[id]) {
            for (const int a : m.insts[id].args) {
                mark[static_cast<std::size_t>(a)] = 1;
            }
        }
    }
    return mark;
}

// The instructions of a program that run over the lattice before its gather: stored cell values and global ops.
std::vector<int> stages_of(const Module& m, const std::vector<char>& mark)
{
    std::vector<int> out;
    for (std::size_t id = 0; id < mark.size(); ++id) {
        const Inst& in = m.insts[id];
        if (mark[id] && ((m.is_stage(static_cast<int>(id)) && in.shape == Shape::Cell) || in.global)) {
            out.push_back(static_cast<int>(id));
        }
    }
    return out;
}

// For every field, the farthest (in cells, along any path of inlined ops) the gather of `roots` reads it from its own
// cell; -1 where it does not read it. A stage is read where it was stored (it saw the state as it was), so a path
// stops at one unless it is a root itself; nets and global ops read what they read before the gather.
std::vector<int> read_distance(const Module& m, const std::vector<int>& roots, std::size_t field_count)
{
    constexpr int kNone = std::numeric_limits<int>::min();
    std::vector<int> entered(m.insts.size(), kNone); // the farthest offset a path enters the instruction at
    std::vector<char> root(m.insts.size(), 0);
    for (const int r : roots) {
        root[static_cast<std::size_t>(r)] = 1;
    }
    std::vector<int> out(field_count, -1);
    for (std::size_t id = m.insts.size(); id-- > 0;) {
        const Inst& in = m.insts[id];
        int at = root[id] ? std::max(entered[id], 0) : (m.is_stage(static_cast<int>(id)) ? kNone : entered[id]);
        if (root[id] && m.is_stage(static_cast<int>(id))) {
            at = 0; // a stored root is entered at its own cell only
        }
        if (at == kNone) {
            continue;
        }
        if (in.op == Op::Field) {
            int& d = out[static_cast<std::size_t>(in.index)];
            d = std::max(d, at);
            continue;
        }
        if (in.net || in.global) {
            continue;
        }
        for (const int a : in.args) {
            int& e = entered[static_cast<std::size_t>(a)];
            e = std::max(e, at + in.radius);
        }
    }
    return out;
}

Program program(Builder& b, const std::string& name, const std::vector<field::Equation>& eqs, const Node* max_dt)
{
    Program p;
    p.name = name;
    std::set<int> writes;
    for (const field::Equation& eq : eqs) {
        p.equations.push_back(Equation{eq.target, b.lower(*eq.rhs)});
        writes.insert(eq.target);
    }
    if (max_dt != nullptr) {
        p.max_dt = b.lower(*max_dt);
    }
    p.writes.assign(writes.begin(), writes.end());
    return p;
}

void complete(const Module& m, Program& p, std::size_t field_count)
{
    std::vector<int> roots;
    for (const Equation& eq : p.equations) {
        roots.push_back(eq.value);
    }
    const std::vector<char> mark = reachable(m, roots);
    p.stages = stages_of(m, mark);
    if (p.max_dt >= 0) {
        p.limit_stages = stages_of(m, reachable(m, {p.max_dt}));
    }
    p.reads.clear();
    std::set<int> reads;
    for (std::size_t id = 0; id < mark.size(); ++id) {
        if (mark[id] && m.insts[id].op == Op::Field) {
            reads.insert(m.insts[id].index);
        }
    }
    p.reads.assign(reads.begin(), reads.end());
    // In place only when no field it writes is read off its own cell: another cell's thread may already have
    // stored what this one would read.
    const std::vector<int> distance = read_distance(m, roots, field_count);
    p.in_place = std::none_of(p.writes.begin(), p.writes.end(),
        [&](int f) { return f >= 0 && static_cast<std::size_t>(f) < distance.size() && distance[static_cast<std::size_t>(f)] > 0; });
}

// ---- JSON ---------------------------------------------------------------------------------------------------------

class JsonText {
public:
    std::string str() const { return out_.str(); }

    JsonText& open(char c)
    {
        comma();
        out_ << c;
        first_.push_back(true);
        return *this;
    }
    JsonText& close(char c)
    {
        first_.pop_back();
        out_ << c;
        return *this;
    }
    // The next element of an array that lists one element a line.
    JsonText& line()
    {
        comma();
        out_ << '\n';
        pending_ = true; // the element follows without another separator
        return *this;
    }
    JsonText& key(const char* k, bool on_new_line = false)
    {
        comma();
        if (on_new_line) {
            out_ << '\n';
        }
        out_ << '"' << k << "\":";
        pending_ = true; // the value follows without a separator
        return *this;
    }
    JsonText& str(const std::string& s)
    {
        comma();
        out_ << '"';
        for (const char c : s) {
            if (c == '"' || c == '\\') {
                out_ << '\\' << c;
            } else {
                out_ << (static_cast<unsigned char>(c) < 0x20 ? ' ' : c);
            }
        }
        out_ << '"';
        return *this;
    }
    JsonText& num(double v)
    {
        comma();
        if (std::isnan(v)) {
            out_ << '0';
        } else if (std::isinf(v)) {
            out_ << (v > 0 ? "1e300" : "-1e300");
        } else {
            char text[40];
            std::snprintf(text, sizeof text, "%.17g", v);
            out_ << text;
        }
        return *this;
    }
    JsonText& integer(long long v)
    {
        comma();
        out_ << v;
        return *this;
    }
    JsonText& boolean(bool v)
    {
        comma();
        out_ << (v ? "true" : "false");
        return *this;
    }
    JsonText& ints(const std::vector<int>& v)
    {
        open('[');
        for (const int x : v) {
            integer(x);
        }
        return close(']');
    }
    JsonText& literal(const char* text)
    {
        comma();
        out_ << text;
        return *this;
    }

private:
    // A separator before a value: none after an opening bracket or a key.
    void comma()
    {
        if (pending_) {
            pending_ = false;
            return;
        }
        if (!first_.empty()) {
            if (!first_.back()) {
                out_ << ',';
            }
            first_.back() = false;
        }
        pending_ = false;
    }

    std::ostringstream out_;
    std::vector<bool> first_;
    bool pending_ = false;
};

void program_json(JsonText& j, const Program& p)
{
    j.open('{').key("name").str(p.name).key("equations").open('[');
    for (const Equation& eq : p.equations) {
        j.open('[').integer(eq.target).integer(eq.value).close(']');
    }
    j.close(']');
    j.key("stages").ints(p.stages).key("limit_stages").ints(p.limit_stages).key("in_place").boolean(p.in_place);
    j.key("max_dt").integer(p.max_dt).key("reads").ints(p.reads).key("writes").ints(p.writes).key("conserves").open('[');
    for (const std::vector<int>& g : p.conserves) {
        j.ints(g);
    }
    j.close(']').key("accumulate").boolean(p.accumulate).close('}');
}

} // namespace

const char* op_name(Op op)
{
    const int i = static_cast<int>(op);
    return i >= 0 && i < kOpCount ? kFacts[static_cast<std::size_t>(i)].name : "?";
}

bool Module::is_stage(int id) const
{
    const Inst& in = insts[static_cast<std::size_t>(id)];
    return in.global || in.net || materialized[static_cast<std::size_t>(id)] != 0;
}

Module lower(const Law& law)
{
    Module m;
    Builder b(m);
    for (const Term& t : law.terms) {
        m.terms.push_back(program(b, t.name, t.equations, t.max_dt.get()));
        m.terms.back().conserves = t.conserves;
        m.learned.push_back(program(b, t.name + "~learned", t.learned, t.learned_max_dt.get()));
        m.learned.back().conserves = t.conserves;
    }
    std::set<int> set_before;
    for (const Derive& d : law.derives) {
        Program p = program(b, "derive " + law.fields[static_cast<std::size_t>(d.target)].name,
            {field::Equation{d.target, d.rhs}}, nullptr);
        p.accumulate = d.accumulate && set_before.contains(d.target);
        set_before.insert(d.target);
        m.derives.push_back(std::move(p));
    }
    plan(m);
    for (auto* group : {&m.terms, &m.learned, &m.derives}) {
        for (Program& p : *group) {
            complete(m, p, law.fields.size());
        }
    }
    for (Program& p : m.derives) {
        p.in_place = false; // a derive replaces its field: the old value of its own cell is not an input
    }
    return m;
}

std::string to_json(const Module& m, const Law& law)
{
    JsonText j;
    j.open('{').key("version").integer(1).key("fields").open('[');
    for (const FieldDecl& f : law.fields) {
        j.open('{').key("name").str(f.name).key("init").num(f.init).key("lo").num(f.lo).key("hi").num(f.hi);
        j.key("clamp").boolean(f.clamp).key("static").boolean(f.is_static).key("derived").boolean(f.derived).close('}');
    }
    j.close(']').key("params").open('[');
    for (std::size_t p = 0; p < law.param_names.size(); ++p) {
        j.open('{').key("name").str(law.param_names[p]).key("value").num(law.param_values[p]).close('}');
    }
    j.close(']').key("nets").open('[');
    for (const NetDecl& n : law.nets) {
        j.open('{').key("name").str(n.name).key("file").str(n.file).key("sizes").ints(n.sizes);
        j.key("act").str(to_string(n.act)).key("half").boolean(n.half).key("loaded").boolean(n.loaded).close('}');
    }
    j.close(']').key("insts").open('[');
    for (std::size_t i = 0; i < m.insts.size(); ++i) {
        const Inst& in = m.insts[i];
        j.line().open('{').key("op").str(op_name(in.op)).key("shape").str(in.shape == Shape::Cell ? "cell" : "scalar");
        j.key("number").num(in.number).key("index").integer(in.index).key("iterations").integer(in.iterations);
        j.key("offset").open('[').integer(in.offset[0]).integer(in.offset[1]).integer(in.offset[2]).close(']');
        j.key("args").ints(in.args).key("radius").integer(in.radius);
        j.key("stage").boolean(m.is_stage(static_cast<int>(i))).close('}');
    }
    j.close(']');
    const std::pair<const char*, const std::vector<Program>*> groups[] = {
        {"terms", &m.terms}, {"learned", &m.learned}, {"derives", &m.derives}};
    for (const auto& [name, programs] : groups) {
        j.key(name, true).open('[');
        for (const Program& p : *programs) {
            j.line();
            program_json(j, p);
        }
        j.close(']');
    }
    j.close('}');
    return j.str() + "\n";
}

std::string describe(const Module& m, const Law& law)
{
    std::ostringstream out;
    out << "IR: " << m.insts.size() << " instructions; " << m.shared << " repeated values merged, " << m.folded
        << " folded or simplified away\n";
    for (std::size_t i = 0; i < m.insts.size(); ++i) {
        const Inst& in = m.insts[i];
        out << "  %" << i << " := " << op_name(in.op);
        switch (in.op) {
        case Op::Const: out << '(' << in.number << ')'; break;
        case Op::Field: out << '(' << law.fields[static_cast<std::size_t>(in.index)].name << ')'; break;
        case Op::Param: out << '(' << law.param_names[static_cast<std::size_t>(in.index)] << ')'; break;
        case Op::Tridiag: out << '_' << "xyz"[in.index]; break;
        default:
            if (in.net) {
                out << '[' << law.nets[static_cast<std::size_t>(in.index)].name << ']';
            }
            break;
        }
        for (std::size_t a = 0; a < in.args.size(); ++a) {
            out << (a == 0 ? " " : ", ") << '%' << in.args[a];
        }
        if (in.op == Op::Shift || in.op == Op::ShiftT) {
            out << " by (" << in.offset[0] << ", " << in.offset[1] << ", " << in.offset[2] << ')';
        }
        out << "    ; " << (in.shape == Shape::Cell ? "per cell" : "one number");
        if (in.radius > 0) {
            out << ", radius " << in.radius;
        }
        if (m.is_stage(static_cast<int>(i))) {
            out << ", " << (in.global ? "whole-lattice stage" : in.net ? "net stage" : "stored stage");
        }
        out << ", cost " << in.cost << '\n';
    }
    const auto show = [&](const Program& p) {
        if (p.equations.empty()) {
            return;
        }
        out << p.name << " (" << (p.in_place ? "in place" : "ping-pong") << ", " << p.stages.size() << " stage(s))\n";
        for (const int s : p.stages) {
            out << "    first %" << s << " (" << op_name(m.insts[static_cast<std::size_t>(s)].op) << ")\n";
        }
        for (const Equation& eq : p.equations) {
            out << "    " << law.fields[static_cast<std::size_t>(eq.target)].name << " += dt * %" << eq.value << '\n';
        }
        if (p.max_dt >= 0) {
            out << "    step at most %" << p.max_dt << '\n';
        }
    };
    for (const auto* group : {&m.terms, &m.learned, &m.derives}) {
        for (const Program& p : *group) {
            show(p);
        }
    }
    return out.str();
}

} // namespace fieldengine::field::ir
