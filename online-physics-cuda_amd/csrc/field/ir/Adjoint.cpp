// The adjoint law of a term. See Adjoint.hpp.
//
// Built as expression trees, not as text: the term's instructions (from the IR, so folded and shared already) become
// the adjoint law's primal nodes, one node per instruction; a reverse sweep over the instructions, highest id first
// (every user of a value before the value), gathers each instruction's cotangent as a list of contributions, and a
// table of pullbacks turns an instruction's summed cotangent into its arguments' contributions. The result is a Law,
// written out by the law printer (shared subtrees become lets). The pullbacks are the derivatives of the reference
// CPU's own definitions of each op (sqrt and log of their clamped arguments, exp capped at 80, min/max/clamp choosing
// a branch by the same comparisons), so where the CPU takes a branch the derivative follows the same branch.
#include "field/ir/Adjoint.hpp"

#include "field/LawText.hpp"
#include "field/ir/LawIR.hpp"

#include <algorithm>
#include <functional>
#include <map>
#include <optional>
#include <stdexcept>
#include <vector>

namespace fieldengine::field::ir {

namespace {

using NodePtr = std::shared_ptr<Node>;

// A handle on an expression node with the law language's arithmetic, for writing derivatives as formulas.
struct X {
    NodePtr n;
};

X node(Op op, std::vector<X> args)
{
    auto n = std::make_shared<Node>();
    n->op = op;
    for (X& a : args) {
        n->args.push_back(std::move(a.n));
    }
    return X{n};
}

X num(double v)
{
    auto n = std::make_shared<Node>();
    n->op = Op::Const;
    n->number = v;
    return X{n};
}

X operator+(X a, X b) { return node(Op::Add, {std::move(a), std::move(b)}); }
X operator-(X a, X b) { return node(Op::Sub, {std::move(a), std::move(b)}); }
X operator*(X a, X b) { return node(Op::Mul, {std::move(a), std::move(b)}); }
X operator/(X a, X b) { return node(Op::Div, {std::move(a), std::move(b)}); }
X operator-(X a) { return node(Op::Neg, {std::move(a)}); }
X step(X a) { return node(Op::Step, {std::move(a)}); }
X call(Op op, X a) { return node(op, {std::move(a)}); }

X shifted(Op op, X a, const int (&offset)[3])
{
    X out = node(op, {std::move(a)});
    for (const int o : offset) {
        out.n->args.push_back(num(static_cast<double>(o)).n);
    }
    return out;
}

// What one instruction passes back to its arguments, given its cotangent `a`, its arguments' primal values `x` and its
// own primal value `y`: one entry per argument (null: nothing, the argument is not differentiable through this op).
using Pullback = std::function<std::vector<std::optional<X>>(const X& a, const std::vector<X>& x, const X& y,
    const Inst& in)>;

const std::map<Op, Pullback>& pullbacks()
{
    static const std::map<Op, Pullback> table = {
        {Op::Neg, [](const X& a, auto&, auto&, auto&) { return std::vector<std::optional<X>>{-a}; }},
        {Op::Add, [](const X& a, auto&, auto&, auto&) { return std::vector<std::optional<X>>{a, a}; }},
        {Op::Sub, [](const X& a, auto&, auto&, auto&) { return std::vector<std::optional<X>>{a, -a}; }},
        {Op::Mul, [](const X& a, const std::vector<X>& x, auto&, auto&) {
             return std::vector<std::optional<X>>{a * x[1], a * x[0]};
         }},
        // y = x0 / x1 (0 where x1 = 0: the division rule makes those terms 0 as well)
        {Op::Div, [](const X& a, const std::vector<X>& x, const X& y, auto&) {
             return std::vector<std::optional<X>>{a / x[1], -(a * y / x[1])};
         }},
        {Op::Pow, [](const X& a, const std::vector<X>& x, const X& y, auto&) {
             return std::vector<std::optional<X>>{a * x[1] * node(Op::Pow, {x[0], x[1] - num(1.0)}),
                 a * y * call(Op::Log, x[0])};
         }},
        // min(x0, x1) is x1 where x1 < x0, else x0; max(x0, x1) is x1 where x0 < x1, else x0.
        {Op::Min, [](const X& a, const std::vector<X>& x, auto&, auto&) {
             const X second = a * step(x[0] - x[1]);
             return std::vector<std::optional<X>>{a - second, second};
         }},
        {Op::Max, [](const X& a, const std::vector<X>& x, auto&, auto&) {
             const X second = a * step(x[1] - x[0]);
             return std::vector<std::optional<X>>{a - second, second};
         }},
        // clamp(v, lo, hi) is lo where v < lo, else top where top < v, else v; top = max(lo, hi) is hi where lo < hi.
        {Op::Clamp, [](const X& a, const std::vector<X>& x, auto&, auto&) {
             const X top = node(Op::Max, {x[1], x[2]});
             const X below = step(x[1] - x[0]);
             const X above = (num(1.0) - below) * step(x[0] - top);
             const X to_top = a * above;
             const X hi_wins = step(x[2] - x[1]);
             return std::vector<std::optional<X>>{a * (num(1.0) - below) * (num(1.0) - step(x[0] - top)),
                 a * below + to_top * (num(1.0) - hi_wins), to_top * hi_wins};
         }},
        {Op::Abs, [](const X& a, const std::vector<X>& x, auto&, auto&) {
             return std::vector<std::optional<X>>{a * (step(x[0]) - step(-x[0]))};
         }},
        // y = sqrt(max(x, 0)): dy/dx = 1 / (2 y) where x > 0
        {Op::Sqrt, [](const X& a, const std::vector<X>& x, const X& y, auto&) {
             return std::vector<std::optional<X>>{a * step(x[0]) / (num(2.0) * y)};
         }},
        // y = exp(min(x, 80))
        {Op::Exp, [](const X& a, const std::vector<X>& x, const X& y, auto&) {
             return std::vector<std::optional<X>>{a * y * step(num(80.0) - x[0])};
         }},
        // y = log(max(x, 1e-30))
        {Op::Log, [](const X& a, const std::vector<X>& x, auto&, auto&) {
             return std::vector<std::optional<X>>{a * step(x[0] - num(1e-30)) / x[0]};
         }},
        {Op::Pos, [](const X& a, const std::vector<X>& x, auto&, auto&) {
             return std::vector<std::optional<X>>{a * step(x[0])};
         }},
        {Op::Step, [](auto&, auto&, auto&, auto&) { return std::vector<std::optional<X>>{std::nullopt}; }},
        // The stencils' transposes: <op a, b> = <a, op^T b>. lap and diffuse (flux form, zero-flux edges) are
        // symmetric; diffuse in a scalar k is k times the unit diffusion where k > 0 (a face with k <= 0 is shut).
        {Op::Shift, [](const X& a, auto&, auto&, const Inst& in) {
             return std::vector<std::optional<X>>{shifted(Op::ShiftT, a, in.offset)};
         }},
        {Op::ShiftT, [](const X& a, auto&, auto&, const Inst& in) {
             return std::vector<std::optional<X>>{shifted(Op::Shift, a, in.offset)};
         }},
        {Op::Lap, [](const X& a, auto&, auto&, auto&) { return std::vector<std::optional<X>>{call(Op::Lap, a)}; }},
        {Op::Gx, [](const X& a, auto&, auto&, auto&) { return std::vector<std::optional<X>>{call(Op::GxT, a)}; }},
        {Op::Gy, [](const X& a, auto&, auto&, auto&) { return std::vector<std::optional<X>>{call(Op::GyT, a)}; }},
        {Op::Gz, [](const X& a, auto&, auto&, auto&) { return std::vector<std::optional<X>>{call(Op::GzT, a)}; }},
        {Op::GxT, [](const X& a, auto&, auto&, auto&) { return std::vector<std::optional<X>>{call(Op::Gx, a)}; }},
        {Op::GyT, [](const X& a, auto&, auto&, auto&) { return std::vector<std::optional<X>>{call(Op::Gy, a)}; }},
        {Op::GzT, [](const X& a, auto&, auto&, auto&) { return std::vector<std::optional<X>>{call(Op::Gz, a)}; }},
        {Op::Diffuse, [](const X& a, const std::vector<X>& x, auto&, auto&) {
             return std::vector<std::optional<X>>{a * step(x[0]) * node(Op::Diffuse, {num(1.0), x[1]}),
                 node(Op::Diffuse, {x[0], a})};
         }},
    };
    return table;
}

bool is_leaf(Op op)
{
    return op == Op::Const || op == Op::Param || op == Op::Field || op == Op::Cell || op == Op::Dt || op == Op::Rim;
}

// Leaves through which nothing is differentiated.
bool constant_leaf(Op op) { return op == Op::Const || op == Op::Cell || op == Op::Dt || op == Op::Rim; }

struct Refused : std::runtime_error {
    using std::runtime_error::runtime_error;
};

} // namespace

std::string adjoint_law(const Law& law, const std::string& term, std::string& error)
{
    const Module m = lower(law);
    std::size_t which = 0;
    if (!term.empty()) {
        while (which < m.terms.size() && m.terms[which].name != term) {
            ++which;
        }
    }
    if (which >= m.terms.size()) {
        error = "no term '" + term + "'";
        return {};
    }
    const Program& p = m.terms[which];
    if (p.equations.empty()) {
        error = "term '" + p.name + "' has no equations";
        return {};
    }
    try {
        // The adjoint law's declarations: the fields the term touches as static inputs (the state its step began
        // from), then a cotangent field for each, then params' cotangent fields (added below as they are met).
        Law out;
        std::vector<int> touched(p.reads);
        touched.insert(touched.end(), p.writes.begin(), p.writes.end());
        std::sort(touched.begin(), touched.end());
        touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
        std::map<int, int> primal_field, cotangent_field; // primal law's field id -> adjoint law's
        for (const int f : touched) {
            FieldDecl d;
            d.name = law.fields[static_cast<std::size_t>(f)].name;
            d.init = law.fields[static_cast<std::size_t>(f)].init;
            d.is_static = true;
            primal_field[f] = static_cast<int>(out.fields.size());
            out.fields.push_back(d);
        }
        for (const int f : touched) {
            FieldDecl d;
            d.name = law.fields[static_cast<std::size_t>(f)].name + "__adj";
            cotangent_field[f] = static_cast<int>(out.fields.size());
            out.fields.push_back(d);
        }
        out.param_names = law.param_names;
        out.param_values = law.param_values;
        const auto field_node = [&](int id) {
            X x = node(Op::Field, {});
            x.n->index = id;
            return x;
        };

        // Which instructions the equations reach, and each one's primal value as a node of the adjoint law.
        const std::size_t n = m.insts.size();
        std::vector<char> live(n, 0);
        for (const Equation& eq : p.equations) {
            live[static_cast<std::size_t>(eq.value)] = 1;
        }
        for (std::size_t id = n; id-- > 0;) {
            if (live[id]) {
                for (const int a : m.insts[id].args) {
                    live[static_cast<std::size_t>(a)] = 1;
                }
            }
        }
        std::vector<X> value(n);
        for (std::size_t id = 0; id < n; ++id) {
            if (!live[id]) {
                continue;
            }
            const Inst& in = m.insts[id];
            if (!is_leaf(in.op) && !pullbacks().contains(in.op)) {
                throw Refused(std::string(op_name(in.op)) + " has no adjoint here");
            }
            if (in.op == Op::Diffuse && m.insts[static_cast<std::size_t>(in.args[0])].shape != Shape::Scalar) {
                throw Refused("diffuse whose k is a field has no adjoint here");
            }
            X v = node(in.op, {});
            v.n->number = in.number;
            v.n->index = in.op == Op::Field ? primal_field.at(in.index) : in.index;
            v.n->iterations = in.iterations;
            for (const int a : in.args) {
                v.n->args.push_back(value[static_cast<std::size_t>(a)].n);
            }
            if (in.op == Op::Shift || in.op == Op::ShiftT) {
                for (const int o : in.offset) {
                    v.n->args.push_back(num(static_cast<double>(o)).n);
                }
            }
            value[id] = v;
        }

        // The reverse sweep. An equation d F = r seeds r's cotangent with F's cotangent field.
        std::vector<std::vector<X>> parts(n);
        for (const Equation& eq : p.equations) {
            parts[static_cast<std::size_t>(eq.value)].push_back(field_node(cotangent_field.at(eq.target)));
        }
        std::map<int, X> field_rate, param_rate;
        for (std::size_t id = n; id-- > 0;) {
            if (parts[id].empty()) {
                continue;
            }
            X a = parts[id][0];
            for (std::size_t k = 1; k < parts[id].size(); ++k) {
                a = a + parts[id][k];
            }
            const Inst& in = m.insts[id];
            if (in.op == Op::Field) {
                field_rate.emplace(in.index, a);
                continue;
            }
            if (in.op == Op::Param) {
                param_rate.emplace(in.index, a);
                continue;
            }
            if (is_leaf(in.op)) {
                continue;
            }
            std::vector<X> x;
            for (const int arg : in.args) {
                x.push_back(value[static_cast<std::size_t>(arg)]);
            }
            const std::vector<std::optional<X>> back = pullbacks().at(in.op)(a, x, value[id], in);
            for (std::size_t k = 0; k < back.size(); ++k) {
                const int arg = in.args[k];
                if (back[k] && !constant_leaf(m.insts[static_cast<std::size_t>(arg)].op)) {
                    parts[static_cast<std::size_t>(arg)].push_back(*back[k]);
                }
            }
        }

        Term t;
        t.name = "adjoint";
        for (const auto& [f, rate] : field_rate) {
            t.equations.push_back(field::Equation{cotangent_field.at(f), rate.n});
        }
        for (const auto& [param, rate] : param_rate) {
            FieldDecl d;
            d.name = law.param_names[static_cast<std::size_t>(param)] + "__adj";
            t.equations.push_back(field::Equation{static_cast<int>(out.fields.size()), rate.n});
            out.fields.push_back(d);
        }
        out.terms.push_back(std::move(t));
        const std::string header = "The adjoint of term '" + p.name + "', generated from its IR by reverse mode.\n"
            "One step of it is one vector-Jacobian product of one step of the term: every field F the term touches\n"
            "gets F__adj += dt (dR/dF)^T A, every param p it reads a field p__adj += dt (dR/dp)^T A whose total over\n"
            "the lattice is dL/dp. The fields without __adj hold the state the primal step began from.";
        return law_text(out, header);
    } catch (const std::exception& e) {
        error = "term '" + p.name + "': " + e.what();
        return {};
    }
}

} // namespace fieldengine::field::ir
