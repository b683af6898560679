// Law text from a Law. See LawText.hpp.
#include "field/LawText.hpp"

#include <cmath>
#include <cstdio>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

namespace fieldengine::field {

namespace {

// Binding strengths as the grammar reads them: sums, products, a prefix sign, powers, then atoms.
enum Level { kSumLevel = 1, kProductLevel = 2, kSignLevel = 3, kPowerLevel = 4, kAtomLevel = 5 };

std::string number_text(double v)
{
    if (!std::isfinite(v)) {
        throw std::runtime_error("a law cannot state a number that is not finite");
    }
    char text[40];
    std::snprintf(text, sizeof text, "%.17g", std::fabs(v));
    return std::signbit(v) && v != 0.0 ? std::string("-") + text : std::string(text);
}

std::string bound_text(double v)
{
    if (v >= 1e30) {
        return "inf";
    }
    if (v <= -1e30) {
        return "-inf";
    }
    return number_text(v);
}

const char* function_name(const Node& n)
{
    switch (n.op) {
    case Op::Abs: return "abs";
    case Op::Sqrt: return "sqrt";
    case Op::Exp: return "exp";
    case Op::Log: return "log";
    case Op::Step: return "step";
    case Op::Pos: return "pos";
    case Op::Min: return "min";
    case Op::Max: return "max";
    case Op::Clamp: return "clamp";
    case Op::Lap: return "lap";
    case Op::Gx: return "gx";
    case Op::Gy: return "gy";
    case Op::Gz: return "gz";
    case Op::GxT: return "gx_t";
    case Op::GyT: return "gy_t";
    case Op::GzT: return "gz_t";
    case Op::Diffuse: return "diffuse";
    case Op::Adv: return "adv";
    case Op::Distance: return "distance";
    case Op::Poisson: return "poisson";
    case Op::MaxOf: return "maxof";
    case Op::MinOf: return "minof";
    case Op::SumOf: return "sumof";
    case Op::Fall: return "fall";
    case Op::Spread: return "spread";
    case Op::Topple: return "topple";
    case Op::Shift: return "shift";
    case Op::ShiftT: return "shift_t";
    case Op::Mlp: return "mlp";
    case Op::Flux: return "flux";
    case Op::LDiffuse: return "ldiffuse";
    case Op::Tridiag: return n.index == 0 ? "tridiag_x" : n.index == 1 ? "tridiag_y" : "tridiag_z";
    default: return nullptr;
    }
}

class Writer {
public:
    Writer(const Law& law, const std::map<const Node*, std::string>* lets) : law_(law), lets_(lets) {}

    // `n` at a position that needs at least `level` to stand without parentheses.
    std::string at(const Node& n, int level) const
    {
        if (lets_ != nullptr) {
            if (const auto it = lets_->find(&n); it != lets_->end()) {
                return it->second;
            }
        }
        const auto [text, own] = bare(n);
        return own >= level ? text : "(" + text + ")";
    }

    // The node's own text (a let's defining expression) and the level it stands at.
    std::pair<std::string, int> bare(const Node& n) const
    {
        const auto arg = [&](std::size_t i, int level) { return at(*n.args[i], level); };
        switch (n.op) {
        case Op::Const: return {number_text(n.number), n.number < 0.0 || std::signbit(n.number) ? kSignLevel : kAtomLevel};
        case Op::Param: return {law_.param_names[static_cast<std::size_t>(n.index)], kAtomLevel};
        case Op::Field: return {law_.fields[static_cast<std::size_t>(n.index)].name, kAtomLevel};
        case Op::Cell: return {"h", kAtomLevel};
        case Op::Dt: return {"dt", kAtomLevel};
        case Op::Rim: return {"rim", kAtomLevel};
        case Op::Neg: return {"-" + arg(0, kSignLevel), kSignLevel};
        case Op::Add: return {arg(0, kSumLevel) + " + " + arg(1, kProductLevel), kSumLevel};
        case Op::Sub: return {arg(0, kSumLevel) + " - " + arg(1, kProductLevel), kSumLevel};
        case Op::Mul: return {arg(0, kProductLevel) + " * " + arg(1, kSignLevel), kProductLevel};
        case Op::Div: return {arg(0, kProductLevel) + " / " + arg(1, kSignLevel), kProductLevel};
        case Op::Pow: return {arg(0, kAtomLevel) + " ^ " + arg(1, kSignLevel), kPowerLevel};
        case Op::Local: throw std::runtime_error("a creature's state cannot be written on the lattice");
        default: break;
        }
        const char* fn = function_name(n);
        if (fn == nullptr) {
            throw std::runtime_error("no law text for this op");
        }
        std::string out = std::string(fn) + "(";
        std::size_t count = n.args.size();
        if (n.op == Op::Mlp || n.op == Op::Flux || n.op == Op::LDiffuse) {
            out += law_.nets[static_cast<std::size_t>(n.index)].name + ", ";
            if (n.op == Op::Flux) {
                count -= 2; // the limiter's bounds, added by the front end from the field's range
            }
        }
        for (std::size_t i = 0; i < count; ++i) {
            const bool offset = (n.op == Op::Shift || n.op == Op::ShiftT) && i > 0;
            out += (i ? ", " : "") + (offset ? number_text(n.args[i]->number) : arg(i, 0));
        }
        if (n.op == Op::Poisson) {
            out += ", " + std::to_string(n.iterations);
        }
        return {out + ")", kAtomLevel};
    }

private:
    const Law& law_;
    const std::map<const Node*, std::string>* lets_;
};

// The subtrees of `roots` that are used more than once (leaves aside), in an order where each comes after those it
// uses: the lets a term needs.
std::vector<const Node*> shared_subtrees(const std::vector<const Node*>& roots)
{
    std::unordered_map<const Node*, int> uses;
    std::vector<const Node*> order; // post-order of first visits
    std::vector<std::pair<const Node*, std::size_t>> stack;
    for (const Node* r : roots) {
        if (uses[r]++ > 0) {
            continue;
        }
        stack.push_back({r, 0});
        while (!stack.empty()) {
            auto& [n, next] = stack.back();
            const bool offsets = n->op == Op::Shift || n->op == Op::ShiftT;
            const std::size_t operands = offsets ? 1 : n->args.size();
            if (next < operands) {
                const Node* child = n->args[next++].get();
                if (uses[child]++ == 0) {
                    stack.push_back({child, 0});
                }
                continue;
            }
            order.push_back(n);
            stack.pop_back();
        }
    }
    std::vector<const Node*> out;
    for (const Node* n : order) {
        if (uses[n] > 1 && !n->args.empty()) {
            out.push_back(n);
        }
    }
    return out;
}

bool name_taken(const Law& law, const std::string& name)
{
    return law.field_id(name) >= 0 || law.param_id(name) >= 0 || law.net_id(name) >= 0;
}

} // namespace

std::string expression_text(const Node& node, const Law& law)
{
    return Writer(law, nullptr).at(node, 0);
}

std::string law_text(const Law& law, const std::string& header)
{
    std::ostringstream out;
    if (!header.empty()) {
        std::istringstream lines(header);
        for (std::string line; std::getline(lines, line);) {
            out << "# " << line << '\n';
        }
    }
    for (const FieldDecl& f : law.fields) {
        out << "field " << f.name << " init " << number_text(f.init);
        if (f.lo > -1e30 || f.hi < 1e30) {
            out << " range " << bound_text(f.lo) << ' ' << bound_text(f.hi);
        }
        out << (f.is_static ? " static" : "") << (f.clamp ? " clamp" : "");
        if (!f.unit.empty()) {
            out << " unit " << f.unit;
        }
        out << '\n';
    }
    for (std::size_t p = 0; p < law.param_names.size(); ++p) {
        out << "param " << law.param_names[p] << " = " << number_text(law.param_values[p]) << '\n';
    }
    for (const NetDecl& n : law.nets) {
        out << "net " << n.name << " in " << n.inputs();
        if (n.layers() > 1) {
            out << " hidden";
            for (int l = 1; l < n.layers(); ++l) {
                out << ' ' << n.sizes[static_cast<std::size_t>(l)];
            }
        }
        out << " out " << n.outputs() << " act " << to_string(n.act) << (n.half ? " half" : "") << '\n';
    }
    int next_let = 0;
    for (const Term& t : law.terms) {
        out << "\nterm " << t.name << '\n';
        std::vector<const Node*> roots;
        for (const auto* group : {&t.equations, &t.learned}) {
            for (const Equation& eq : *group) {
                roots.push_back(eq.rhs.get());
            }
        }
        std::map<const Node*, std::string> lets;
        for (const Node* s : shared_subtrees(roots)) {
            std::string name;
            do {
                name = "v__" + std::to_string(++next_let);
            } while (name_taken(law, name));
            out << "  let " << name << " = " << Writer(law, &lets).bare(*s).first << '\n';
            lets.emplace(s, name);
        }
        const Writer w(law, &lets);
        for (const Equation& eq : t.equations) {
            out << "  d " << law.fields[static_cast<std::size_t>(eq.target)].name << " = " << w.at(*eq.rhs, 0) << '\n';
        }
        for (const Equation& eq : t.learned) {
            out << "  learned d " << law.fields[static_cast<std::size_t>(eq.target)].name << " = " << w.at(*eq.rhs, 0)
                << '\n';
        }
        for (const std::vector<int>& group : t.conserves) {
            out << "  conserves";
            for (std::size_t i = 0; i < group.size(); ++i) {
                out << (i ? " + " : " ") << law.fields[static_cast<std::size_t>(group[i])].name;
            }
            out << '\n';
        }
        if (t.max_dt) {
            out << "  max_dt = " << expression_text(*t.max_dt, law) << '\n';
        }
        if (t.learned_max_dt) {
            out << "  learned max_dt = " << expression_text(*t.learned_max_dt, law) << '\n';
        }
    }
    if (!law.derives.empty()) {
        out << '\n';
    }
    for (const Derive& d : law.derives) {
        out << "derive " << law.fields[static_cast<std::size_t>(d.target)].name << (d.accumulate ? " += " : " = ")
            << expression_text(*d.rhs, law) << '\n';
    }
    return out.str();
}

} // namespace fieldengine::field
