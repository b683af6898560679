// The adjoint law generator. See Adjoint.hpp.
#include "field/ir/Adjoint.hpp"

#include "field/ir/LawIR.hpp"

#include <cmath>
#include <cstdio>
#include <map>
#include <set>
#include <sstream>
#include <vector>

namespace fieldengine::field::ir {

namespace {

// A number as law text: every digit of the double, a negative one as (0 - x) (the grammar has no unary minus on a
// literal that survives every context).
std::string literal(double v)
{
    char buf[48];
    std::snprintf(buf, sizeof(buf), "%.17g", std::fabs(v));
    return v < 0.0 ? std::string("(0 - ") + buf + ")" : std::string(buf);
}

struct Writer {
    const Law& law;
    const Module& m;
    std::string error;

    [[nodiscard]] const Inst& at(int id) const { return m.insts[static_cast<std::size_t>(id)]; }

    // How the adjoint law names a primal value: leaves by their own names, every other value by its let.
    [[nodiscard]] std::string name(int id) const
    {
        const Inst& in = at(id);
        switch (in.op) {
        case Op::Const: return literal(in.number);
        case Op::Param: return law.param_names[static_cast<std::size_t>(in.index)];
        case Op::Field: return law.fields[static_cast<std::size_t>(in.index)].name;
        case Op::Cell: return "h";
        case Op::Dt: return "dt";
        case Op::Rim: return "rim";
        default: return "fwd" + std::to_string(id);
        }
    }

    [[nodiscard]] static std::string shifts(const Inst& in)
    {
        return std::to_string(in.offset[0]) + ", " + std::to_string(in.offset[1]) + ", " + std::to_string(in.offset[2]);
    }

    // The primal value of a non-leaf instruction as an expression of its arguments' names; false when it has no form.
    bool primal(int id, std::string& out)
    {
        const Inst& in = at(id);
        auto a = [&](std::size_t k) { return name(in.args[k]); };
        switch (in.op) {
        case Op::Neg: out = "(0 - " + a(0) + ")"; return true;
        case Op::Add: out = "(" + a(0) + " + " + a(1) + ")"; return true;
        case Op::Sub: out = "(" + a(0) + " - " + a(1) + ")"; return true;
        case Op::Mul: out = "(" + a(0) + " * " + a(1) + ")"; return true;
        case Op::Div: out = "(" + a(0) + " / " + a(1) + ")"; return true;
        case Op::Pow: out = "(" + a(0) + " ^ " + a(1) + ")"; return true;
        case Op::Min: out = "min(" + a(0) + ", " + a(1) + ")"; return true;
        case Op::Max: out = "max(" + a(0) + ", " + a(1) + ")"; return true;
        case Op::Clamp: out = "clamp(" + a(0) + ", " + a(1) + ", " + a(2) + ")"; return true;
        case Op::Diffuse: out = "diffuse(" + a(0) + ", " + a(1) + ")"; return true;
        case Op::Shift: out = "shift(" + a(0) + ", " + shifts(in) + ")"; return true;
        case Op::ShiftT: out = "shift_t(" + a(0) + ", " + shifts(in) + ")"; return true;
        case Op::Abs: case Op::Sqrt: case Op::Exp: case Op::Log: case Op::Step: case Op::Pos: case Op::Lap: case Op::Gx:
        case Op::Gy: case Op::Gz: case Op::GxT: case Op::GyT: case Op::GzT:
            out = std::string(op_name(in.op)) + "(" + a(0) + ")";
            return true;
        default:
            error = std::string(op_name(in.op)) + " has no adjoint here";
            return false;
        }
    }
};

} // namespace

std::string adjoint_law(const Law& law, const std::string& term, std::string& error)
{
    const Module m = lower(law);
    std::size_t t = 0;
    while (!term.empty() && t < m.terms.size() && m.terms[t].name != term) {
        ++t;
    }
    if (t >= m.terms.size()) {
        error = term.empty() ? "the law has no term" : "no term '" + term + "'";
        return {};
    }
    const Program& p = m.terms[t];
    if (p.equations.empty()) {
        error = "term '" + p.name + "' has no equations";
        return {};
    }
    Writer w{law, m, {}};

    // Every instruction the equations reach, in id order (ids are topological: an instruction's arguments come first).
    std::set<int> reached;
    std::vector<int> todo;
    for (const ir::Equation& eq : p.equations) {
        todo.push_back(eq.value);
    }
    while (!todo.empty()) {
        const int id = todo.back();
        todo.pop_back();
        if (reached.insert(id).second) {
            for (const int a : w.at(id).args) {
                todo.push_back(a);
            }
        }
    }

    std::ostringstream body;
    for (const int id : reached) {
        const Op op = w.at(id).op;
        if (op == Op::Const || op == Op::Param || op == Op::Field || op == Op::Cell || op == Op::Dt || op == Op::Rim) {
            continue;
        }
        std::string e;
        if (!w.primal(id, e)) {
            error = "term '" + p.name + "': " + w.error;
            return {};
        }
        body << "  let fwd" << id << " = " << e << "\n";
    }

    // Reverse sweep: each instruction's adjoint is the sum of what its readers pass it; the equations' values are
    // seeded with their targets' adjoint fields.
    std::map<int, std::vector<std::string>> incoming;
    for (const ir::Equation& eq : p.equations) {
        incoming[eq.value].push_back(law.fields[static_cast<std::size_t>(eq.target)].name + "__adj");
    }
    std::map<int, std::string> field_adjoint, param_adjoint; // field / param index -> its adjoint's let
    for (auto it = reached.rbegin(); it != reached.rend(); ++it) {
        const int id = *it;
        const auto found = incoming.find(id);
        if (found == incoming.end()) {
            continue;
        }
        const Inst& in = w.at(id);
        const std::string A = "adj" + std::to_string(id);
        body << "  let " << A << " = ";
        for (std::size_t k = 0; k < found->second.size(); ++k) {
            body << (k ? " + " : "") << found->second[k];
        }
        body << "\n";
        auto pass = [&](std::size_t k, const std::string& e) {
            const Op to = w.at(in.args[k]).op;
            if (to != Op::Const && to != Op::Cell && to != Op::Dt && to != Op::Rim) {
                incoming[in.args[k]].push_back(e);
            }
        };
        auto a = [&](std::size_t k) { return w.name(in.args[k]); };
        const std::string self = w.name(id);
        switch (in.op) {
        case Op::Field: field_adjoint[in.index] = A; break;
        case Op::Param: param_adjoint[in.index] = A; break;
        case Op::Const: case Op::Cell: case Op::Dt: case Op::Rim: case Op::Step: break; // constant, or flat almost everywhere
        case Op::Neg: pass(0, "(0 - " + A + ")"); break;
        case Op::Add: pass(0, A); pass(1, A); break;
        case Op::Sub: pass(0, A); pass(1, "(0 - " + A + ")"); break;
        case Op::Mul: pass(0, "(" + A + " * " + a(1) + ")"); pass(1, "(" + A + " * " + a(0) + ")"); break;
        case Op::Div: pass(0, "(" + A + " / " + a(1) + ")"); pass(1, "(0 - " + A + " * " + self + " / " + a(1) + ")"); break;
        case Op::Pow:
            pass(0, "(" + A + " * " + a(1) + " * (" + a(0) + " ^ (" + a(1) + " - 1)))");
            pass(1, "(" + A + " * " + self + " * log(" + a(0) + "))");
            break;
        case Op::Min: // min(x, y) is y where x > y (std::min's comparison), else x
            pass(0, "(" + A + " * (1 - step(" + a(0) + " - " + a(1) + ")))");
            pass(1, "(" + A + " * step(" + a(0) + " - " + a(1) + "))");
            break;
        case Op::Max: // max(x, y) is y where y > x, else x
            pass(0, "(" + A + " * (1 - step(" + a(1) + " - " + a(0) + ")))");
            pass(1, "(" + A + " * step(" + a(1) + " - " + a(0) + "))");
            break;
        case Op::Abs: pass(0, "(" + A + " * (step(" + a(0) + ") - step(0 - " + a(0) + ")))"); break;
        case Op::Sqrt: pass(0, "(" + A + " * step(" + a(0) + ") * 0.5 / " + self + ")"); break;
        case Op::Exp: pass(0, "(" + A + " * " + self + " * step(80 - " + a(0) + "))"); break;
        case Op::Log: pass(0, "(" + A + " * step(" + a(0) + " - 1e-30) / " + a(0) + ")"); break;
        case Op::Pos: pass(0, "(" + A + " * step(" + a(0) + "))"); break;
        case Op::Clamp: {
            const std::string hi = "max(" + a(1) + ", " + a(2) + ")";
            pass(0, "(" + A + " * step(" + a(0) + " - " + a(1) + ") * step(" + hi + " - " + a(0) + "))");
            pass(1, "(" + A + " * (1 - step(" + a(0) + " - " + a(1) + ")))");
            pass(2, "(" + A + " * step(" + a(0) + " - " + hi + ") * step(" + a(2) + " - " + a(1) + "))");
            break;
        }
        // Linear operators pass their transposes; lap and a scalar-k diffuse are symmetric.
        case Op::Shift: pass(0, "shift_t(" + A + ", " + Writer::shifts(in) + ")"); break;
        case Op::ShiftT: pass(0, "shift(" + A + ", " + Writer::shifts(in) + ")"); break;
        case Op::Lap: pass(0, "lap(" + A + ")"); break;
        case Op::Gx: pass(0, "gx_t(" + A + ")"); break;
        case Op::Gy: pass(0, "gy_t(" + A + ")"); break;
        case Op::Gz: pass(0, "gz_t(" + A + ")"); break;
        case Op::GxT: pass(0, "gx(" + A + ")"); break;
        case Op::GyT: pass(0, "gy(" + A + ")"); break;
        case Op::GzT: pass(0, "gz(" + A + ")"); break;
        case Op::Diffuse:
            if (w.at(in.args[0]).shape != Shape::Scalar) {
                error = "term '" + p.name + "': diffuse whose k is a field has no adjoint here";
                return {};
            }
            // In F: itself. In k (one number): the unit-k flux form of F where k > 0 (the face mean of k <= 0 is 0).
            pass(1, "diffuse(" + a(0) + ", " + A + ")");
            pass(0, "(" + A + " * step(" + a(0) + ") * diffuse(1, " + a(1) + "))");
            break;
        default:
            error = "term '" + p.name + "': " + std::string(op_name(in.op)) + " has no adjoint here";
            return {};
        }
    }

    std::set<int> fields(p.reads.begin(), p.reads.end());
    fields.insert(p.writes.begin(), p.writes.end());
    std::ostringstream out;
    out << "# Adjoint of term '" << p.name << "': a step of it is one vector-Jacobian product of one step of the term,\n"
        << "# X__adj += dt (dR/dX)^T A for every field X, and for a param p the field p__adj (its total is dL/dp).\n"
        << "# The primal fields hold the state the primal step began from.\n";
    for (const int f : fields) {
        const FieldDecl& d = law.fields[static_cast<std::size_t>(f)];
        out << "field " << d.name << " init " << literal(d.init) << " range -1e30 1e30 static\n";
    }
    for (const int f : fields) {
        out << "field " << law.fields[static_cast<std::size_t>(f)].name << "__adj init 0 range -1e30 1e30\n";
    }
    for (const auto& [index, let] : param_adjoint) {
        out << "field " << law.param_names[static_cast<std::size_t>(index)] << "__adj init 0 range -1e30 1e30\n";
    }
    for (std::size_t i = 0; i < law.param_names.size(); ++i) {
        out << "param " << law.param_names[i] << " = " << literal(law.param_values[i]) << "\n";
    }
    out << "\nterm adjoint\n" << body.str();
    for (const auto& [index, let] : field_adjoint) {
        out << "  d " << law.fields[static_cast<std::size_t>(index)].name << "__adj = " << let << "\n";
    }
    for (const auto& [index, let] : param_adjoint) {
        out << "  d " << law.param_names[static_cast<std::size_t>(index)] << "__adj = " << let << "\n";
    }
    return out.str();
}

} // namespace fieldengine::field::ir
