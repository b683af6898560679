// Interval bounds of law expressions. See LawBounds.hpp; the rules are the reference's (src/field/LawBounds.cpp:13-269),
// so a term is idle here exactly when it would be there.
#include "field/LawBounds.hpp"

#include <algorithm>
#include <cmath>

namespace fieldengine::field {

namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();

Bounds point(double v) { return Bounds{v, v}; }
Bounds unbounded() { return Bounds{}; }

// A NaN end (inf - inf and the like) becomes the open end on its side.
Bounds tidy(double lo, double hi) { return Bounds{std::isnan(lo) ? -kInf : lo, std::isnan(hi) ? kInf : hi}; }

// How far the values spread: 0 for a flat range (also a flat infinite one).
double spread(const Bounds& b) { return b.lo == b.hi ? 0.0 : b.hi - b.lo; }

double magnitude(const Bounds& b) { return std::max(std::abs(b.lo), std::abs(b.hi)); }

// The hull of the four corner values.
Bounds hull(double a, double b, double c, double d)
{
    return Bounds{std::min(std::min(a, b), std::min(c, d)), std::max(std::max(a, b), std::max(c, d))};
}

Bounds symmetric(double most) { return Bounds{-most, most}; }

class Evaluator {
public:
    Evaluator(const std::vector<Bounds>& fields, const std::vector<double>& params, double h) : fields_(fields), params_(params), h_(h) {}

    Bounds of(const Node& n) const
    {
        switch (n.op) {
        case Op::Const: return point(n.number);
        case Op::Param: return point(static_cast<float>(params_[static_cast<std::size_t>(n.index)])); // the card's float
        case Op::Cell: return point(h_);
        case Op::Rim: return Bounds{0.0, 1.0};
        case Op::Field: return fields_[static_cast<std::size_t>(n.index)];
        case Op::Shift: case Op::MaxOf: case Op::MinOf: return arg(n, 0);
        default: break;
        }
        if (n.args.size() == 1 || n.op == Op::Lap || n.op == Op::Gx || n.op == Op::Gy || n.op == Op::Gz) {
            return unary(n);
        }
        return wider(n);
    }

private:
    Bounds arg(const Node& n, std::size_t i) const { return of(*n.args[i]); }

    Bounds unary(const Node& n) const
    {
        const Bounds a = arg(n, 0);
        switch (n.op) {
        case Op::Neg: return Bounds{-a.hi, -a.lo};
        case Op::Abs: return Bounds{a.lo <= 0.0 && a.hi >= 0.0 ? 0.0 : std::min(std::abs(a.lo), std::abs(a.hi)), magnitude(a)};
        case Op::Sqrt: return Bounds{std::sqrt(std::max(a.lo, 0.0)), std::sqrt(std::max(a.hi, 0.0))};
        case Op::Exp: return Bounds{std::exp(std::min(a.lo, 80.0)), std::exp(std::min(a.hi, 80.0))};
        case Op::Log: return Bounds{std::log(std::max(a.lo, 1e-30)), std::log(std::max(a.hi, 1e-30))};
        case Op::Step: return Bounds{a.lo > 0.0 ? 1.0 : 0.0, a.hi > 0.0 ? 1.0 : 0.0};
        case Op::Pos: return Bounds{std::max(a.lo, 0.0), std::max(a.hi, 0.0)};
        case Op::Lap: return symmetric(6.0 * spread(a) / (h_ * h_));
        case Op::Gx: case Op::Gy: case Op::Gz: return symmetric(spread(a) / h_);
        case Op::Adv: case Op::Fall: case Op::Spread: case Op::Topple: case Op::SumOf: return a.none() ? a : unbounded();
        default: return unbounded();
        }
    }

    Bounds wider(const Node& n) const
    {
        switch (n.op) {
        case Op::Add: { const Bounds a = arg(n, 0), b = arg(n, 1); return tidy(a.lo + b.lo, a.hi + b.hi); }
        case Op::Sub: { const Bounds a = arg(n, 0), b = arg(n, 1); return tidy(a.lo - b.hi, a.hi - b.lo); }
        case Op::Mul: {
            const Bounds a = arg(n, 0), b = arg(n, 1);
            return hull(times(a.lo, b.lo), times(a.lo, b.hi), times(a.hi, b.lo), times(a.hi, b.hi));
        }
        case Op::Div: {
            const Bounds a = arg(n, 0), b = arg(n, 1);
            if (a.none() || b.none()) {
                return point(0.0); // a division by a zero divisor is 0 in the law language
            }
            if (b.lo <= 0.0 && b.hi >= 0.0) {
                return unbounded();
            }
            const Bounds q = hull(a.lo / b.lo, a.lo / b.hi, a.hi / b.lo, a.hi / b.hi);
            return tidy(q.lo, q.hi);
        }
        case Op::Pow: {
            const Bounds a = arg(n, 0), b = arg(n, 1);
            if (a.none() && b.lo > 0.0) {
                return point(0.0);
            }
            const bool fixed_positive_power = b.lo == b.hi && b.lo > 0.0;
            return a.lo >= 0.0 && fixed_positive_power ? tidy(std::pow(a.lo, b.lo), std::pow(a.hi, b.lo)) : unbounded();
        }
        case Op::Min: { const Bounds a = arg(n, 0), b = arg(n, 1); return Bounds{std::min(a.lo, b.lo), std::min(a.hi, b.hi)}; }
        case Op::Max: { const Bounds a = arg(n, 0), b = arg(n, 1); return Bounds{std::max(a.lo, b.lo), std::max(a.hi, b.hi)}; }
        case Op::Clamp: {
            const Bounds a = arg(n, 0), lo = arg(n, 1), hi = arg(n, 2);
            const double top = std::max(lo.hi, hi.hi);
            return Bounds{std::min(std::max(a.lo, lo.lo), top), std::min(std::max(a.hi, lo.hi), top)};
        }
        case Op::Diffuse: {
            const Bounds k = arg(n, 0), f = arg(n, 1);
            return symmetric(times(12.0 * magnitude(k), spread(f)) / (h_ * h_));
        }
        case Op::Flux: {
            // A net's antisymmetric flux is zero across every face when none of its inputs differs anywhere.
            for (std::size_t i = 0; i + 2 < n.args.size(); ++i) {
                const Bounds b = arg(n, i);
                if (!(b.lo == b.hi)) {
                    return unbounded();
                }
            }
            return point(0.0);
        }
        case Op::LDiffuse: { const Bounds f = arg(n, 1); return f.lo == f.hi ? point(0.0) : unbounded(); }
        case Op::Adv: case Op::Fall: case Op::Spread: case Op::Topple: case Op::SumOf: {
            const Bounds a = arg(n, 0);
            return a.none() ? a : unbounded();
        }
        default: return unbounded();
        }
    }

    const std::vector<Bounds>& fields_;
    const std::vector<double>& params_;
    double h_;
};

// Whether a rate is heat running down its own gradient (diffuse of the field it changes): it keeps the field between its
// coldest and hottest cell at any stable step, so it moves no bound.
bool within_own_range(const Equation& eq)
{
    const Node& r = *eq.rhs;
    return r.op == Op::Diffuse && r.args[1]->op == Op::Field && r.args[1]->index == eq.target;
}

} // namespace

double times(double a, double b)
{
    return a == 0.0 || b == 0.0 ? 0.0 : a * b;
}

Bounds bounds_of(const Node& n, const std::vector<Bounds>& fields, const std::vector<double>& params, double h)
{
    return Evaluator(fields, params, h).of(n);
}

bool always_zero(const Node& n, const std::vector<char>& zero, const std::vector<double>& params)
{
    const auto z = [&](std::size_t i) { return always_zero(*n.args[i], zero, params); };
    switch (n.op) {
    case Op::Const: return n.number == 0.0;
    case Op::Param: return params[static_cast<std::size_t>(n.index)] == 0.0;
    case Op::Field: return zero[static_cast<std::size_t>(n.index)] != 0;
    // zero in, zero out
    case Op::Neg: case Op::Abs: case Op::Sqrt: case Op::Step: case Op::Pos: case Op::Shift: case Op::Lap: case Op::Gx:
    case Op::Gy: case Op::Gz: case Op::MaxOf: case Op::MinOf: case Op::SumOf: case Op::Adv: case Op::Fall: case Op::Spread:
    case Op::Topple:
        return z(0);
    // a zero factor is enough
    case Op::Mul: case Op::Div: case Op::Diffuse: return z(0) || z(1);
    // both sides must be zero
    case Op::Add: case Op::Sub: case Op::Min: case Op::Max: return z(0) && z(1);
    case Op::Flux:
        for (std::size_t i = 0; i + 2 < n.args.size(); ++i) {
            if (!z(i)) {
                return false;
            }
        }
        return true;
    case Op::LDiffuse: return z(1);
    default: return false;
    }
}

void flow_bounds(const std::vector<Equation>& equations, double dt, int substeps, const Law& law, double h, std::vector<Bounds>& flowing)
{
    // A substep moves each cell by sub * (its rate), and its rate lies within bounds_of over the ranges at the
    // substep's start. Two passes a substep: an in-place term's later equation may read an earlier one's new value.
    // Substeps are carried one by one (growth compounds), until one moves nothing.
    const int count = std::max(substeps, 1);
    const double sub = dt / count;
    const std::vector<Bounds> entry = flowing;
    for (int s = 0; s < count; ++s) {
        const std::vector<Bounds> at_start = flowing;
        std::vector<Bounds> cur = flowing;
        for (int pass = 0; pass < 2; ++pass) {
            std::vector<Bounds> grown = cur;
            for (const Equation& eq : equations) {
                if (within_own_range(eq)) {
                    continue;
                }
                const std::size_t t = static_cast<std::size_t>(eq.target);
                const Bounds rate = bounds_of(*eq.rhs, cur, law.param_values, h);
                grown[t].lo = std::min(grown[t].lo, at_start[t].lo + times(sub, std::min(rate.lo, 0.0)));
                grown[t].hi = std::max(grown[t].hi, at_start[t].hi + times(sub, std::max(rate.hi, 0.0)));
            }
            cur = std::move(grown);
        }
        bool moved = false;
        for (std::size_t f = 0; f < cur.size(); ++f) {
            moved = moved || cur[f].lo != flowing[f].lo || cur[f].hi != flowing[f].hi;
        }
        flowing = std::move(cur);
        if (!moved) {
            break;
        }
    }
    // The cells move in float, the bounds in double: a bound that moved is widened by a few roundings a substep (by a
    // fraction of itself, so a bound at exactly 0 stays there). A clamped field is clamped after the substeps, so its
    // bounds are too; an unclamped one may leave its range (a violation, not a bound).
    std::vector<char> done(flowing.size(), 0);
    for (const Equation& eq : equations) {
        const std::size_t t = static_cast<std::size_t>(eq.target);
        if (done[t]) {
            continue;
        }
        done[t] = 1;
        Bounds& b = flowing[t];
        if (b.lo != entry[t].lo || b.hi != entry[t].hi) {
            b.lo -= 1e-6 * count * std::abs(b.lo);
            b.hi += 1e-6 * count * std::abs(b.hi);
        }
        const FieldDecl& decl = law.fields[t];
        if (decl.clamp) {
            b.lo = std::clamp(b.lo, decl.lo, decl.hi);
            b.hi = std::clamp(b.hi, decl.lo, decl.hi);
        }
    }
}

} // namespace fieldengine::field
