// The law language's front end: source text to a Law of expression trees, or the list of what is wrong, by line.
//
// The language is the reference's (grammar, keywords, functions, refusals: src/field/Law.cpp:47-410,1619-2439), so
// a law written for the reference compiles here to the same trees; the code is organised differently:
//
//   - a scanner turns each line into a Statement (lexemes with bounds-safe lookahead; past the end reads as the end);
//   - expressions are read by precedence climbing over one binding-power table, and the built-in functions are a
//     table of CallRule entries (op, arity, and what their arguments must be), checked in one place;
//   - statements are dispatched by keyword through two tables, one per pass (declarations first, so a term may use a
//     field declared later or in another source; then params, terms, lets, derives);
//   - a refusal inside an expression is an exception carrying the message, so the first problem met on a line is the
//     one reported, exactly as the reference's "first error wins" parser reports it.
//
// Creature blocks (the reference's demo agents, read at points and never on the lattice) are passed over up to the
// next top-level keyword; `show` lines are read into Law::visuals.
#include "field/Law.hpp"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

namespace fieldengine::field {

namespace {

using NodePtr = std::shared_ptr<Node>;

NodePtr node_of(Op op, std::vector<NodePtr> args = {})
{
    auto n = std::make_shared<Node>();
    n->op = op;
    n->args = std::move(args);
    return n;
}

NodePtr number_node(double v)
{
    NodePtr n = node_of(Op::Const);
    n->number = v;
    return n;
}

template <typename Visit>
void walk(const Node& n, const Visit& visit)
{
    visit(n);
    for (const NodePtr& a : n.args) {
        walk(*a, visit);
    }
}

std::vector<int> fields_read(const std::vector<const Node*>& roots)
{
    std::set<int> out;
    for (const Node* r : roots) {
        walk(*r, [&](const Node& n) {
            if (n.op == Op::Field) {
                out.insert(n.index);
            }
        });
    }
    return {out.begin(), out.end()};
}

bool mentions_dt(const Node& root)
{
    bool found = false;
    walk(root, [&](const Node& n) { found = found || n.op == Op::Dt; });
    return found;
}

// ---- scanning ---------------------------------------------------------------------------------------------------------

enum class Lx : unsigned char { Num, Word, Punct, Eol };

struct Lexeme {
    Lx kind = Lx::Eol;
    std::string text;
    double value = 0.0;
};

// One line's lexemes. Reading past the last one gives the end marker, so a statement may look ahead freely.
class Statement {
public:
    std::string where;

    [[nodiscard]] const Lexeme& operator[](std::size_t i) const { return i < lex_.size() ? lex_[i] : end_; }
    [[nodiscard]] bool word(std::size_t i) const { return (*this)[i].kind == Lx::Word; }
    [[nodiscard]] bool word(std::size_t i, std::string_view w) const { return word(i) && (*this)[i].text == w; }
    [[nodiscard]] bool punct(std::size_t i, char c) const
    {
        return (*this)[i].kind == Lx::Punct && (*this)[i].text[0] == c;
    }
    [[nodiscard]] bool end(std::size_t i) const { return (*this)[i].kind == Lx::Eol; }
    [[nodiscard]] bool empty() const { return lex_.empty(); }

    std::vector<Lexeme> lex_;

private:
    static inline const Lexeme end_{};
};

constexpr std::string_view kPunctuation = "+-*/^(),=";

bool digit(char c) { return c >= '0' && c <= '9'; }
bool word_start(char c) { return std::isalpha(static_cast<unsigned char>(c)) != 0 || c == '_'; }
bool word_char(char c) { return std::isalnum(static_cast<unsigned char>(c)) != 0 || c == '_'; }

// A line to lexemes; a comment runs from '#' to the end. On a character the language has no use for, or a number
// out of double's range, `error` says which and the line yields nothing.
std::vector<Lexeme> scan(std::string_view line, std::string& error)
{
    std::vector<Lexeme> out;
    std::size_t i = 0;
    const std::size_t n = line.size();
    while (i < n && line[i] != '#') {
        const char c = line[i];
        std::size_t j = i + 1;
        if (std::isspace(static_cast<unsigned char>(c)) != 0) {
            i = j;
            continue;
        }
        if (digit(c) || (c == '.' && j < n && digit(line[j]))) {
            // Digits and points greedily, then an exponent only when a digit follows its sign.
            j = i;
            while (j < n && (digit(line[j]) || line[j] == '.')) {
                ++j;
            }
            if (j < n && (line[j] == 'e' || line[j] == 'E')) {
                std::size_t k = j + 1;
                k += (k < n && (line[k] == '+' || line[k] == '-')) ? 1 : 0;
                if (k < n && digit(line[k])) {
                    for (j = k; j < n && digit(line[j]); ++j) {
                    }
                }
            }
            Lexeme num{Lx::Num, std::string(line.substr(i, j - i)), 0.0};
            errno = 0;
            num.value = std::strtod(num.text.c_str(), nullptr); // the longest prefix that is a number
            if (errno == ERANGE) {
                error = "bad number '" + num.text + "'";
                return {};
            }
            out.push_back(std::move(num));
        } else if (word_start(c)) {
            while (j < n && word_char(line[j])) {
                ++j;
            }
            out.push_back(Lexeme{Lx::Word, std::string(line.substr(i, j - i)), 0.0});
        } else if (kPunctuation.find(c) != std::string_view::npos) {
            out.push_back(Lexeme{Lx::Punct, std::string(1, c), 0.0});
        } else {
            error = std::string("unexpected character '") + c + "'";
            return {};
        }
        i = j;
    }
    return out;
}

// ---- built-in functions -------------------------------------------------------------------------------------------

// What a call's arguments must be, checked after they are read.
enum class Shape { Plain, Axis, NetRows, NetFaces, NetWindow, Sweeps, Offsets };

struct CallRule {
    Op op;
    int fewest;
    int most;
    Shape shape = Shape::Plain;
};

const std::unordered_map<std::string_view, CallRule>& call_rules()
{
    // Arity counts a net's name as an argument (mlp, flux, ldiffuse take one first).
    static const std::unordered_map<std::string_view, CallRule> rules = {
        {"abs", {Op::Abs, 1, 1}}, {"sqrt", {Op::Sqrt, 1, 1}}, {"exp", {Op::Exp, 1, 1}}, {"log", {Op::Log, 1, 1}},
        {"step", {Op::Step, 1, 1}}, {"pos", {Op::Pos, 1, 1}}, {"min", {Op::Min, 2, 2}}, {"max", {Op::Max, 2, 2}},
        {"clamp", {Op::Clamp, 3, 3}}, {"lap", {Op::Lap, 1, 1}}, {"gx", {Op::Gx, 1, 1}}, {"gy", {Op::Gy, 1, 1}},
        {"gz", {Op::Gz, 1, 1}}, {"gx_t", {Op::GxT, 1, 1}}, {"gy_t", {Op::GyT, 1, 1}}, {"gz_t", {Op::GzT, 1, 1}},
        {"diffuse", {Op::Diffuse, 2, 2}}, {"adv", {Op::Adv, 4, 4}}, {"distance", {Op::Distance, 1, 1}},
        {"maxof", {Op::MaxOf, 1, 1}}, {"minof", {Op::MinOf, 1, 1}}, {"sumof", {Op::SumOf, 1, 1}},
        {"fall", {Op::Fall, 3, 3}}, {"spread", {Op::Spread, 3, 3}}, {"topple", {Op::Topple, 3, 3}},
        {"poisson", {Op::Poisson, 1, 2, Shape::Sweeps}}, {"shift", {Op::Shift, 4, 4, Shape::Offsets}},
        {"shift_t", {Op::ShiftT, 4, 4, Shape::Offsets}}, {"tridiag_x", {Op::Tridiag, 4, 4, Shape::Axis}},
        {"tridiag_y", {Op::Tridiag, 4, 4, Shape::Axis}}, {"tridiag_z", {Op::Tridiag, 4, 4, Shape::Axis}},
        {"mlp", {Op::Mlp, 2, 33, Shape::NetRows}}, {"flux", {Op::Flux, 2, 9, Shape::NetFaces}},
        {"ldiffuse", {Op::LDiffuse, 3, 3, Shape::NetWindow}},
    };
    return rules;
}

bool takes_net(std::string_view fn)
{
    const auto it = call_rules().find(fn);
    return it != call_rules().end() &&
        (it->second.shape == Shape::NetRows || it->second.shape == Shape::NetFaces || it->second.shape == Shape::NetWindow);
}

// Names no declaration may take: the statement words and every built-in function.
bool reserved(std::string_view name)
{
    static const std::set<std::string_view> words = {"field", "param", "term", "derive", "creature", "show", "d",
        "conserves", "max_dt", "h", "dt", "pi", "inf", "init", "range", "unit", "static", "clamp", "net", "learned",
        "plan", "observe", "let"};
    return words.contains(name) || call_rules().contains(name);
}

// ---- expressions --------------------------------------------------------------------------------------------------

struct Refusal : std::runtime_error {
    using std::runtime_error::runtime_error;
};

using Lets = std::map<std::string, NodePtr, std::less<>>;

// Binding powers: + - bind loosest, then * /, then a prefix sign, then ^ (right-associative, with its exponent read
// at the prefix sign's level so 2^-1 is a power and -2^2 is -(2^2)).
constexpr int kSum = 10;
constexpr int kProduct = 20;
constexpr int kSign = 30;
constexpr int kPower = 40;

class ExprReader {
public:
    ExprReader(const Statement& s, std::size_t at, const Law& law, const Lets* lets)
        : s_(s), at_(at), law_(law), lets_(lets)
    {
    }

    // The whole rest of the line as one expression; `trailing` is the refusal when more follows it.
    NodePtr read_line(const char* trailing = "trailing tokens after the expression")
    {
        NodePtr e = read(0);
        if (!s_.end(at_)) {
            throw Refusal(trailing);
        }
        return e;
    }

private:
    // The binding power of the infix operator at the cursor; 0 when there is none.
    [[nodiscard]] int infix_power() const
    {
        if (s_[at_].kind != Lx::Punct) {
            return 0;
        }
        switch (s_[at_].text[0]) {
        case '+': case '-': return kSum;
        case '*': case '/': return kProduct;
        case '^': return kPower;
        default: return 0;
        }
    }

    NodePtr read(int floor)
    {
        NodePtr left = operand();
        for (int power = infix_power(); power > floor; power = infix_power()) {
            const char sym = s_[at_++].text[0];
            // Left-associative operators read their right side one notch tighter; ^ reads its exponent at the sign's
            // level, which takes a further ^ (right-associative) and a sign but no product or sum.
            NodePtr right = read(sym == '^' ? kSign : power);
            const Op op = sym == '+' ? Op::Add : sym == '-' ? Op::Sub : sym == '*' ? Op::Mul : sym == '/' ? Op::Div : Op::Pow;
            left = node_of(op, {left, right});
        }
        return left;
    }

    NodePtr operand()
    {
        const Lexeme& lx = s_[at_];
        if (s_.punct(at_, '-') || s_.punct(at_, '+')) {
            ++at_;
            NodePtr inner = read(kSign);
            return lx.text[0] == '-' ? node_of(Op::Neg, {inner}) : inner;
        }
        switch (lx.kind) {
        case Lx::Num:
            ++at_;
            return number_node(lx.value);
        case Lx::Word:
            ++at_;
            return s_.punct(at_, '(') ? call(lx.text) : named(lx.text);
        case Lx::Punct:
            if (lx.text[0] == '(') {
                ++at_;
                NodePtr inner = read(0);
                expect(')');
                return inner;
            }
            throw Refusal("unexpected '" + lx.text + "'");
        case Lx::Eol:
            break;
        }
        throw Refusal("expression ends too early");
    }

    void expect(char c)
    {
        if (!s_.punct(at_, c)) {
            throw Refusal(std::string("expected '") + c + "'");
        }
        ++at_;
    }

    NodePtr call(const std::string& fn)
    {
        ++at_; // '('
        int net = -1;
        if (takes_net(fn)) {
            net = s_.word(at_) ? law_.net_id(s_[at_].text) : -1;
            if (net < 0) {
                throw Refusal("'" + fn + "' takes the name of a declared net first");
            }
            ++at_;
            if (!s_.punct(at_, ',')) {
                throw Refusal("expected ',' after the net's name");
            }
            ++at_;
        }
        std::vector<NodePtr> args;
        if (!s_.punct(at_, ')')) {
            args.push_back(read(0));
            while (s_.punct(at_, ',')) {
                ++at_;
                args.push_back(read(0));
            }
        }
        expect(')');
        const auto rule = call_rules().find(fn);
        if (rule == call_rules().end()) {
            throw Refusal("unknown function '" + fn + "'");
        }
        const CallRule& r = rule->second;
        const int given = static_cast<int>(args.size()) + (net >= 0 ? 1 : 0);
        if (given < r.fewest || given > r.most) {
            std::string span = std::to_string(r.fewest);
            if (r.most != r.fewest) {
                span += "-" + std::to_string(r.most);
            }
            throw Refusal("'" + fn + "' takes " + span + " argument(s), got " + std::to_string(given));
        }
        NodePtr out = node_of(r.op);
        switch (r.shape) {
        case Shape::Plain:
            break;
        case Shape::Axis:
            out->index = fn.back() - 'x';
            break;
        case Shape::NetRows: case Shape::NetFaces: case Shape::NetWindow:
            check_net(fn, r, net, args);
            out->index = net;
            break;
        case Shape::Sweeps:
            out->iterations = 24;
            if (args.size() == 2) {
                const Node& count = *args[1];
                if (count.op != Op::Const || count.number < 1.0) {
                    throw Refusal("poisson's iteration count must be a positive number");
                }
                out->iterations = static_cast<int>(count.number);
            }
            args.resize(1);
            break;
        case Shape::Offsets:
            for (std::size_t a = 1; a < 4; ++a) {
                args[a] = number_node(whole_cells(*args[a]));
            }
            break;
        }
        out->args = std::move(args);
        return out;
    }

    // A shift offset: a whole number of cells, at most 8, written as a number or a negated number.
    static double whole_cells(const Node& arg)
    {
        const bool negated = arg.op == Op::Neg;
        const Node& v = negated ? *arg.args[0] : arg;
        if (v.op != Op::Const || v.number != std::floor(v.number) || std::abs(v.number) > 8.0) {
            throw Refusal("shift's offsets must be whole numbers of cells (at most 8)");
        }
        return negated ? -v.number : v.number;
    }

    void check_net(const std::string& fn, const CallRule& r, int net, std::vector<NodePtr>& args) const
    {
        const NetDecl& decl = law_.nets[static_cast<std::size_t>(net)];
        const int rows = static_cast<int>(args.size());
        // mlp feeds its arguments in; flux feeds each of its fields from both cells of a face; ldiffuse feeds k and
        // four cells of the face's window.
        const int feeds = r.shape == Shape::NetRows ? rows : r.shape == Shape::NetFaces ? 2 * rows : 5;
        if (decl.inputs() != feeds) {
            throw Refusal("net '" + decl.name + "' takes " + std::to_string(decl.inputs()) + " inputs; " + fn +
                " here gives it " + std::to_string(feeds));
        }
        if (r.shape == Shape::NetWindow && decl.outputs() != 4) {
            throw Refusal("ldiffuse's net gives 4 weights (one per cell of the face's window), '" + decl.name +
                "' gives " + std::to_string(decl.outputs()));
        }
        if (r.shape != Shape::NetRows) {
            const std::size_t moved = r.shape == Shape::NetWindow ? 1 : 0;
            if (args[moved]->op != Op::Field) {
                throw Refusal(fn + " moves a field: its " + (moved == 0 ? "first" : "second") +
                    " argument after the net is a field's name");
            }
        }
        if (r.shape == Shape::NetFaces) {
            // The limiter keeps the moved field within its declared range: the bounds ride along as two constants.
            const FieldDecl& moved = law_.fields[static_cast<std::size_t>(args[0]->index)];
            args.push_back(number_node(moved.lo));
            args.push_back(number_node(moved.hi));
        }
    }

    NodePtr named(const std::string& name) const
    {
        static const std::unordered_map<std::string_view, Op> builtins = {{"h", Op::Cell}, {"dt", Op::Dt}, {"rim", Op::Rim}};
        if (const auto b = builtins.find(name); b != builtins.end()) {
            return node_of(b->second);
        }
        if (name == "pi") {
            return number_node(3.14159265358979323846);
        }
        if (lets_ != nullptr) {
            if (const auto l = lets_->find(name); l != lets_->end()) {
                return l->second; // every use of a let is the same node
            }
        }
        for (const auto& [op, id] : {std::pair{Op::Param, law_.param_id(name)}, std::pair{Op::Field, law_.field_id(name)}}) {
            if (id >= 0) {
                NodePtr n = node_of(op);
                n->index = id;
                return n;
            }
        }
        throw Refusal("unknown name '" + name + "' (not a field, param, h or pi)");
    }

    const Statement& s_;
    std::size_t at_;
    const Law& law_;
    const Lets* lets_;
};

// ---- a param's value: the reference CPU's scalar arithmetic on a one-cell lattice (h = 1, no step) ------------------

std::optional<double> param_value(const Node& n, const std::vector<double>& params)
{
    std::vector<std::optional<double>> a;
    a.reserve(n.args.size());
    for (const NodePtr& arg : n.args) {
        a.push_back(param_value(*arg, params));
    }
    const auto all = [&] { return std::all_of(a.begin(), a.end(), [](const auto& v) { return v.has_value(); }); };
    switch (n.op) {
    case Op::Const: return n.number;
    case Op::Param: return params[static_cast<std::size_t>(n.index)];
    case Op::Cell: return 1.0;
    case Op::Dt: case Op::Local: return 0.0;
    // Every stencil of a one-cell lattice reads one uniform value: no difference, no flux.
    case Op::Lap: case Op::Gx: case Op::Gy: case Op::Gz: case Op::Diffuse: case Op::Adv: return 0.0;
    // Zero times anything is zero, whatever the anything is.
    case Op::Mul:
        if ((a[0] && *a[0] == 0.0) || (a[1] && *a[1] == 0.0)) {
            return 0.0;
        }
        break;
    default:
        break;
    }
    if (!all()) {
        return std::nullopt;
    }
    const double x = a.empty() ? 0.0 : *a[0];
    const double y = a.size() > 1 ? *a[1] : 0.0;
    switch (n.op) {
    case Op::Neg: return -x;
    case Op::Abs: return std::fabs(x);
    case Op::Sqrt: return std::sqrt(0.0 < x ? x : 0.0);
    case Op::Exp: return std::exp(80.0 < x ? 80.0 : x);
    case Op::Log: return std::log(x < 1e-30 ? 1e-30 : x);
    case Op::Step: return x > 0.0 ? 1.0 : 0.0;
    case Op::Pos: return x < 0.0 ? 0.0 : x;
    case Op::Add: return x + y;
    case Op::Sub: return x - y;
    case Op::Mul: return x * y;
    case Op::Div: return y == 0.0 ? 0.0 : x / y;
    case Op::Pow: return std::pow(x, y);
    case Op::Min: return y < x ? y : x;
    case Op::Max: return x < y ? y : x;
    case Op::Clamp: {
        const double top = y < *a[2] ? *a[2] : y;
        return x < y ? y : (top < x ? top : x);
    }
    case Op::Shift: case Op::MaxOf: case Op::MinOf: case Op::SumOf: return x; // one cell: the value itself
    default: return std::nullopt; // rim (an array even here), nets
    }
}

// A bound in a declaration: a number or inf, either optionally negated.
bool read_bound(const Statement& s, std::size_t& i, double& out)
{
    const bool negative = s.punct(i, '-');
    const std::size_t at = negative ? i + 1 : i;
    double v = 0.0;
    if (s[at].kind == Lx::Num) {
        v = s[at].value;
    } else if (s.word(at, "inf")) {
        v = 1e30;
    } else {
        return false;
    }
    out = negative ? -v : v;
    i = at + 1;
    return true;
}

// ---- statements ---------------------------------------------------------------------------------------------------

bool starts_block(std::string_view kw)
{
    return kw == "field" || kw == "param" || kw == "term" || kw == "derive" || kw == "show" || kw == "net";
}

class Compiler {
public:
    explicit Compiler(std::vector<std::string>& errors) : errors_(errors) {}

    Law law;

    void declare(const Statement& s)
    {
        static const std::unordered_map<std::string_view, void (Compiler::*)(const Statement&)> pass1 = {
            {"net", &Compiler::declare_net}, {"field", &Compiler::declare_field},
            {"derive", &Compiler::declare_derived}, {"creature", &Compiler::declare_creature}};
        if (!s.word(0)) {
            return;
        }
        if (const auto h = pass1.find(s[0].text); h != pass1.end()) {
            (this->*(h->second))(s);
        }
    }

    void define(const Statement& s)
    {
        static const std::unordered_map<std::string_view, void (Compiler::*)(const Statement&)> pass2 = {
            {"show", &Compiler::define_visual}, {"param", &Compiler::define_param}, {"term", &Compiler::define_term},
            {"let", &Compiler::define_let}, {"derive", &Compiler::define_derive}, {"learned", &Compiler::define_learned},
            {"d", &Compiler::define_rate}, {"conserves", &Compiler::define_conserves}, {"max_dt", &Compiler::define_max_dt}};
        if (!s.word(0)) {
            refuse(s, "a line starts with a keyword (field, param, term, let, derive, creature, d, conserves, max_dt)");
            return;
        }
        const std::string& kw = s[0].text;
        if (starts_block(kw)) {
            in_creature_ = false;
        }
        if (kw == "creature") {
            term_ = nullptr;
            in_creature_ = s.word(1) && std::find(creatures_.begin(), creatures_.end(), s[1].text) != creatures_.end();
            return;
        }
        if (in_creature_) {
            return; // a creature's own line
        }
        if (kw == "net" || kw == "field") {
            term_ = nullptr;
            return; // read in the first pass
        }
        const auto h = pass2.find(kw);
        if (h == pass2.end()) {
            refuse(s, "unknown keyword '" + kw + "'");
            return;
        }
        (this->*(h->second))(s);
    }

    void validate();

private:
    void refuse(const Statement& s, const std::string& why) { errors_.push_back(s.where + ": " + why); }

    // Reads an expression from `at` to the end of the line; null (the refusal reported) when it cannot.
    NodePtr expression(const Statement& s, std::size_t at, const Lets* lets,
        const char* trailing = "trailing tokens after the expression")
    {
        try {
            return ExprReader(s, at, law, lets).read_line(trailing);
        } catch (const Refusal& r) {
            refuse(s, r.what());
            return nullptr;
        }
    }

    // ---- the first pass: what names exist ----

    void declare_net(const Statement& s)
    {
        if (!s.word(1) || reserved(s[1].text) || law.net_id(s[1].text) >= 0) {
            refuse(s, "expected 'net NAME in N hidden H... out M act tanh|relu|silu [half] [file PATH]'");
            return;
        }
        NetDecl net;
        net.name = s[1].text;
        net.where = s.where;
        net.file = "assets/nets/" + net.name + ".net";
        int in = 0, out = 0;
        std::vector<int> hidden;
        std::string section; // the word numbers after it belong to
        bool bad = false;
        for (std::size_t i = 2; !s.end(i) && !bad;) {
            const Lexeme& lx = s[i++];
            if (lx.kind == Lx::Num) {
                const int v = static_cast<int>(lx.value);
                if (section == "in") {
                    in = v;
                } else if (section == "out") {
                    out = v;
                } else if (section == "hidden") {
                    hidden.push_back(v);
                } else {
                    bad = true;
                }
                continue;
            }
            section = lx.text;
            if (section == "act") {
                bad = !s.word(i) || !parse_activation(s[i].text, net.act);
                ++i;
            } else if (section == "half") {
                net.half = true;
            } else if (section == "file") {
                // A path is glued back from its lexemes up to the end or the next `half` / `act`.
                std::string path;
                for (; !s.end(i) && !s.word(i, "half") && !s.word(i, "act"); ++i) {
                    path += s[i].kind == Lx::Num ? std::to_string(static_cast<long long>(s[i].value)) : s[i].text;
                }
                bad = path.empty();
                net.file = path;
            } else {
                bad = section != "in" && section != "hidden" && section != "out";
            }
        }
        const bool sized = in >= 1 && in <= 64 && out >= 1 && out <= 64 && hidden.size() <= 6 &&
            std::all_of(hidden.begin(), hidden.end(), [](int w) { return w >= 1 && w <= 128; });
        if (bad || !sized) {
            refuse(s, "malformed net (in 1..64, hidden up to 6 layers of 1..128, out 1..64)");
            return;
        }
        net.sizes = {in};
        net.sizes.insert(net.sizes.end(), hidden.begin(), hidden.end());
        net.sizes.push_back(out);
        init_net(net);
        std::string why;
        if (!load_net(net, why) && why.rfind("no weights at", 0) != 0) {
            refuse(s, why); // a weight file that is there but does not fit
            return;
        }
        law.nets.push_back(std::move(net));
    }

    void declare_field(const Statement& s)
    {
        if (!s.word(1) || reserved(s[1].text)) {
            refuse(s, "'field' needs a name that is not a keyword");
            return;
        }
        if (law.field_id(s[1].text) >= 0) {
            refuse(s, "field '" + s[1].text + "' declared twice");
            return;
        }
        FieldDecl decl;
        decl.name = s[1].text;
        bool ok = true;
        for (std::size_t i = 2; !s.end(i) && ok;) {
            const std::string key = s[i++].text;
            if (key == "init") {
                ok = read_bound(s, i, decl.init);
            } else if (key == "range") {
                ok = read_bound(s, i, decl.lo) && read_bound(s, i, decl.hi) && decl.lo <= decl.hi;
            } else if (key == "unit") {
                ok = s.word(i);
                decl.unit = s[i++].text;
            } else if (key == "static" || key == "clamp") {
                (key == "static" ? decl.is_static : decl.clamp) = true;
            } else {
                ok = false;
            }
        }
        if (!ok) {
            refuse(s, "malformed field declaration (expected init N, range LO HI, unit U, static, clamp)");
            return;
        }
        law.fields.push_back(std::move(decl));
    }

    void declare_derived(const Statement& s)
    {
        // `derive NAME ...` of a name no field has yet declares it, derived.
        if (s.word(1) && law.field_id(s[1].text) < 0 && !reserved(s[1].text)) {
            FieldDecl decl;
            decl.name = s[1].text;
            decl.derived = true;
            law.fields.push_back(std::move(decl));
        }
    }

    void declare_creature(const Statement& s)
    {
        if (!s.word(1) || !s.end(2) || reserved(s[1].text)) {
            refuse(s, "expected 'creature NAME'");
            return;
        }
        if (std::find(creatures_.begin(), creatures_.end(), s[1].text) != creatures_.end()) {
            refuse(s, "creature '" + s[1].text + "' declared twice");
            return;
        }
        creatures_.push_back(s[1].text);
    }

    // ---- the second pass: what they mean ----

    void define_visual(const Statement& s)
    {
        term_ = nullptr;
        FieldVisual v;
        v.field = s.word(1) ? law.field_id(s[1].text) : -1;
        bool ok = v.field >= 0;
        bool coloured = false;
        for (std::size_t i = 2; ok && !s.end(i);) {
            const std::string key = s[i++].text;
            if (key == "colour") {
                double rgb[3] = {0.0, 0.0, 0.0};
                ok = read_bound(s, i, rgb[0]) && read_bound(s, i, rgb[1]) && read_bound(s, i, rgb[2]);
                for (int c = 0; c < 3; ++c) {
                    v.colour[c] = static_cast<float>(rgb[c]);
                }
                coloured = true;
            } else if (key == "above" || key == "full" || key == "alpha") {
                ok = read_bound(s, i, key == "above" ? v.above : key == "full" ? v.full : v.alpha);
            } else {
                ok = false;
            }
        }
        if (!ok || !coloured || v.full <= v.above) {
            refuse(s, "expected 'show FIELD colour R G B above LO full HI [alpha A]' with a declared field and HI > LO");
            return;
        }
        law.visuals.push_back(v);
    }

    void define_param(const Statement& s)
    {
        term_ = nullptr;
        if (!s.word(1) || reserved(s[1].text) || !s.punct(2, '=')) {
            refuse(s, "expected 'param NAME = expression'");
            return;
        }
        if (law.param_id(s[1].text) >= 0 || law.field_id(s[1].text) >= 0) {
            refuse(s, "'" + s[1].text + "' is already a param or a field");
            return;
        }
        const NodePtr e = expression(s, 3, nullptr);
        if (!e) {
            return;
        }
        if (!always_scalar(*e)) {
            refuse(s, "a param cannot read a field");
            return;
        }
        law.param_names.push_back(s[1].text);
        law.param_values.push_back(param_value(*e, law.param_values).value_or(0.0));
    }

    void define_term(const Statement& s)
    {
        if (!s.word(1) || !s.end(2)) {
            refuse(s, "expected 'term NAME'");
            term_ = nullptr;
            return;
        }
        for (const Term& t : law.terms) {
            if (t.name == s[1].text) {
                refuse(s, "term '" + s[1].text + "' declared twice");
            }
        }
        Term t;
        t.name = s[1].text;
        t.where = s.where;
        law.terms.push_back(std::move(t));
        term_ = &law.terms.back();
        lets_.clear();
    }

    // The term a rate line belongs to; null (refused) outside one.
    Term* inside_term(const Statement& s, const std::string& what)
    {
        if (term_ == nullptr) {
            refuse(s, "'" + what + "' outside a term");
        }
        return term_;
    }

    void define_let(const Statement& s)
    {
        if (inside_term(s, "let") == nullptr) {
            return;
        }
        if (!s.word(1) || !s.punct(2, '=')) {
            refuse(s, "expected 'let NAME = expression'");
            return;
        }
        const std::string& name = s[1].text;
        if (reserved(name) || law.field_id(name) >= 0 || law.param_id(name) >= 0 || lets_.contains(name)) {
            refuse(s, "'" + name + "' is already a keyword, field, param or let");
            return;
        }
        if (NodePtr e = expression(s, 3, &lets_)) {
            lets_.emplace(name, std::move(e));
        }
    }

    void define_derive(const Statement& s)
    {
        term_ = nullptr;
        const int target = s.word(1) ? law.field_id(s[1].text) : -1;
        const bool adds = s.punct(2, '+') && s.punct(3, '=');
        if (target < 0 || !(adds || s.punct(2, '='))) {
            refuse(s, "expected 'derive FIELD = expression' or 'derive FIELD += expression'");
            return;
        }
        if (!adds && set_by_derive_.contains(target)) {
            refuse(s, "field '" + s[1].text + "' is derived twice with '=' (use '+=' to add to it)");
            return;
        }
        const NodePtr e = expression(s, adds ? 4 : 3, nullptr);
        if (!e) {
            return;
        }
        law.fields[static_cast<std::size_t>(target)].derived = true;
        if (!adds) {
            set_by_derive_.insert(target);
        }
        Derive d;
        d.target = target;
        d.accumulate = adds;
        d.where = s.where;
        d.rhs = e;
        d.reads = fields_read({e.get()});
        law.derives.push_back(std::move(d));
    }

    // A step limit: one number that does not read the step it sets.
    NodePtr step_limit(const Statement& s, std::size_t at, bool learned)
    {
        static constexpr const char* kLearnedLimit = "learned max_dt must be a number that does not read dt";
        NodePtr e = learned ? expression(s, at, nullptr, kLearnedLimit) : expression(s, at, nullptr);
        if (!e) {
            return nullptr;
        }
        if (learned && (!always_scalar(*e) || mentions_dt(*e))) {
            refuse(s, kLearnedLimit);
            return nullptr;
        }
        if (!always_scalar(*e)) {
            refuse(s, "max_dt must be a number: use maxof/minof/sumof to reduce a field");
            return nullptr;
        }
        if (mentions_dt(*e)) {
            refuse(s, "max_dt sets the step: it cannot read dt");
            return nullptr;
        }
        return e;
    }

    void define_learned(const Statement& s)
    {
        Term* t = inside_term(s, "learned");
        if (t == nullptr) {
            return;
        }
        if (s.word(1, "d")) {
            const int target = s.word(2) ? law.field_id(s[2].text) : -1;
            if (target < 0 || !s.punct(3, '=')) {
                refuse(s, "expected 'learned d FIELD = expression'");
                return;
            }
            if (NodePtr e = expression(s, 4, &lets_)) {
                t->learned.push_back(Equation{target, std::move(e)});
            }
        } else if (s.word(1, "max_dt") && s.punct(2, '=')) {
            if (NodePtr e = step_limit(s, 3, true)) {
                t->learned_max_dt = std::move(e);
            }
        } else {
            refuse(s, "expected 'learned d FIELD = expression' or 'learned max_dt = expression'");
        }
    }

    void define_rate(const Statement& s)
    {
        Term* t = inside_term(s, "d");
        if (t == nullptr) {
            return;
        }
        const int target = s.word(1) ? law.field_id(s[1].text) : -1;
        if (target < 0 || !s.punct(2, '=')) {
            refuse(s, "expected 'd FIELD = expression' with a declared field");
            return;
        }
        if (NodePtr e = expression(s, 3, &lets_)) {
            t->equations.push_back(Equation{target, std::move(e)});
        }
    }

    void define_conserves(const Statement& s)
    {
        Term* t = inside_term(s, "conserves");
        if (t == nullptr) {
            return;
        }
        // `a + b` is one group (their summed total); a name after another without `+` opens a new group.
        bool joined = false;
        for (std::size_t i = 1; !s.end(i); ++i) {
            if (s.punct(i, '+')) {
                joined = true;
                continue;
            }
            const int f = s.word(i) ? law.field_id(s[i].text) : -1;
            if (f < 0) {
                refuse(s, "conserves: '" + s[i].text + "' is not a field");
                continue;
            }
            if (joined && !t->conserves.empty()) {
                t->conserves.back().push_back(f);
            } else {
                t->conserves.push_back({f});
            }
            joined = false;
        }
    }

    void define_max_dt(const Statement& s)
    {
        Term* t = inside_term(s, "max_dt");
        if (t == nullptr) {
            return;
        }
        if (!s.punct(1, '=')) {
            refuse(s, "expected 'max_dt = expression'");
            return;
        }
        if (NodePtr e = step_limit(s, 2, false)) {
            t->max_dt = std::move(e);
        }
    }

    std::vector<std::string>& errors_;
    std::vector<std::string> creatures_;
    bool in_creature_ = false;
    Term* term_ = nullptr;
    Lets lets_;
    std::set<int> set_by_derive_;
};

void Compiler::validate()
{
    for (Term& t : law.terms) {
        const auto at = [&](const std::string& m) { errors_.push_back(t.where + ": term '" + t.name + "' " + m); };
        std::set<int> writes;
        std::vector<const Node*> roots;
        for (const Equation& eq : t.equations) {
            roots.push_back(eq.rhs.get());
            writes.insert(eq.target);
            const FieldDecl& f = law.fields[static_cast<std::size_t>(eq.target)];
            if (f.is_static) {
                at("writes static field '" + f.name + "'");
            }
            if (f.derived) {
                at("writes derived field '" + f.name + "'");
            }
        }
        if (t.equations.empty()) {
            at("has no equations");
        }
        for (const Equation& eq : t.learned) {
            roots.push_back(eq.rhs.get());
            if (!writes.contains(eq.target)) {
                // The classical equations are the fallback, so they must cover what the learned ones write.
                at("has a learned equation for '" + law.fields[static_cast<std::size_t>(eq.target)].name +
                    "', which its own equations do not write");
            }
        }
        for (const std::vector<int>& group : t.conserves) {
            std::string names;
            for (const int f : group) {
                names += (names.empty() ? "" : " + ") + law.fields[static_cast<std::size_t>(f)].name;
            }
            if (std::none_of(group.begin(), group.end(), [&](int f) { return writes.contains(f); })) {
                at("conserves '" + names + "' but never writes it");
            }
        }
        t.reads = fields_read(roots);
        t.writes.assign(writes.begin(), writes.end());
    }
    // The derives of one field run together where it first appears: its '=' first, then its '+='s in source order.
    std::vector<Derive> ordered;
    std::set<int> placed;
    for (const Derive& first : law.derives) {
        if (!placed.insert(first.target).second) {
            continue;
        }
        for (const bool adds : {false, true}) {
            for (const Derive& d : law.derives) {
                if (d.target == first.target && d.accumulate == adds) {
                    ordered.push_back(d);
                }
            }
        }
    }
    law.derives = std::move(ordered);
}

NodePtr deep_copy(const NodePtr& n, std::map<const Node*, NodePtr>& done)
{
    if (!n) {
        return nullptr;
    }
    if (const auto it = done.find(n.get()); it != done.end()) {
        return it->second;
    }
    auto c = std::make_shared<Node>(*n);
    for (NodePtr& a : c->args) {
        a = deep_copy(a, done);
    }
    done.emplace(n.get(), c);
    return c;
}

} // namespace

int Law::field_id(std::string_view name) const
{
    const auto it = std::find_if(fields.begin(), fields.end(), [&](const FieldDecl& f) { return f.name == name; });
    return it == fields.end() ? -1 : static_cast<int>(it - fields.begin());
}

int Law::net_id(std::string_view name) const
{
    const auto it = std::find_if(nets.begin(), nets.end(), [&](const NetDecl& n) { return n.name == name; });
    return it == nets.end() ? -1 : static_cast<int>(it - nets.begin());
}

int Law::param_id(std::string_view name) const
{
    const auto it = std::find(param_names.begin(), param_names.end(), name);
    return it == param_names.end() ? -1 : static_cast<int>(it - param_names.begin());
}

int Law::creature_id(std::string_view name) const
{
    const auto it = std::find_if(creatures.begin(), creatures.end(), [&](const CreatureDef& c) { return c.name == name; });
    return it == creatures.end() ? -1 : static_cast<int>(it - creatures.begin());
}

std::optional<LawSource> read_law_file(const std::string& path)
{
    std::ifstream in(path);
    if (!in) {
        return std::nullopt;
    }
    std::ostringstream text;
    text << in.rdbuf();
    return LawSource{path, text.str()};
}

Law clone_law(const Law& law)
{
    Law copy = law;
    std::map<const Node*, NodePtr> done;
    const auto fix = [&](std::vector<Equation>& eqs) {
        for (Equation& eq : eqs) {
            eq.rhs = deep_copy(eq.rhs, done);
        }
    };
    for (Term& t : copy.terms) {
        fix(t.equations);
        fix(t.learned);
        t.max_dt = deep_copy(t.max_dt, done);
        t.learned_max_dt = deep_copy(t.learned_max_dt, done);
    }
    for (Derive& d : copy.derives) {
        d.rhs = deep_copy(d.rhs, done);
    }
    return copy;
}

CompileResult compile_law(const std::vector<LawSource>& sources)
{
    CompileResult result;
    std::vector<Statement> lines;
    for (const LawSource& src : sources) {
        std::istringstream text(src.text);
        std::string raw;
        for (int number = 1; std::getline(text, raw); ++number) {
            Statement s;
            s.where = src.name + ":" + std::to_string(number);
            std::string error;
            s.lex_ = scan(raw, error);
            if (!error.empty()) {
                result.errors.push_back(s.where + ": " + error);
            } else if (!s.empty()) {
                lines.push_back(std::move(s));
            }
        }
    }
    Compiler c(result.errors);
    // First pass: nets, fields and derived fields, so every name exists before any expression reads it.
    for (const Statement& s : lines) {
        c.declare(s);
    }
    for (const Statement& s : lines) {
        c.define(s);
    }
    c.validate();
    if (result.errors.empty()) {
        result.law = std::move(c.law);
    }
    return result;
}

} // namespace fieldengine::field
