// The law language front end. See Law.hpp for what it follows.
#include "field/Law.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <functional>
#include <map>
#include <set>
#include <sstream>

namespace fieldengine::field {

const char* to_string(Activation a)
{
    switch (a) {
    case Activation::Tanh: return "tanh";
    case Activation::Relu: return "relu";
    case Activation::Silu: return "silu";
    case Activation::Linear: return "linear";
    case Activation::Gelu: return "gelu";
    }
    return "?";
}

bool parse_activation(const std::string& text, Activation& out)
{
    static const std::map<std::string, Activation> table = {{"tanh", Activation::Tanh}, {"relu", Activation::Relu},
        {"silu", Activation::Silu}, {"linear", Activation::Linear}, {"gelu", Activation::Gelu}};
    const auto it = table.find(text);
    if (it == table.end()) {
        return false;
    }
    out = it->second;
    return true;
}

int NetDecl::widest() const
{
    int w = 0;
    for (const int s : sizes) {
        w = std::max(w, s);
    }
    return w;
}

std::size_t NetDecl::weight_offset(int layer) const
{
    std::size_t at = 0;
    for (int l = 0; l < layer; ++l) {
        at += static_cast<std::size_t>(sizes[static_cast<std::size_t>(l)] + 1) * static_cast<std::size_t>(sizes[static_cast<std::size_t>(l) + 1]);
    }
    return at;
}

std::size_t NetDecl::parameter_count() const
{
    return weight_offset(layers());
}

int Law::field_id(std::string_view name) const
{
    for (std::size_t i = 0; i < fields.size(); ++i) {
        if (fields[i].name == name) {
            return static_cast<int>(i);
        }
    }
    return -1;
}

int Law::net_id(std::string_view name) const
{
    for (std::size_t i = 0; i < nets.size(); ++i) {
        if (nets[i].name == name) {
            return static_cast<int>(i);
        }
    }
    return -1;
}

int Law::param_id(std::string_view name) const
{
    for (std::size_t i = 0; i < param_names.size(); ++i) {
        if (param_names[i] == name) {
            return static_cast<int>(i);
        }
    }
    return -1;
}

namespace {

NodePtr make(Op op, std::vector<NodePtr> args = {})
{
    auto n = std::make_shared<Node>();
    n->op = op;
    n->args = std::move(args);
    return n;
}

NodePtr constant(double v)
{
    auto n = make(Op::Const);
    n->number = v;
    return n;
}

void collect_reads(const Node& node, std::set<int>& out)
{
    if (node.op == Op::Field) {
        out.insert(node.index);
    }
    for (const NodePtr& a : node.args) {
        collect_reads(*a, out);
    }
}

bool reads_dt(const Node& node)
{
    if (node.op == Op::Dt) {
        return true;
    }
    return std::any_of(node.args.begin(), node.args.end(), [](const NodePtr& a) { return reads_dt(*a); });
}

// ---- tokens ---------------------------------------------------------------------------------------------------------

struct Token {
    enum Kind { Number, Name, Symbol, End } kind = End;
    std::string text;
    double number = 0.0;
};

bool is_digit(char c) { return std::isdigit(static_cast<unsigned char>(c)) != 0; }

std::vector<Token> tokenize(std::string_view line, std::string& error)
{
    std::vector<Token> out;
    std::size_t i = 0;
    while (i < line.size()) {
        const char c = line[i];
        if (c == '#') {
            break;
        }
        if (std::isspace(static_cast<unsigned char>(c))) {
            ++i;
            continue;
        }
        if (is_digit(c) || (c == '.' && i + 1 < line.size() && is_digit(line[i + 1]))) {
            std::size_t j = i;
            while (j < line.size() && (is_digit(line[j]) || line[j] == '.')) {
                ++j;
            }
            if (j < line.size() && (line[j] == 'e' || line[j] == 'E')) {
                std::size_t k = j + 1;
                if (k < line.size() && (line[k] == '+' || line[k] == '-')) {
                    ++k;
                }
                if (k < line.size() && is_digit(line[k])) {
                    j = k;
                    while (j < line.size() && is_digit(line[j])) {
                        ++j;
                    }
                }
            }
            Token t{Token::Number, std::string(line.substr(i, j - i)), 0.0};
            char* end = nullptr;
            t.number = std::strtod(t.text.c_str(), &end);
            if (end == t.text.c_str() || *end != '\0') {
                error = "bad number '" + t.text + "'";
                return {};
            }
            out.push_back(t);
            i = j;
            continue;
        }
        if (std::isalpha(static_cast<unsigned char>(c)) || c == '_') {
            std::size_t j = i;
            while (j < line.size() && (std::isalnum(static_cast<unsigned char>(line[j])) || line[j] == '_')) {
                ++j;
            }
            out.push_back(Token{Token::Name, std::string(line.substr(i, j - i)), 0.0});
            i = j;
            continue;
        }
        if (std::string_view("+-*/^(),=").find(c) != std::string_view::npos) {
            out.push_back(Token{Token::Symbol, std::string(1, c), 0.0});
            ++i;
            continue;
        }
        error = std::string("unexpected character '") + c + "'";
        return {};
    }
    out.push_back(Token{});
    return out;
}

// ---- expressions ----------------------------------------------------------------------------------------------------

struct FunctionSpec {
    Op op;
    int min_args;
    int max_args;
};

const std::map<std::string, FunctionSpec, std::less<>>& functions()
{
    // Names and arities of the language's functions (src/field/Law.cpp:126-143). The first argument of mlp, flux and
    // ldiffuse is a net's name.
    static const std::map<std::string, FunctionSpec, std::less<>> table = {
        {"abs", {Op::Abs, 1, 1}}, {"sqrt", {Op::Sqrt, 1, 1}}, {"exp", {Op::Exp, 1, 1}}, {"log", {Op::Log, 1, 1}},
        {"step", {Op::Step, 1, 1}}, {"pos", {Op::Pos, 1, 1}}, {"min", {Op::Min, 2, 2}}, {"max", {Op::Max, 2, 2}},
        {"clamp", {Op::Clamp, 3, 3}}, {"lap", {Op::Lap, 1, 1}}, {"gx", {Op::Gx, 1, 1}}, {"gy", {Op::Gy, 1, 1}},
        {"gz", {Op::Gz, 1, 1}}, {"diffuse", {Op::Diffuse, 2, 2}}, {"adv", {Op::Adv, 4, 4}},
        {"distance", {Op::Distance, 1, 1}}, {"poisson", {Op::Poisson, 1, 2}}, {"maxof", {Op::MaxOf, 1, 1}},
        {"minof", {Op::MinOf, 1, 1}}, {"sumof", {Op::SumOf, 1, 1}}, {"fall", {Op::Fall, 3, 3}},
        {"spread", {Op::Spread, 3, 3}}, {"topple", {Op::Topple, 3, 3}}, {"shift", {Op::Shift, 4, 4}},
        {"shift_t", {Op::ShiftT, 4, 4}}, {"gx_t", {Op::GxT, 1, 1}}, {"gy_t", {Op::GyT, 1, 1}}, {"gz_t", {Op::GzT, 1, 1}},
        {"tridiag_x", {Op::Tridiag, 4, 4}}, {"tridiag_y", {Op::Tridiag, 4, 4}}, {"tridiag_z", {Op::Tridiag, 4, 4}},
        {"mlp", {Op::Mlp, 2, 33}}, {"flux", {Op::Flux, 2, 9}}, {"ldiffuse", {Op::LDiffuse, 3, 3}},
    };
    return table;
}

class Parser {
public:
    Parser(const std::vector<Token>& t, std::size_t start, const Law& law,
        const std::map<std::string, NodePtr, std::less<>>* lets = nullptr)
        : t_(t), pos_(start), law_(law), lets_(lets)
    {
    }

    NodePtr expression()
    {
        NodePtr left = product();
        while (ok() && (peek("+") || peek("-"))) {
            const bool add = take().text == "+";
            NodePtr right = product();
            left = make(add ? Op::Add : Op::Sub, {left, right});
        }
        return left;
    }

    [[nodiscard]] bool at_end() const { return t_[pos_].kind == Token::End; }
    [[nodiscard]] bool ok() const { return error_.empty(); }
    [[nodiscard]] const std::string& error() const { return error_; }

private:
    NodePtr product()
    {
        NodePtr left = unary();
        while (ok() && (peek("*") || peek("/"))) {
            const bool mul = take().text == "*";
            NodePtr right = unary();
            left = make(mul ? Op::Mul : Op::Div, {left, right});
        }
        return left;
    }

    NodePtr unary()
    {
        if (peek("-")) {
            take();
            return make(Op::Neg, {unary()});
        }
        if (peek("+")) {
            take();
            return unary();
        }
        return power();
    }

    NodePtr power()
    {
        NodePtr base = primary();
        if (ok() && peek("^")) {
            take();
            return make(Op::Pow, {base, unary()}); // right-associative, and -x binds looser than ^
        }
        return base;
    }

    NodePtr primary()
    {
        const Token& tok = t_[pos_];
        if (tok.kind == Token::Number) {
            ++pos_;
            return constant(tok.number);
        }
        if (tok.kind == Token::Symbol && tok.text == "(") {
            ++pos_;
            NodePtr inner = expression();
            expect(")");
            return inner;
        }
        if (tok.kind == Token::Name) {
            ++pos_;
            return peek("(") ? call(tok.text) : name(tok.text);
        }
        fail(tok.kind == Token::End ? "expression ends too early" : "unexpected '" + tok.text + "'");
        return constant(0.0);
    }

    NodePtr call(const std::string& fn)
    {
        take(); // (
        std::vector<NodePtr> args;
        int net = -1;
        const bool netted = fn == "mlp" || fn == "flux" || fn == "ldiffuse";
        if (netted) {
            const Token& nm = t_[pos_];
            net = nm.kind == Token::Name ? law_.net_id(nm.text) : -1;
            if (net < 0) {
                fail("'" + fn + "' takes the name of a declared net first");
                return constant(0.0);
            }
            ++pos_;
            if (!peek(",")) {
                fail("expected ',' after the net's name");
                return constant(0.0);
            }
            take();
        }
        if (!peek(")")) {
            args.push_back(expression());
            while (ok() && peek(",")) {
                take();
                args.push_back(expression());
            }
        }
        expect(")");
        const auto found = functions().find(fn);
        if (found == functions().end()) {
            fail("unknown function '" + fn + "'");
            return constant(0.0);
        }
        const FunctionSpec& spec = found->second;
        const int count = static_cast<int>(args.size()) + (netted ? 1 : 0);
        if (count < spec.min_args || count > spec.max_args) {
            fail("'" + fn + "' takes " + std::to_string(spec.min_args) +
                (spec.max_args != spec.min_args ? "-" + std::to_string(spec.max_args) : "") + " argument(s), got " +
                std::to_string(count));
            return constant(0.0);
        }
        NodePtr node = make(spec.op);
        if (spec.op == Op::Tridiag) {
            node->index = fn.back() - 'x';
        }
        if (netted) {
            node->index = net;
            const NetDecl& decl = law_.nets[static_cast<std::size_t>(net)];
            const int given = static_cast<int>(args.size());
            const int wanted = spec.op == Op::Mlp ? given : (spec.op == Op::Flux ? 2 * given : 5);
            if (decl.inputs() != wanted) {
                fail("net '" + decl.name + "' takes " + std::to_string(decl.inputs()) + " inputs; " + fn + " here gives it " +
                    std::to_string(wanted));
                return constant(0.0);
            }
            if (spec.op == Op::LDiffuse && decl.outputs() != 4) {
                fail("ldiffuse's net gives 4 weights (one per cell of the face's window), '" + decl.name + "' gives " +
                    std::to_string(decl.outputs()));
                return constant(0.0);
            }
            const std::size_t moved = spec.op == Op::LDiffuse ? 1 : 0;
            if ((spec.op == Op::Flux || spec.op == Op::LDiffuse) && args[moved]->op != Op::Field) {
                fail(fn + " moves a field: its " + (moved == 0 ? "first" : "second") + " argument after the net is a field's name");
                return constant(0.0);
            }
            if (spec.op == Op::Flux) {
                // The limiter keeps the moved field within its declared range: carried as two trailing constants.
                const FieldDecl& f = law_.fields[static_cast<std::size_t>(args[0]->index)];
                args.push_back(constant(f.lo));
                args.push_back(constant(f.hi));
            }
            node->args = std::move(args);
            return node;
        }
        if (spec.op == Op::Poisson) {
            node->args = {args[0]};
            node->iterations = 24;
            if (args.size() == 2) {
                if (args[1]->op != Op::Const || args[1]->number < 1.0) {
                    fail("poisson's iteration count must be a positive number");
                } else {
                    node->iterations = static_cast<int>(args[1]->number);
                }
            }
            return node;
        }
        if (spec.op == Op::Shift || spec.op == Op::ShiftT) {
            for (std::size_t a = 1; a < 4; ++a) {
                double sign = 1.0;
                const Node* arg = args[a].get();
                if (arg->op == Op::Neg) {
                    sign = -1.0;
                    arg = arg->args[0].get();
                }
                if (arg->op != Op::Const || arg->number != std::floor(arg->number) || std::abs(arg->number) > 8.0) {
                    fail("shift's offsets must be whole numbers of cells (at most 8)");
                    return constant(0.0);
                }
                args[a] = constant(sign * arg->number);
            }
        }
        node->args = std::move(args);
        return node;
    }

    NodePtr name(const std::string& text)
    {
        if (text == "h") {
            return make(Op::Cell);
        }
        if (text == "dt") {
            return make(Op::Dt);
        }
        if (text == "rim") {
            return make(Op::Rim);
        }
        if (text == "pi") {
            return constant(3.14159265358979323846);
        }
        if (lets_ != nullptr) {
            if (const auto found = lets_->find(text); found != lets_->end()) {
                return found->second;
            }
        }
        if (const int p = law_.param_id(text); p >= 0) {
            auto n = make(Op::Param);
            n->index = p;
            return n;
        }
        if (const int f = law_.field_id(text); f >= 0) {
            auto n = make(Op::Field);
            n->index = f;
            return n;
        }
        fail("unknown name '" + text + "' (not a field, param, h or pi)");
        return constant(0.0);
    }

    [[nodiscard]] bool peek(std::string_view s) const { return t_[pos_].kind == Token::Symbol && t_[pos_].text == s; }
    const Token& take() { return t_[pos_++]; }
    void expect(std::string_view s)
    {
        if (!peek(s)) {
            fail("expected '" + std::string(s) + "'");
            return;
        }
        ++pos_;
    }
    void fail(const std::string& m)
    {
        if (error_.empty()) {
            error_ = m;
        }
    }

    const std::vector<Token>& t_;
    std::size_t pos_;
    const Law& law_;
    const std::map<std::string, NodePtr, std::less<>>* lets_;
    std::string error_;
};

bool is_reserved(std::string_view name)
{
    static const std::set<std::string, std::less<>> words = {"field", "param", "term", "derive", "creature", "show", "d",
        "conserves", "max_dt", "h", "dt", "pi", "inf", "init", "range", "unit", "static", "clamp", "net", "learned", "plan",
        "observe", "let"};
    return words.contains(name);
}

// A bound in a declaration: a number, optionally negative, or inf / -inf.
bool parse_bound(const std::vector<Token>& t, std::size_t& i, double& out)
{
    double sign = 1.0;
    if (t[i].kind == Token::Symbol && t[i].text == "-") {
        sign = -1.0;
        ++i;
    }
    if (t[i].kind == Token::Name && t[i].text == "inf") {
        out = sign * 1e30;
        ++i;
        return true;
    }
    if (t[i].kind != Token::Number) {
        return false;
    }
    out = sign * t[i].number;
    ++i;
    return true;
}

struct Line {
    std::string where;
    std::vector<Token> tokens;
};

bool top_level(const std::string& word)
{
    return word == "field" || word == "param" || word == "term" || word == "derive" || word == "show" || word == "net" ||
        word == "creature";
}

} // namespace

bool always_scalar(const Node& node)
{
    switch (node.op) {
    case Op::Const: case Op::Param: case Op::Cell: case Op::MaxOf: case Op::MinOf: case Op::SumOf: case Op::Dt:
        return true;
    case Op::Flux: case Op::LDiffuse: case Op::Field: case Op::Distance: case Op::Poisson: case Op::Fall: case Op::Spread:
    case Op::Topple: case Op::Tridiag: case Op::ShiftT: case Op::GxT: case Op::GyT: case Op::GzT: case Op::Rim:
        return false;
    default:
        return std::all_of(node.args.begin(), node.args.end(), [](const NodePtr& a) { return always_scalar(*a); });
    }
}

std::optional<double> evaluate_scalar(const Node& n, const std::vector<double>& params, double h, double dt)
{
    std::vector<double> a;
    for (const NodePtr& arg : n.args) {
        const std::optional<double> v = evaluate_scalar(*arg, params, h, dt);
        if (!v) {
            return std::nullopt;
        }
        a.push_back(*v);
    }
    switch (n.op) {
    case Op::Const: return n.number;
    case Op::Param: return params[static_cast<std::size_t>(n.index)];
    case Op::Cell: return h;
    case Op::Dt: return dt;
    case Op::Neg: return -a[0];
    case Op::Abs: return std::abs(a[0]);
    case Op::Sqrt: return std::sqrt(0.0 < a[0] ? a[0] : 0.0); // std::max(0.0, a): a NaN becomes 0
    case Op::Exp: return std::exp(80.0 < a[0] ? 80.0 : a[0]);
    case Op::Log: return std::log(a[0] < 1e-30 ? 1e-30 : a[0]);
    case Op::Step: return a[0] > 0.0 ? 1.0 : 0.0;
    case Op::Pos: return a[0] < 0.0 ? 0.0 : a[0];
    case Op::Add: return a[0] + a[1];
    case Op::Sub: return a[0] - a[1];
    case Op::Mul: return a[0] == 0.0 || a[1] == 0.0 ? 0.0 : a[0] * a[1];
    case Op::Div: return a[1] == 0.0 ? 0.0 : a[0] / a[1];
    case Op::Pow: return std::pow(a[0], a[1]);
    case Op::Min: return a[1] < a[0] ? a[1] : a[0];
    case Op::Max: return a[0] < a[1] ? a[1] : a[0];
    case Op::Clamp: {
        const double hi = a[1] < a[2] ? a[2] : a[1];
        return a[0] < a[1] ? a[1] : (hi < a[0] ? hi : a[0]);
    }
    case Op::Shift: return a[0];
    case Op::Lap: case Op::Gx: case Op::Gy: case Op::Gz: return 0.0;
    default: return std::nullopt;
    }
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

CompileResult compile_law(const std::vector<LawSource>& sources)
{
    CompileResult result;
    Law law;
    std::vector<Line> lines;
    for (const LawSource& src : sources) {
        std::istringstream stream(src.text);
        std::string raw;
        int number = 0;
        while (std::getline(stream, raw)) {
            ++number;
            const std::string where = src.name + ":" + std::to_string(number);
            std::string error;
            std::vector<Token> tokens = tokenize(raw, error);
            if (!error.empty()) {
                result.errors.push_back(where + ": " + error);
                continue;
            }
            if (tokens.size() > 1) {
                lines.push_back(Line{where, std::move(tokens)});
            }
        }
    }

    // Pass 1: declarations (fields, nets, fields a derive introduces), so terms may use names declared later.
    bool in_creature = false;
    for (const Line& line : lines) {
        const std::vector<Token>& t = line.tokens;
        if (t[0].kind != Token::Name) {
            continue;
        }
        if (top_level(t[0].text)) {
            in_creature = t[0].text == "creature";
        }
        if (in_creature) {
            continue;
        }
        if (t[0].text == "net") {
            if (t[1].kind != Token::Name || is_reserved(t[1].text) || law.net_id(t[1].text) >= 0) {
                result.errors.push_back(line.where + ": expected 'net NAME in N hidden H... out M act tanh|relu|silu|gelu [half] [file PATH]'");
                continue;
            }
            NetDecl net;
            net.name = t[1].text;
            net.where = line.where;
            net.file = "assets/nets/" + net.name + ".net";
            std::vector<int> hidden;
            int in = 0, out = 0;
            bool bad = false;
            std::size_t i = 2;
            std::string section;
            while (t[i].kind != Token::End && !bad) {
                const Token& k = t[i++];
                if (k.kind == Token::Number) {
                    if (section == "in") {
                        in = static_cast<int>(k.number);
                    } else if (section == "hidden") {
                        hidden.push_back(static_cast<int>(k.number));
                    } else if (section == "out") {
                        out = static_cast<int>(k.number);
                    } else {
                        bad = true;
                    }
                    continue;
                }
                section = k.text;
                if (section == "act") {
                    bad = t[i].kind != Token::Name || !parse_activation(t[i].text, net.act);
                    ++i;
                } else if (section == "half") {
                    net.half = true;
                } else if (section == "file") {
                    std::string path;
                    while (t[i].kind != Token::End && !(t[i].kind == Token::Name && (t[i].text == "half" || t[i].text == "act"))) {
                        path += t[i].kind == Token::Number ? std::to_string(static_cast<long long>(t[i].number)) : t[i].text;
                        ++i;
                    }
                    bad = path.empty();
                    net.file = path;
                } else if (section != "in" && section != "hidden" && section != "out") {
                    bad = true;
                }
            }
            if (bad || in < 1 || out < 1 || in > 64 || out > 64 || hidden.size() > 6 ||
                std::any_of(hidden.begin(), hidden.end(), [](int h) { return h < 1 || h > 128; })) {
                result.errors.push_back(line.where + ": malformed net (in 1..64, hidden up to 6 layers of 1..128, out 1..64)");
                continue;
            }
            net.sizes.push_back(in);
            net.sizes.insert(net.sizes.end(), hidden.begin(), hidden.end());
            net.sizes.push_back(out);
            init_net(net);
            std::string why;
            if (!load_net(net, why) && why.rfind("no weights at", 0) != 0) {
                result.errors.push_back(line.where + ": " + why);
                continue;
            }
            law.nets.push_back(std::move(net));
            continue;
        }
        if (t[0].text == "field") {
            if (t[1].kind != Token::Name || is_reserved(t[1].text)) {
                result.errors.push_back(line.where + ": 'field' needs a name that is not a keyword");
                continue;
            }
            if (law.field_id(t[1].text) >= 0) {
                result.errors.push_back(line.where + ": field '" + t[1].text + "' declared twice");
                continue;
            }
            FieldDecl decl;
            decl.name = t[1].text;
            std::size_t i = 2;
            bool bad = false;
            while (t[i].kind != Token::End && !bad) {
                const std::string key = t[i].text;
                ++i;
                if (key == "init") {
                    bad = !parse_bound(t, i, decl.init);
                } else if (key == "range") {
                    bad = !parse_bound(t, i, decl.lo) || !parse_bound(t, i, decl.hi) || decl.lo > decl.hi;
                } else if (key == "unit") {
                    bad = t[i].kind != Token::Name;
                    decl.unit = t[i].text;
                    ++i;
                } else if (key == "static") {
                    decl.is_static = true;
                } else if (key == "clamp") {
                    decl.clamp = true;
                } else {
                    bad = true;
                }
            }
            if (bad) {
                result.errors.push_back(line.where + ": malformed field declaration (expected init N, range LO HI, unit U, static, clamp)");
                continue;
            }
            law.fields.push_back(decl);
        } else if (t[0].text == "derive" && t[1].kind == Token::Name && law.field_id(t[1].text) < 0 && !is_reserved(t[1].text)) {
            FieldDecl decl;
            decl.name = t[1].text;
            decl.derived = true;
            law.fields.push_back(decl);
        }
    }

    // Pass 2: params in order, terms with their lets, derives.
    Term* current = nullptr;
    std::map<std::string, NodePtr, std::less<>> lets;
    std::set<int> derived_targets;
    in_creature = false;
    for (const Line& line : lines) {
        const std::vector<Token>& t = line.tokens;
        auto fail = [&](const std::string& m) { result.errors.push_back(line.where + ": " + m); };
        if (t[0].kind != Token::Name) {
            fail("a line starts with a keyword (field, param, term, let, derive, d, conserves, max_dt)");
            continue;
        }
        const std::string& kw = t[0].text;
        if (top_level(kw)) {
            in_creature = kw == "creature";
            if (kw != "term") {
                current = nullptr;
            }
        }
        if (in_creature || kw == "show" || kw == "net" || kw == "field") {
            continue; // creatures and pictures are not the lattice's; fields and nets were pass 1's
        }
        if (kw == "param") {
            if (t[1].kind != Token::Name || is_reserved(t[1].text) || !(t[2].kind == Token::Symbol && t[2].text == "=")) {
                fail("expected 'param NAME = expression'");
                continue;
            }
            if (law.param_id(t[1].text) >= 0 || law.field_id(t[1].text) >= 0) {
                fail("'" + t[1].text + "' is already a param or a field");
                continue;
            }
            Parser p(t, 3, law);
            NodePtr e = p.expression();
            if (!p.ok() || !p.at_end()) {
                fail(p.ok() ? "trailing tokens after the expression" : p.error());
                continue;
            }
            if (!always_scalar(*e)) {
                fail("a param cannot read a field");
                continue;
            }
            law.param_names.push_back(t[1].text);
            law.param_values.push_back(evaluate_scalar(*e, law.param_values, 1.0, 0.0).value_or(0.0));
            continue;
        }
        if (kw == "term") {
            if (t[1].kind != Token::Name || t[2].kind != Token::End) {
                fail("expected 'term NAME'");
                current = nullptr;
                continue;
            }
            for (const Term& existing : law.terms) {
                if (existing.name == t[1].text) {
                    fail("term '" + t[1].text + "' declared twice");
                }
            }
            law.terms.push_back(Term{});
            law.terms.back().name = t[1].text;
            law.terms.back().where = line.where;
            current = &law.terms.back();
            lets.clear();
            continue;
        }
        if (kw == "let") {
            if (current == nullptr) {
                fail("'let' outside a term");
                continue;
            }
            if (t[1].kind != Token::Name || !(t[2].kind == Token::Symbol && t[2].text == "=")) {
                fail("expected 'let NAME = expression'");
                continue;
            }
            const std::string& nm = t[1].text;
            if (is_reserved(nm) || law.field_id(nm) >= 0 || law.param_id(nm) >= 0 || lets.contains(nm)) {
                fail("'" + nm + "' is already a keyword, field, param or let");
                continue;
            }
            Parser p(t, 3, law, &lets);
            NodePtr e = p.expression();
            if (!p.ok() || !p.at_end()) {
                fail(p.ok() ? "trailing tokens after the expression" : p.error());
                continue;
            }
            lets.emplace(nm, e);
            continue;
        }
        if (kw == "derive") {
            current = nullptr;
            const int target = t[1].kind == Token::Name ? law.field_id(t[1].text) : -1;
            const bool accumulate = t[2].kind == Token::Symbol && t[2].text == "+" && t[3].kind == Token::Symbol && t[3].text == "=";
            if (target < 0 || !(accumulate || (t[2].kind == Token::Symbol && t[2].text == "="))) {
                fail("expected 'derive FIELD = expression' or 'derive FIELD += expression'");
                continue;
            }
            if (!accumulate && derived_targets.contains(target)) {
                fail("field '" + t[1].text + "' is derived twice with '=' (use '+=' to add to it)");
                continue;
            }
            Parser p(t, accumulate ? 4 : 3, law);
            NodePtr e = p.expression();
            if (!p.ok() || !p.at_end()) {
                fail(p.ok() ? "trailing tokens after the expression" : p.error());
                continue;
            }
            law.fields[static_cast<std::size_t>(target)].derived = true;
            if (!accumulate) {
                derived_targets.insert(target);
            }
            std::set<int> reads;
            collect_reads(*e, reads);
            Derive d;
            d.target = target;
            d.accumulate = accumulate;
            d.where = line.where;
            d.rhs = e;
            d.reads.assign(reads.begin(), reads.end());
            law.derives.push_back(std::move(d));
            continue;
        }
        if (kw == "learned") {
            if (current == nullptr) {
                fail("'learned' outside a term");
                continue;
            }
            if (t[1].kind == Token::Name && t[1].text == "d") {
                const int target = t[2].kind == Token::Name ? law.field_id(t[2].text) : -1;
                if (target < 0 || !(t[3].kind == Token::Symbol && t[3].text == "=")) {
                    fail("expected 'learned d FIELD = expression'");
                    continue;
                }
                Parser p(t, 4, law, &lets);
                NodePtr e = p.expression();
                if (!p.ok() || !p.at_end()) {
                    fail(p.ok() ? "trailing tokens after the expression" : p.error());
                    continue;
                }
                current->learned.push_back(Equation{target, e});
            } else if (t[1].kind == Token::Name && t[1].text == "max_dt" && t[2].kind == Token::Symbol && t[2].text == "=") {
                Parser p(t, 3, law);
                NodePtr e = p.expression();
                if (!p.ok() || !p.at_end() || !always_scalar(*e) || reads_dt(*e)) {
                    fail(p.ok() ? "learned max_dt must be a number that does not read dt" : p.error());
                    continue;
                }
                current->learned_max_dt = e;
            } else {
                fail("expected 'learned d FIELD = expression' or 'learned max_dt = expression'");
            }
            continue;
        }
        if (kw == "d" || kw == "conserves" || kw == "max_dt") {
            if (current == nullptr) {
                fail("'" + kw + "' outside a term");
                continue;
            }
            if (kw == "d") {
                const int target = t[1].kind == Token::Name ? law.field_id(t[1].text) : -1;
                if (target < 0 || !(t[2].kind == Token::Symbol && t[2].text == "=")) {
                    fail("expected 'd FIELD = expression' with a declared field");
                    continue;
                }
                Parser p(t, 3, law, &lets);
                NodePtr e = p.expression();
                if (!p.ok() || !p.at_end()) {
                    fail(p.ok() ? "trailing tokens after the expression" : p.error());
                    continue;
                }
                current->equations.push_back(Equation{target, e});
            } else if (kw == "conserves") {
                bool join = false;
                for (std::size_t i = 1; t[i].kind != Token::End; ++i) {
                    if (t[i].kind == Token::Symbol && t[i].text == "+") {
                        join = true;
                        continue;
                    }
                    const int f = t[i].kind == Token::Name ? law.field_id(t[i].text) : -1;
                    if (f < 0) {
                        fail("conserves: '" + t[i].text + "' is not a field");
                        continue;
                    }
                    if (join && !current->conserves.empty()) {
                        current->conserves.back().push_back(f);
                    } else {
                        current->conserves.push_back({f});
                    }
                    join = false;
                }
            } else {
                if (!(t[1].kind == Token::Symbol && t[1].text == "=")) {
                    fail("expected 'max_dt = expression'");
                    continue;
                }
                Parser p(t, 2, law);
                NodePtr e = p.expression();
                if (!p.ok() || !p.at_end()) {
                    fail(p.ok() ? "trailing tokens after the expression" : p.error());
                    continue;
                }
                if (!always_scalar(*e)) {
                    fail("max_dt must be a number: use maxof/minof/sumof to reduce a field");
                    continue;
                }
                if (reads_dt(*e)) {
                    fail("max_dt sets the step: it cannot read dt");
                    continue;
                }
                current->max_dt = e;
            }
            continue;
        }
        fail("unknown keyword '" + kw + "'");
    }

    // What each term reads and writes, and the contracts that would otherwise hold vacuously.
    for (Term& term : law.terms) {
        std::set<int> reads, writes;
        for (const Equation& eq : term.equations) {
            collect_reads(*eq.rhs, reads);
            writes.insert(eq.target);
            const FieldDecl& decl = law.fields[static_cast<std::size_t>(eq.target)];
            if (decl.is_static) {
                result.errors.push_back(term.where + ": term '" + term.name + "' writes static field '" + decl.name + "'");
            }
            if (decl.derived) {
                result.errors.push_back(term.where + ": term '" + term.name + "' writes derived field '" + decl.name + "'");
            }
        }
        if (term.equations.empty()) {
            result.errors.push_back(term.where + ": term '" + term.name + "' has no equations");
        }
        for (const Equation& eq : term.learned) {
            collect_reads(*eq.rhs, reads);
            if (!writes.contains(eq.target)) {
                result.errors.push_back(term.where + ": term '" + term.name + "' has a learned equation for '" +
                    law.fields[static_cast<std::size_t>(eq.target)].name + "', which its own equations do not write");
            }
        }
        for (const std::vector<int>& group : term.conserves) {
            bool touched = false;
            std::string names;
            for (const int f : group) {
                touched = touched || writes.contains(f);
                names += (names.empty() ? "" : " + ") + law.fields[static_cast<std::size_t>(f)].name;
            }
            if (!touched) {
                result.errors.push_back(term.where + ": term '" + term.name + "' conserves '" + names + "' but never writes it");
            }
        }
        term.reads.assign(reads.begin(), reads.end());
        term.writes.assign(writes.begin(), writes.end());
    }

    // Derives of one field run together at the position it first appears: the '=' first, then the '+='s.
    {
        std::vector<int> order;
        for (const Derive& d : law.derives) {
            if (std::find(order.begin(), order.end(), d.target) == order.end()) {
                order.push_back(d.target);
            }
        }
        std::vector<Derive> grouped;
        for (const int target : order) {
            for (int pass = 0; pass < 2; ++pass) {
                for (const Derive& d : law.derives) {
                    if (d.target == target && d.accumulate == (pass == 1)) {
                        grouped.push_back(d);
                    }
                }
            }
        }
        law.derives = std::move(grouped);
    }

    if (result.errors.empty()) {
        result.law = std::move(law);
    }
    return result;
}

// ---- nets -----------------------------------------------------------------------------------------------------------

void init_net(NetDecl& net)
{
    // Glorot-uniform hidden layers from a hash of the net's name (no RNG state: the same law gives the same net), and a
    // zero last layer, so the output is exactly 0 until trained.
    net.weights.assign(net.parameter_count(), 0.0f);
    std::uint64_t seed = 0xcbf29ce484222325ull;
    for (const char c : net.name) {
        seed = (seed ^ static_cast<unsigned char>(c)) * 0x100000001b3ull;
    }
    const int layers = net.layers();
    for (int l = 0; l + 1 < layers; ++l) {
        const int ni = net.sizes[static_cast<std::size_t>(l)];
        const int no = net.sizes[static_cast<std::size_t>(l) + 1];
        const double limit = std::sqrt(6.0 / static_cast<double>(ni + no));
        float* w = net.weights.data() + net.weight_offset(l);
        for (int k = 0; k < ni * no; ++k) {
            std::uint64_t z = seed + 0x9e3779b97f4a7c15ull * static_cast<std::uint64_t>(l * 1000003 + k + 1);
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            z ^= z >> 31;
            const double u = static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0);
            w[k] = static_cast<float>((2.0 * u - 1.0) * limit);
        }
    }
    net.in_shift.assign(static_cast<std::size_t>(net.inputs()), 0.0f);
    net.in_scale.assign(static_cast<std::size_t>(net.inputs()), 1.0f);
    net.out_scale = 1.0f;
    net.loaded = false;
}

bool load_net(NetDecl& net, std::string& why)
{
    // The text format: sections introduced by a word (net NAME, sizes ..., act A, in_shift ..., in_scale ...,
    // out_scale S, weights ...), numbers separated by white space.
    std::ifstream in(net.file);
    if (!in) {
        why = "no weights at " + net.file;
        return false;
    }
    std::string word, section;
    std::vector<int> sizes;
    std::vector<float> shift, scale, weights;
    float out_scale = 1.0f;
    Activation act = net.act;
    while (in >> word) {
        if (word == "net") {
            in >> word;
            section.clear();
        } else if (word == "act") {
            in >> word;
            if (!parse_activation(word, act)) {
                why = net.file + ": unknown activation '" + word + "'";
                return false;
            }
            section.clear();
        } else if (word == "out_scale") {
            in >> out_scale;
            section.clear();
        } else if (word == "sizes" || word == "in_shift" || word == "in_scale" || word == "weights") {
            section = word;
        } else {
            char* end = nullptr;
            const double v = std::strtod(word.c_str(), &end);
            if (end == word.c_str() || *end != '\0') {
                why = net.file + ": unexpected '" + word + "'";
                return false;
            }
            if (section == "sizes") {
                sizes.push_back(static_cast<int>(v));
            } else if (section == "in_shift") {
                shift.push_back(static_cast<float>(v));
            } else if (section == "in_scale") {
                scale.push_back(static_cast<float>(v));
            } else if (section == "weights") {
                weights.push_back(static_cast<float>(v));
            } else {
                why = net.file + ": a number outside a section";
                return false;
            }
        }
    }
    if (sizes != net.sizes) {
        why = net.file + ": its sizes are not the declared net's";
        return false;
    }
    if (weights.size() != net.parameter_count()) {
        why = net.file + ": " + std::to_string(weights.size()) + " weights, the net has " + std::to_string(net.parameter_count());
        return false;
    }
    if (act != net.act) {
        why = net.file + ": trained with another activation than the law declares"; // src/field/nn/Net.cpp:324-327
        return false;
    }
    net.weights = std::move(weights);
    shift.resize(static_cast<std::size_t>(net.inputs()), 0.0f);
    scale.resize(static_cast<std::size_t>(net.inputs()), 1.0f);
    net.in_shift = std::move(shift);
    net.in_scale = std::move(scale);
    net.out_scale = out_scale;
    net.loaded = true;
    return true;
}

bool save_net(const NetDecl& net, const std::string& path, std::string& why)
{
    std::ofstream out(path);
    if (!out) {
        why = "cannot write " + path;
        return false;
    }
    out.precision(9);
    out << "net " << net.name << "\nsizes";
    for (const int s : net.sizes) {
        out << ' ' << s;
    }
    out << "\nact " << to_string(net.act) << "\nin_shift";
    for (const float v : net.in_shift) {
        out << ' ' << v;
    }
    out << "\nin_scale";
    for (const float v : net.in_scale) {
        out << ' ' << v;
    }
    out << "\nout_scale " << net.out_scale << "\nweights";
    for (std::size_t i = 0; i < net.weights.size(); ++i) {
        out << (i % 16 == 0 ? "\n" : " ") << net.weights[i];
    }
    out << "\n";
    return static_cast<bool>(out);
}

} // namespace fieldengine::field
