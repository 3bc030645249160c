// The .sdf loader subset and the object-program compiler.
//
// Loader: the statement forms of the scene language the render path's
// scenes use (target/docs/reference_language.md): let, for, if, import of
// the standard library, settings blocks (render, camera, sky gradient,
// light quad/point), materials matte/plastic/metal/glass, modifiers
// at/turn/scale/paint, with-blocks, joins merge/cut/clip with a radius and
// calls of the standard.sdf primitives (scenes/standard.sdf:1-43). What it
// leaves behind is the reference's flat list of placed and painted leaves
// under CSG nodes (source/sdf/scene.h:237-310).
//
// Compiler: one straight-line SSA program per object, as the reference
// compileNode/inlineLeaf/hardJoin/smoothJoin (source/sdf/scene.cpp:122-287)
// and the union split (scene.cpp:324-345, scene.h:897), then prune
// (source/sdf/program.cpp:338-407).
#include "scene.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <optional>
#include <stdexcept>

namespace kdn {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr size_t kFewestUnionSteps = 512;  // source/sdf/scene.h:897

// ---------------------------------------------------------------- lexer --
enum class Tok { End, Number, Ident, String, Colour, Sym };

struct Token {
    Tok kind;
    std::string text;
    double number = 0;
    int line = 1, column = 1;
};

struct LoadError : std::runtime_error {
    int line, column;
    LoadError(int l, int c, const std::string& m) : std::runtime_error(m), line(l), column(c) {}
};

std::vector<Token> lex(const std::string& s) {
    std::vector<Token> out;
    size_t i = 0;
    int line = 1, col = 1;
    auto adv = [&](size_t n) {
        for (size_t k = 0; k < n && i < s.size(); ++k, ++i) {
            if (s[i] == '\n') { ++line; col = 1; } else { ++col; }
        }
    };
    while (i < s.size()) {
        char ch = s[i];
        if (std::isspace((unsigned char)ch)) { adv(1); continue; }
        if (ch == '/' && i + 1 < s.size() && s[i + 1] == '/') {
            while (i < s.size() && s[i] != '\n') adv(1);
            continue;
        }
        if (ch == '/' && i + 1 < s.size() && s[i + 1] == '*') {
            adv(2);
            while (i + 1 < s.size() && !(s[i] == '*' && s[i + 1] == '/')) adv(1);
            adv(2);
            continue;
        }
        Token t;
        t.line = line;
        t.column = col;
        if (std::isdigit((unsigned char)ch) || (ch == '.' && i + 1 < s.size() && std::isdigit((unsigned char)s[i + 1]))) {
            size_t j = i;
            while (j < s.size() && (std::isdigit((unsigned char)s[j]) || s[j] == '.')) {
                if (s[j] == '.' && j + 1 < s.size() && s[j + 1] == '.') break;  // a range ".."
                ++j;
            }
            if (j < s.size() && (s[j] == 'e' || s[j] == 'E')) {
                ++j;
                if (j < s.size() && (s[j] == '+' || s[j] == '-')) ++j;
                while (j < s.size() && std::isdigit((unsigned char)s[j])) ++j;
            }
            t.kind = Tok::Number;
            t.text = s.substr(i, j - i);
            t.number = std::strtod(t.text.c_str(), nullptr);
            adv(j - i);
        } else if (std::isalpha((unsigned char)ch) || ch == '_') {
            size_t j = i;
            while (j < s.size() && (std::isalnum((unsigned char)s[j]) || s[j] == '_')) ++j;
            t.kind = Tok::Ident;
            t.text = s.substr(i, j - i);
            adv(j - i);
        } else if (ch == '"') {
            size_t j = i + 1;
            while (j < s.size() && s[j] != '"') ++j;
            t.kind = Tok::String;
            t.text = s.substr(i + 1, j - i - 1);
            adv(j - i + 1);
        } else if (ch == '#') {
            size_t j = i + 1;
            while (j < s.size() && std::isxdigit((unsigned char)s[j])) ++j;
            t.kind = Tok::Colour;
            t.text = s.substr(i + 1, j - i - 1);
            if (t.text.size() != 6) throw LoadError(line, col, "a colour is # and six hex digits");
            adv(j - i);
        } else {
            static const char* two[] = {"<=", ">=", "==", "!=", "&&", "||", ".."};
            t.kind = Tok::Sym;
            t.text = std::string(1, ch);
            for (const char* w : two) {
                if (s.compare(i, 2, w) == 0) { t.text = w; break; }
            }
            adv(t.text.size());
        }
        out.push_back(t);
    }
    Token end;
    end.kind = Tok::End;
    end.line = line;
    end.column = col;
    out.push_back(end);
    return out;
}

// --------------------------------------------------------------- values --
struct Value {
    enum Kind { Number, List, MaterialRef, Word } kind = Number;
    double number = 0;
    std::vector<Value> items;
    uint32_t material = 0;
    std::string word;
};

Value num(double v) { Value x; x.kind = Value::Number; x.number = v; return x; }
Value list(std::vector<Value> items) { Value x; x.kind = Value::List; x.items = std::move(items); return x; }

double srgbToLinear(double c) {
    return c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4);
}

// ------------------------------------------------------------ placement --
// A local→world affine map x_world = A x_local + b, and the distance factor.
struct Frame {
    double A[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double b[3] = {0, 0, 0};
    double stretch = 1;
    std::optional<uint32_t> material;
};

void invert3(const double m[3][3], double out[3][3]) {
    const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                       m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                       m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    const double inv = 1.0 / det;
    out[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * inv;
    out[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * inv;
    out[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * inv;
    out[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * inv;
    out[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * inv;
    out[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * inv;
    out[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * inv;
    out[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * inv;
    out[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * inv;
}

// Compose: frame ∘ (x ↦ M x + t).
Frame composed(const Frame& f, const double M[3][3], const double t[3]) {
    Frame g = f;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            g.A[i][j] = f.A[i][0] * M[0][j] + f.A[i][1] * M[1][j] + f.A[i][2] * M[2][j];
        }
        g.b[i] = f.b[i] + f.A[i][0] * t[0] + f.A[i][1] * t[1] + f.A[i][2] * t[2];
    }
    return g;
}

// An exact zero where a rotation entry is within rounding of zero keeps
// turns by quarter circles exact, so the compiler skips zero weights as
// the reference localCoordinate does (scene.cpp:111-120).
double cleaned(double v) { return std::abs(v) < 1e-15 ? 0.0 : v; }

// ------------------------------------------------------------ evaluator --
struct Scope {
    std::map<std::string, Value> names;
};

class Loader {
public:
    Loader(const std::vector<Token>& tokens, Scene& scene) : t_(tokens), scene_(scene) {}

    void run() {
        scopes_.emplace_back();
        set("pi", num(kPi));
        set("time", num(0));
        // bare words the settings take: axes, and a grade's curves
        for (const char* axis : {"x", "y", "z", "filmic", "clamped"}) {
            Value w;
            w.kind = Value::Word;
            w.word = axis;
            set(axis, w);
        }
        Frame top;
        while (peek().kind != Tok::End) {
            statement(top, nullptr);
        }
    }

private:
    const std::vector<Token>& t_;
    size_t p_ = 0;
    Scene& scene_;
    std::vector<Scope> scopes_;

    const Token& peek(int k = 0) const { return t_[std::min(p_ + k, t_.size() - 1)]; }
    const Token& next() { return t_[p_ < t_.size() - 1 ? p_++ : p_]; }
    [[noreturn]] void fail(const Token& at, const std::string& m) { throw LoadError(at.line, at.column, m); }
    bool isSym(const char* s, int k = 0) const { return peek(k).kind == Tok::Sym && peek(k).text == s; }
    bool isIdent(const char* s, int k = 0) const { return peek(k).kind == Tok::Ident && peek(k).text == s; }
    void expect(const char* s) {
        if (!isSym(s)) fail(peek(), std::string("expected '") + s + "'");
        next();
    }
    std::string ident() {
        if (peek().kind != Tok::Ident) fail(peek(), "expected a name");
        return next().text;
    }

    Value* find(const std::string& n) {
        for (size_t i = scopes_.size(); i-- > 0;) {
            auto it = scopes_[i].names.find(n);
            if (it != scopes_[i].names.end()) return &it->second;
        }
        return nullptr;
    }
    // `let` of a bound name changes that binding; a new name binds here.
    void set(const std::string& n, const Value& v) {
        if (Value* old = find(n)) { *old = v; return; }
        scopes_.back().names[n] = v;
    }

    // ------------------------------------------------------ expressions --
    Value expr() { return orExpr(); }
    Value orExpr() {
        Value a = andExpr();
        while (isSym("||")) { next(); Value b = andExpr(); a = num((truth(a) || truth(b)) ? 1 : 0); }
        return a;
    }
    Value andExpr() {
        Value a = cmpExpr();
        while (isSym("&&")) { next(); Value b = cmpExpr(); a = num((truth(a) && truth(b)) ? 1 : 0); }
        return a;
    }
    Value cmpExpr() {
        Value a = addExpr();
        for (;;) {
            const char* ops[] = {"<", "<=", ">", ">=", "==", "!="};
            int which = -1;
            for (int k = 0; k < 6; ++k) if (isSym(ops[k])) which = k;
            if (which < 0) return a;
            next();
            Value b = addExpr();
            double x = scalar(a), y = scalar(b);
            bool r = which == 0 ? x < y : which == 1 ? x <= y : which == 2 ? x > y : which == 3 ? x >= y : which == 4 ? x == y : x != y;
            a = num(r ? 1 : 0);
        }
    }
    Value addExpr() {
        Value a = mulExpr();
        while (isSym("+") || isSym("-")) {
            char op = next().text[0];
            Value b = mulExpr();
            a = elementwise(a, b, [op](double x, double y) { return op == '+' ? x + y : x - y; });
        }
        return a;
    }
    Value mulExpr() {
        Value a = unary();
        while (isSym("*") || isSym("/")) {
            char op = next().text[0];
            Value b = unary();
            a = elementwise(a, b, [op](double x, double y) { return op == '*' ? x * y : x / y; });
        }
        return a;
    }
    Value unary() {
        if (isSym("-")) { next(); Value a = unary(); return map1(a, [](double x) { return -x; }); }
        if (isSym("!")) { next(); Value a = unary(); return num(truth(a) ? 0 : 1); }
        return postfix();
    }
    Value postfix() {
        Value v = primary();
        for (;;) {
            if (isSym("[")) {
                const Token& at = next();
                size_t k = size_t(scalar(expr()));
                expect("]");
                if (v.kind != Value::List || k >= v.items.size()) fail(at, "index outside the list");
                v = v.items[k];
            } else if (isSym(".") && peek(1).kind == Tok::Ident) {
                const Token& at = next();
                std::string sw = next().text;
                std::vector<Value> picked;
                for (char c : sw) {
                    int k = c == 'x' ? 0 : c == 'y' ? 1 : c == 'z' ? 2 : c == 'w' ? 3 : -1;
                    if (k < 0 || v.kind != Value::List || size_t(k) >= v.items.size()) fail(at, "bad swizzle ." + sw);
                    picked.push_back(v.items[k]);
                }
                v = picked.size() == 1 ? picked[0] : list(picked);
            } else {
                return v;
            }
        }
    }
    Value primary() {
        const Token& tk = peek();
        if (tk.kind == Tok::Number) { next(); return num(tk.number); }
        if (tk.kind == Tok::Colour) {
            next();
            std::vector<Value> c;
            for (int k = 0; k < 3; ++k) {
                int byte = std::stoi(tk.text.substr(2 * k, 2), nullptr, 16);
                c.push_back(num(srgbToLinear(byte / 255.0)));
            }
            return list(c);
        }
        if (isSym("(")) { next(); Value v = expr(); expect(")"); return v; }
        if (isSym("[")) {
            next();
            std::vector<Value> items;
            while (!isSym("]")) {
                items.push_back(expr());
                if (isSym(",")) next(); else break;
            }
            expect("]");
            return list(items);
        }
        if (tk.kind == Tok::Ident) {
            next();
            if (isSym("(")) return call(tk);
            Value* v = find(tk.text);
            if (!v) fail(tk, "unknown name " + tk.text);
            return *v;
        }
        fail(tk, "expected a value");
    }

    struct Args {
        std::vector<Value> positional;
        std::map<std::string, Value> named;
    };
    Args args() {
        Args a;
        expect("(");
        while (!isSym(")")) {
            if (peek().kind == Tok::Ident && isSym("=", 1)) {
                std::string n = next().text;
                next();
                a.named[n] = expr();
            } else {
                a.positional.push_back(expr());
            }
            if (isSym(",")) next(); else break;
        }
        expect(")");
        return a;
    }

    static bool truth(const Value& v) { return v.kind == Value::Number ? v.number != 0 : true; }
    double scalar(const Value& v) {
        if (v.kind != Value::Number) fail(peek(), "expected a number");
        return v.number;
    }
    template <class F>
    Value map1(const Value& a, F f) {
        if (a.kind == Value::Number) return num(f(a.number));
        if (a.kind != Value::List) fail(peek(), "expected a number or a list");
        Value out = list({});
        for (const Value& x : a.items) out.items.push_back(map1(x, f));
        return out;
    }
    template <class F>
    Value elementwise(const Value& a, const Value& b, F f) {
        if (a.kind == Value::Number && b.kind == Value::Number) return num(f(a.number, b.number));
        if (a.kind == Value::List && b.kind == Value::List) {
            if (a.items.size() != b.items.size()) fail(peek(), "lists of different lengths");
            Value out = list({});
            for (size_t i = 0; i < a.items.size(); ++i) out.items.push_back(elementwise(a.items[i], b.items[i], f));
            return out;
        }
        if (a.kind == Value::List && b.kind == Value::Number) return map1(a, [&](double x) { return f(x, b.number); });
        if (a.kind == Value::Number && b.kind == Value::List) return map1(b, [&](double y) { return f(a.number, y); });
        fail(peek(), "expected numbers or lists");
    }
    std::vector<double> vec(const Value& v, size_t n, const Token& at) {
        if (v.kind == Value::Number) return std::vector<double>(n, v.number);
        if (v.kind != Value::List || v.items.size() != n) fail(at, "expected " + std::to_string(n) + " numbers");
        std::vector<double> out;
        for (const Value& x : v.items) {
            if (x.kind != Value::Number) fail(at, "expected numbers");
            out.push_back(x.number);
        }
        return out;
    }

    Value call(const Token& name) {
        const std::string& n = name.text;
        Args a = args();
        auto arg = [&](size_t i, const char* named, std::optional<Value> dflt = std::nullopt) -> Value {
            if (i < a.positional.size()) return a.positional[i];
            auto it = a.named.find(named);
            if (it != a.named.end()) return it->second;
            if (dflt) return *dflt;
            fail(name, n + " needs " + named);
        };
        // math built-ins, folded at load time as the reference evaluator
        // folds known values (source/syntax/evaluator.cpp:880-920)
        if (n == "sqrt" || n == "abs" || n == "floor" || n == "sin" || n == "cos" || n == "sign") {
            Value x = arg(0, "a");
            return map1(x, [&](double v) {
                if (n == "sqrt") return std::sqrt(v);
                if (n == "abs") return std::abs(v);
                if (n == "floor") return std::floor(v);
                if (n == "sin") return std::sin(v);
                if (n == "cos") return std::cos(v);
                return v > 0 ? 1.0 : v < 0 ? -1.0 : 0.0;
            });
        }
        if (n == "min" || n == "max" || n == "pow" || n == "atan2") {
            return elementwise(arg(0, "a"), arg(1, "b"), [&](double x, double y) {
                if (n == "min") return std::min(x, y);
                if (n == "max") return std::max(x, y);
                if (n == "pow") return std::pow(x, y);
                return std::atan2(x, y);
            });
        }
        if (n == "clamp") {
            Value lo = arg(1, "lo"), hi = arg(2, "hi");
            Value x = elementwise(arg(0, "x"), lo, [](double p, double q) { return std::max(p, q); });
            return elementwise(x, hi, [](double p, double q) { return std::min(p, q); });
        }
        if (n == "mix") {
            Value x = arg(0, "a"), y = arg(1, "b"), s = arg(2, "t");
            Value d = elementwise(y, x, [](double p, double q) { return p - q; });
            return elementwise(x, elementwise(d, s, [](double p, double q) { return p * q; }), [](double p, double q) { return p + q; });
        }
        if (n == "select") return truth(arg(0, "c")) ? arg(1, "a") : arg(2, "b");
        if (n == "length" || n == "dot" || n == "normalize" || n == "cross") {
            std::vector<double> x = vec(arg(0, "a"), arg(0, "a").kind == Value::List ? arg(0, "a").items.size() : 1, name);
            if (n == "length") {
                double s = 0;
                for (double v : x) s += v * v;
                return num(std::sqrt(s));
            }
            if (n == "normalize") {
                double s = 0;
                for (double v : x) s += v * v;
                s = std::sqrt(s);
                std::vector<Value> o;
                for (double v : x) o.push_back(num(v / s));
                return list(o);
            }
            std::vector<double> y = vec(arg(1, "b"), x.size(), name);
            if (n == "dot") {
                double s = 0;
                for (size_t i = 0; i < x.size(); ++i) s += x[i] * y[i];
                return num(s);
            }
            return list({num(x[1] * y[2] - x[2] * y[1]), num(x[2] * y[0] - x[0] * y[2]), num(x[0] * y[1] - x[1] * y[0])});
        }
        if (n == "count") {
            Value x = arg(0, "a");
            return num(x.kind == Value::List ? double(x.items.size()) : 1.0);
        }
        if (n == "matte" || n == "plastic" || n == "metal" || n == "glass") return material(name, a);
        fail(name, "unknown function " + n);
    }

    Value material(const Token& name, const Args& a) {
        kdn_material m{};
        const std::string& n = name.text;
        m.kind = n == "matte" ? KDN_MATTE : n == "plastic" ? KDN_PLASTIC : n == "metal" ? KDN_METAL : KDN_GLASS;
        m.roughness = n == "matte" ? 1.0f : 0.0f;
        m.refraction = 1.5f;
        m.strength = 1.0f;
        m.albedo[0] = m.albedo[1] = m.albedo[2] = 1.0f;
        auto named = [&](const char* k) -> const Value* {
            auto it = a.named.find(k);
            return it == a.named.end() ? nullptr : &it->second;
        };
        size_t next = 0;
        if (m.kind != KDN_GLASS) {
            const Value* al = next < a.positional.size() ? &a.positional[next++] : named("albedo");
            if (al) {
                std::vector<double> c = vec(*al, 3, name);
                for (int k = 0; k < 3; ++k) m.albedo[k] = float(c[k]);
            }
        }
        if (m.kind == KDN_PLASTIC || m.kind == KDN_METAL) {
            const Value* r = next < a.positional.size() ? &a.positional[next++] : named("roughness");
            if (r) m.roughness = float(scalar(*r));
        }
        if (m.kind == KDN_GLASS) {
            const Value* r = next < a.positional.size() ? &a.positional[next++] : named("refraction");
            if (r) m.refraction = float(scalar(*r));
            if (const Value* v = named("roughness")) m.roughness = float(scalar(*v));
            if (const Value* v = named("dispersion")) m.dispersion = float(scalar(*v));
        }
        if (const Value* g = named("glow")) {
            std::vector<double> c = vec(*g, 3, name);
            for (int k = 0; k < 3; ++k) m.glow[k] = float(c[k]);
        }
        if (const Value* s = named("strength")) m.strength = float(scalar(*s));
        scene_.materials.push_back(m);
        Value v;
        v.kind = Value::MaterialRef;
        v.material = uint32_t(scene_.materials.size() - 1);
        return v;
    }

    // ------------------------------------------------------- statements --
    // Run one statement. `parent` collects the nodes it emits inside a
    // join or with-block; at the top level (nullptr) each is an object.
    void statement(const Frame& frame, std::vector<uint32_t>* parent) {
        const Token& tk = peek();
        if (tk.kind == Tok::Sym && tk.text == ";") { next(); return; }
        if (tk.kind != Tok::Ident) fail(tk, "expected a statement");
        const std::string& w = tk.text;
        if (w == "import") {
            next();
            if (peek().kind != Tok::String) fail(peek(), "import takes a path");
            std::string path = next().text;
            if (path != "standard.sdf" && path.size() >= 12 && path.compare(path.size() - 12, 12, "standard.sdf") != 0) {
                fail(tk, "only the standard library is imported in this loader: " + path);
            }
            expect(";");
            return;
        }
        if (w == "let") {
            next();
            std::string n = ident();
            expect("=");
            Value v = expr();
            expect(";");
            if (n == "time") fail(tk, "time cannot be changed");
            set(n, v);
            return;
        }
        if (w == "for") { forLoop(frame, parent); return; }
        if (w == "if") {
            next();
            bool c = truth(expr());
            if (c) block(frame, parent, false);
            else skipBlock();
            return;
        }
        if (parent == nullptr && (w == "render" || w == "camera" || w == "grade") && isSym("{", 1)) { settings(); return; }
        if (parent == nullptr && w == "sky") { sky(); return; }
        if (parent == nullptr && w == "light") { light(); return; }
        emit(frame, parent);
    }

    void forLoop(const Frame& frame, std::vector<uint32_t>* parent) {
        next();
        struct Range { std::string name; long lo, hi; };
        std::vector<Range> ranges;
        for (;;) {
            Range r;
            r.name = ident();
            if (!isIdent("in")) fail(peek(), "expected in");
            next();
            r.lo = long(std::lround(scalar(addExpr())));
            expect("..");
            r.hi = long(std::lround(scalar(addExpr())));
            ranges.push_back(r);
            if (isSym(",")) { next(); continue; }
            break;
        }
        size_t body = p_;
        std::vector<long> at(ranges.size());
        std::function<void(size_t)> loop = [&](size_t k) {
            if (k == ranges.size()) {
                p_ = body;
                scopes_.emplace_back();
                for (size_t j = 0; j < ranges.size(); ++j) scopes_.back().names[ranges[j].name] = num(double(at[j]));
                block(frame, parent, true);
                scopes_.pop_back();
                return;
            }
            for (long v = ranges[k].lo; v < ranges[k].hi; ++v) {
                at[k] = v;
                loop(k + 1);
            }
        };
        bool any = true;
        for (const Range& r : ranges) any = any && r.hi > r.lo;
        if (any) loop(0);
        p_ = body;
        skipBlock();
    }

    void skipBlock() {
        expect("{");
        int depth = 1;
        while (depth > 0) {
            if (peek().kind == Tok::End) fail(peek(), "unclosed block");
            const Token& x = next();
            if (x.kind == Tok::Sym && x.text == "{") ++depth;
            if (x.kind == Tok::Sym && x.text == "}") --depth;
        }
    }

    void block(const Frame& frame, std::vector<uint32_t>* parent, bool scoped) {
        expect("{");
        if (!scoped) scopes_.emplace_back();
        while (!isSym("}")) {
            if (peek().kind == Tok::End) fail(peek(), "unclosed block");
            statement(frame, parent);
        }
        next();
        if (!scoped) scopes_.pop_back();
    }

    void fields(const std::function<void(const std::string&, const Value&, const Token&)>& take) {
        expect("{");
        while (!isSym("}")) {
            const Token& at = peek();
            std::string n = ident();
            expect("=");
            Value v = expr();
            expect(";");
            take(n, v, at);
        }
        next();
    }

    void settings() {
        const Token& kind = next();
        if (kind.text == "render") {
            fields([&](const std::string& n, const Value& v, const Token& at) {
                double x = scalar(v);
                if (n == "width") scene_.render.width = int(x);
                else if (n == "height") scene_.render.height = int(x);
                else if (n == "seed") scene_.render.seed = int(x);
                else if (n == "samples") scene_.render.samples = int(x);
                else if (n == "surfaces") scene_.render.surfaces = std::max(1, int(x));
                else if (n == "far") scene_.render.far_ = float(x);
                else if (n == "batch") scene_.render.batch = int(x);
                else if (n == "tolerance") scene_.render.tolerance = float(x);
                else if (n == "split") scene_.render.split = int(x);
                else if (n == "collisions") {}
                else fail(at, "render has no field " + n);
            });
        } else if (kind.text == "camera") {
            fields([&](const std::string& n, const Value& v, const Token& at) {
                if (n == "eye" || n == "target" || n == "up") {
                    std::vector<double> c = vec(v, 3, at);
                    float* dst = n == "eye" ? scene_.camera.eye : n == "target" ? scene_.camera.target : scene_.camera.up;
                    for (int k = 0; k < 3; ++k) dst[k] = float(c[k]);
                } else if (n == "fov") {
                    scene_.camera.fov = float(scalar(v));
                } else if (n == "aperture" || n == "focus") {
                    if (scalar(v) != 0) fail(at, "this loader takes a pinhole camera only");
                } else {
                    fail(at, "camera has no field " + n);
                }
            });
        } else if (kind.text == "grade") {
            // readGrade (syntax/evaluator.cpp:1971-1985)
            fields([&](const std::string& n, const Value& v, const Token& at) {
                if (n == "exposure") {
                    scene_.render.exposure = float(scalar(v));
                } else if (n == "curve") {
                    const std::string w = v.kind == Value::Word ? v.word : std::string();
                    if (w != "filmic" && w != "clamped") fail(at, "a curve is filmic or clamped");
                    scene_.render.curve = w == "filmic" ? KDN_CURVE_FILMIC : KDN_CURVE_CLAMPED;
                } else {
                    fail(at, "grade has no field " + n);
                }
            });
        } else {
            fields([](const std::string&, const Value&, const Token&) {});
        }
    }

    void sky() {
        next();
        const Token& kind = peek();
        if (!isIdent("gradient")) fail(kind, "this loader takes a gradient sky only");
        next();
        fields([&](const std::string& n, const Value& v, const Token& at) {
            std::vector<double> c = vec(v, 3, at);
            float* dst = n == "zenith" ? scene_.render.sky_zenith : n == "horizon" ? scene_.render.sky_horizon : n == "ground" ? scene_.render.sky_ground : nullptr;
            if (!dst) fail(at, "sky gradient has no field " + n);
            for (int k = 0; k < 3; ++k) dst[k] = float(c[k]);
        });
    }

    void light() {
        next();
        const Token& kind = next();
        // the reference's defaults (readLight, syntax/evaluator.cpp:2114-2140)
        kdn_light l{};
        l.direction[2] = 1;
        l.along[0] = 1;
        l.across[1] = 1;
        l.angle = 0.0093f;
        l.power[0] = l.power[1] = l.power[2] = 1;
        if (kind.text == "quad") l.kind = KDN_LIGHT_QUAD;
        else if (kind.text == "point") l.kind = KDN_LIGHT_POINT;
        else if (kind.text == "directional") l.kind = KDN_LIGHT_DIRECTIONAL;
        else if (kind.text == "sphere") l.kind = KDN_LIGHT_SPHERE;
        else if (kind.text == "sun") l.kind = KDN_LIGHT_SUN;
        else if (kind.text == "moon") l.kind = KDN_LIGHT_MOON;
        else fail(kind, "unknown light kind (point, directional, sphere, quad, sun, moon)");
        fields([&](const std::string& n, const Value& v, const Token& at) {
            if (n == "reaches" || n == "caustics") return;
            if (n == "radius" || n == "angle") {
                const std::vector<double> c = vec(v, 1, at);
                (n == "radius" ? l.radius : l.angle) = float(c[0]);
                return;
            }
            std::vector<double> c = vec(v, 3, at);
            float* dst = n == "at" ? l.at : n == "along" ? l.along : n == "across" ? l.across : n == "power" ? l.power
                       : n == "direction" ? l.direction : nullptr;
            if (!dst) fail(at, "light has no field " + n);
            for (int k = 0; k < 3; ++k) dst[k] = float(c[k]);
        });
        scene_.lights.push_back(l);
    }

    // Modifiers then a join, a with-block or a primitive call.
    void emit(const Frame& outer, std::vector<uint32_t>* parent) {
        Frame f = outer;
        bool with = false;
        if (isIdent("with")) { next(); with = true; }
        for (;;) {
            if (isIdent("at") && isSym("(", 1)) {
                const Token& at = next();
                Args a = args();
                std::vector<double> v = vec(a.positional.at(0), 3, at);
                const double I[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
                const double t[3] = {v[0], v[1], v[2]};
                f = composed(f, I, t);
            } else if (isIdent("turn") && isSym("(", 1)) {
                const Token& at = next();
                Args a = args();
                if (a.positional.size() != 2) fail(at, "turn takes an axis and an angle");
                std::vector<double> axis;
                const Value& ax = a.positional[0];
                if (ax.kind == Value::Word) axis = {ax.word == "x" ? 1.0 : 0.0, ax.word == "y" ? 1.0 : 0.0, ax.word == "z" ? 1.0 : 0.0};
                else axis = vec(ax, 3, at);
                double len = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
                for (double& c : axis) c /= len;
                double ang = scalar(a.positional[1]), c = std::cos(ang), s = std::sin(ang), k = 1 - c;
                double x = axis[0], y = axis[1], z = axis[2];
                const double R[3][3] = {{cleaned(c + x * x * k), cleaned(x * y * k - z * s), cleaned(x * z * k + y * s)},
                                        {cleaned(y * x * k + z * s), cleaned(c + y * y * k), cleaned(y * z * k - x * s)},
                                        {cleaned(z * x * k - y * s), cleaned(z * y * k + x * s), cleaned(c + z * z * k)}};
                const double t[3] = {0, 0, 0};
                f = composed(f, R, t);
            } else if (isIdent("scale") && isSym("(", 1)) {
                const Token& at = next();
                Args a = args();
                std::vector<double> s = vec(a.positional.at(0), 3, at);
                const double S[3][3] = {{s[0], 0, 0}, {0, s[1], 0}, {0, 0, s[2]}};
                const double t[3] = {0, 0, 0};
                f = composed(f, S, t);
                f.stretch *= std::min({std::abs(s[0]), std::abs(s[1]), std::abs(s[2])});
            } else if (isIdent("paint") && isSym("(", 1)) {
                const Token& at = next();
                Args a = args();
                if (a.positional.empty() || a.positional[0].kind != Value::MaterialRef) fail(at, "paint takes a material");
                f.material = a.positional[0].material;
            } else if ((isIdent("coat") || isIdent("rule") || isIdent("group")) && isSym("(", 1)) {
                fail(peek(), "this loader does not take " + peek().text);
            } else {
                break;
            }
        }
        if (with) {
            if (parent == nullptr) {
                block(f, nullptr, false);
            } else {
                std::vector<uint32_t> kids;
                block(f, &kids, false);
                attach(parent, joined(JoinMerge, 0, kids));
            }
            return;
        }
        const Token& head = peek();
        if (isIdent("repeat") && isSym("(", 1)) {
            repeat(f, parent);
            return;
        }
        if (isIdent("merge") || isIdent("cut") || isIdent("clip")) {
            next();
            Join j = head.text == "merge" ? JoinMerge : head.text == "cut" ? JoinCut : JoinClip;
            double radius = 0;
            if (isSym("(")) {
                next();
                radius = scalar(expr());
                expect(")");
            }
            std::vector<uint32_t> kids;
            block(f, &kids, false);
            if (kids.empty()) return;
            attach(parent, joined(j, radius, kids));
            return;
        }
        if (head.kind != Tok::Ident) fail(head, "expected a shape");
        next();
        Args a = args();
        expect(";");
        attach(parent, leafNode(head, a, f));
    }

    // repeat(count = [nx, ny, nz], step = [sx, sy, sz]) { ... } (runRepeat,
    // syntax/evaluator.cpp:2457-2577): the block's shapes, then copies of
    // them at every other place of the grid, place (i, j, k) offset by
    // i sx, j sy, k sz along the frame's axes (z outermost, x innermost),
    // all merged with radius 0. A copy is the same tree with its leaves'
    // origins moved.
    void repeat(const Frame& f, std::vector<uint32_t>* parent) {
        const Token& at = next();
        Args a = args();
        if (!a.positional.empty() || a.named.size() != 2 || !a.named.count("count") || !a.named.count("step")) {
            fail(at, "repeat takes count and step");
        }
        const std::vector<double> counts = vec(a.named.at("count"), 3, at), step = vec(a.named.at("step"), 3, at);
        int n[3];
        double steps[3][3];
        for (int axis = 0; axis < 3; ++axis) {
            if (!(counts[axis] >= 1) || counts[axis] != std::floor(counts[axis])) fail(at, "a repeat count is a whole number of one or more");
            n[axis] = (int)counts[axis];
            for (int row = 0; row < 3; ++row) steps[axis][row] = f.A[row][axis] * step[axis];
        }
        if ((double)n[0] * n[1] * n[2] > 1e6) fail(at, "a repeat of more than a million copies");
        std::vector<uint32_t> first;
        block(f, &first, false);
        if (first.empty()) fail(at, "this emits no shape");
        std::vector<uint32_t> copies = first;
        for (int k = 0; k < n[2]; ++k) {
            for (int j = 0; j < n[1]; ++j) {
                for (int i = 0; i < n[0]; ++i) {
                    if (i == 0 && j == 0 && k == 0) continue;
                    const int place[3] = {i, j, k};
                    double offset[3];
                    for (int row = 0; row < 3; ++row) offset[row] = steps[0][row] * place[0] + steps[1][row] * place[1] + steps[2][row] * place[2];
                    for (uint32_t node : first) copies.push_back(copied(node, offset));
                }
            }
        }
        if (parent == nullptr || copies.size() == 1) {
            for (uint32_t node : copies) attach(parent, node);
        } else {
            attach(parent, joined(JoinMerge, 0, copies));
        }
    }

    uint32_t copied(uint32_t index, const double offset[3]) {
        const Node node = scene_.nodes[index];
        if (node.join != JoinLeaf) {
            Node copy = node;
            copy.children.clear();
            for (uint32_t child : node.children) copy.children.push_back(copied(child, offset));
            scene_.nodes.push_back(copy);
            return uint32_t(scene_.nodes.size() - 1);
        }
        Leaf leaf = scene_.leaves[node.leaf];
        for (int i = 0; i < 3; ++i) leaf.origin[i] += offset[i];
        scene_.leaves.push_back(leaf);
        Node copy = node;
        copy.leaf = uint32_t(scene_.leaves.size() - 1);
        scene_.nodes.push_back(copy);
        return uint32_t(scene_.nodes.size() - 1);
    }

    uint32_t joined(Join j, double radius, const std::vector<uint32_t>& kids) {
        if (kids.size() == 1) return kids[0];
        Node n;
        n.join = j;
        n.radius = radius;
        n.leaf = 0;
        n.children = kids;
        scene_.nodes.push_back(n);
        return uint32_t(scene_.nodes.size() - 1);
    }

    void attach(std::vector<uint32_t>* parent, uint32_t node) {
        if (parent) parent->push_back(node);
        else scene_.objects.push_back(node);
    }

    uint32_t leafNode(const Token& head, const Args& a, const Frame& f) {
        Leaf leaf{};
        const std::string& n = head.text;
        auto pos = [&](size_t i) -> const Value& {
            if (i >= a.positional.size()) fail(head, n + " needs more arguments");
            return a.positional[i];
        };
        if (n == "sphere") {
            leaf.primitive = KDN_PRIM_SPHERE;
            leaf.args[0] = scalar(pos(0));
        } else if (n == "box") {
            leaf.primitive = KDN_PRIM_BOX;
            std::vector<double> s = vec(pos(0), 3, head);
            for (int k = 0; k < 3; ++k) leaf.args[k] = s[k];
            leaf.args[3] = a.positional.size() > 1 ? scalar(pos(1)) : 0.0;
        } else if (n == "torus") {
            leaf.primitive = KDN_PRIM_TORUS;
            leaf.args[0] = scalar(pos(0));
            leaf.args[1] = scalar(pos(1));
        } else if (n == "plane") {
            leaf.primitive = KDN_PRIM_PLANE;
            std::vector<double> nn = vec(pos(0), 3, head);
            for (int k = 0; k < 3; ++k) leaf.args[k] = nn[k];
            leaf.args[3] = a.positional.size() > 1 ? scalar(pos(1)) : 0.0;
        } else if (n == "cylinder") {
            leaf.primitive = KDN_PRIM_CYLINDER;
            leaf.args[0] = scalar(pos(0));
            leaf.args[1] = scalar(pos(1));
        } else if (n == "capsule") {
            leaf.primitive = KDN_PRIM_CAPSULE;
            std::vector<double> p0 = vec(pos(0), 3, head), p1 = vec(pos(1), 3, head);
            for (int k = 0; k < 3; ++k) { leaf.args[k] = p0[k]; leaf.args[3 + k] = p1[k]; }
            leaf.args[6] = scalar(pos(2));
        } else if (n == "ellipsoid") {
            leaf.primitive = KDN_PRIM_ELLIPSOID;
            std::vector<double> r = vec(pos(0), 3, head);
            for (int k = 0; k < 3; ++k) leaf.args[k] = r[k];
        } else {
            fail(head, "unknown shape " + n + " (this loader knows the standard primitives)");
        }
        double inv[3][3];
        invert3(f.A, inv);
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) leaf.linear[i][j] = cleaned(inv[i][j]);
            leaf.origin[i] = f.b[i];
        }
        leaf.stretch = f.stretch;
        leaf.material = f.material ? *f.material : KDN_NO_MATERIAL;
        scene_.leaves.push_back(leaf);
        Node node;
        node.join = JoinLeaf;
        node.radius = 0;
        node.leaf = uint32_t(scene_.leaves.size() - 1);
        scene_.nodes.push_back(node);
        return uint32_t(scene_.nodes.size() - 1);
    }
};

// ------------------------------------------------------------- compiler --
struct Builder {
    Program program;
    uint32_t leaf = 0;
    uint32_t point[3] = {0, 1, 2};
};

uint32_t emitStep(Builder& b, uint8_t op, uint32_t a = 0, uint32_t bb = 0, uint32_t c = 0, double value = 0) {
    b.program.steps.push_back(Step{op, a, bb, c, value, b.leaf});
    return uint32_t(b.program.steps.size() - 1);
}
uint32_t emitValue(Builder& b, double v) { return emitStep(b, KDN_CONSTANT, 0, 0, 0, v); }
uint32_t emit2(Builder& b, uint8_t op, uint32_t x, uint32_t y) { return emitStep(b, op, x, y); }
uint32_t emit1(Builder& b, uint8_t op, uint32_t x) { return emitStep(b, op, x); }

// The local coordinate from the placement: offsets p - origin, then rows
// of the linear part, skipping zero weights and not multiplying by one
// (source/sdf/scene.cpp:95-148).
void startInlining(Builder& b, const Leaf& leaf, uint32_t local[3]) {
    uint32_t offsets[3];
    for (int axis = 0; axis < 3; ++axis) {
        offsets[axis] = leaf.origin[axis] == 0 ? b.point[axis]
                                               : emit2(b, KDN_SUBTRACT, b.point[axis], emitValue(b, leaf.origin[axis]));
    }
    for (int i = 0; i < 3; ++i) {
        std::optional<uint32_t> sum;
        for (int j = 0; j < 3; ++j) {
            const double w = leaf.linear[i][j];
            if (w == 0) continue;
            uint32_t term = w == 1 ? offsets[j] : emit2(b, KDN_MULTIPLY, emitValue(b, w), offsets[j]);
            sum = sum ? emit2(b, KDN_ADD, *sum, term) : term;
        }
        local[i] = sum ? *sum : emitValue(b, 0);
    }
}

// dot(v, v) folded left with Add, as the reference dotValue/foldList
// (source/syntax/evaluator.cpp:258-301).
uint32_t sumOfSquares(Builder& b, const std::vector<uint32_t>& v) {
    uint32_t total = emit2(b, KDN_MULTIPLY, v[0], v[0]);
    for (size_t i = 1; i < v.size(); ++i) total = emit2(b, KDN_ADD, total, emit2(b, KDN_MULTIPLY, v[i], v[i]));
    return total;
}

uint32_t lengthOf(Builder& b, const std::vector<uint32_t>& v) {
    if (v.size() == 1) return emit1(b, KDN_ABSOLUTE, v[0]);
    return emit1(b, KDN_SQUARE_ROOT, sumOfSquares(b, v));
}

// length(max(v, 0)) with the positive squares the cut would produce from
// mul(max(x,0), max(x,0)) (plugins/geometry/tree/cut.cpp:210-224); b of each
// PositiveSquare names the zero constant so sqrt(pos²) unsquares to max.
uint32_t positiveLength(Builder& b, const std::vector<uint32_t>& v, uint32_t zero) {
    uint32_t total = emit2(b, KDN_POSITIVE_SQUARE, v[0], zero);
    for (size_t i = 1; i < v.size(); ++i) total = emit2(b, KDN_ADD, total, emit2(b, KDN_POSITIVE_SQUARE, v[i], zero));
    return emit1(b, KDN_SQUARE_ROOT, total);
}

// The standard.sdf primitives (scenes/standard.sdf:1-43) over local
// coordinates, their number arguments known.
uint32_t primitiveSteps(Builder& b, const Leaf& leaf, const uint32_t p[3]) {
    const double* a = leaf.args;
    switch (leaf.primitive) {
        case KDN_PRIM_SPHERE:  // length(p) - r
            return emit2(b, KDN_SUBTRACT, lengthOf(b, {p[0], p[1], p[2]}), emitValue(b, a[0]));
        case KDN_PRIM_BOX: {  // q = abs(p) - size/2 + r; length(max(q,0)) + min(max(q.x,max(q.y,q.z)),0) - r
            uint32_t q[3];
            for (int k = 0; k < 3; ++k) q[k] = emit2(b, KDN_SUBTRACT, emit1(b, KDN_ABSOLUTE, p[k]), emitValue(b, a[k] / 2 - a[3]));
            const uint32_t zero = emitValue(b, 0);
            const uint32_t outside = positiveLength(b, {q[0], q[1], q[2]}, zero);
            const uint32_t inner = emit2(b, KDN_MINIMUM, emit2(b, KDN_MAXIMUM, q[0], emit2(b, KDN_MAXIMUM, q[1], q[2])), zero);
            const uint32_t sum = emit2(b, KDN_ADD, outside, inner);
            return a[3] == 0 ? sum : emit2(b, KDN_SUBTRACT, sum, emitValue(b, a[3]));
        }
        case KDN_PRIM_TORUS: {  // length([length(p.xy) - big, p.z]) - small
            const uint32_t ring = emit2(b, KDN_SUBTRACT, lengthOf(b, {p[0], p[1]}), emitValue(b, a[0]));
            return emit2(b, KDN_SUBTRACT, lengthOf(b, {ring, p[2]}), emitValue(b, a[1]));
        }
        case KDN_PRIM_PLANE: {  // dot(p, normal) - offset
            std::optional<uint32_t> sum;
            for (int k = 0; k < 3; ++k) {
                if (a[k] == 0) continue;
                uint32_t term = a[k] == 1 ? p[k] : a[k] == -1 ? emit1(b, KDN_NEGATE, p[k]) : emit2(b, KDN_MULTIPLY, p[k], emitValue(b, a[k]));
                sum = sum ? emit2(b, KDN_ADD, *sum, term) : term;
            }
            uint32_t dot = sum ? *sum : emitValue(b, 0);
            return a[3] == 0 ? dot : emit2(b, KDN_SUBTRACT, dot, emitValue(b, a[3]));
        }
        case KDN_PRIM_CYLINDER: {  // d = abs([length(p.xy), p.z]) - [r, h/2]; min(max(d.x,d.y),0) + length(max(d,0))
            const uint32_t dx = emit2(b, KDN_SUBTRACT, emit1(b, KDN_ABSOLUTE, lengthOf(b, {p[0], p[1]})), emitValue(b, a[1]));
            const uint32_t dz = emit2(b, KDN_SUBTRACT, emit1(b, KDN_ABSOLUTE, p[2]), emitValue(b, a[0] / 2));
            const uint32_t zero = emitValue(b, 0);
            const uint32_t inner = emit2(b, KDN_MINIMUM, emit2(b, KDN_MAXIMUM, dx, dz), zero);
            return emit2(b, KDN_ADD, inner, positiveLength(b, {dx, dz}, zero));
        }
        case KDN_PRIM_CAPSULE: {  // pa = p - a; ba = b - a; t = clamp(dot(pa,ba)/dot(ba,ba),0,1); length(pa - ba t) - r
            const double ba[3] = {a[3] - a[0], a[4] - a[1], a[5] - a[2]};
            const double bb = ba[0] * ba[0] + ba[1] * ba[1] + ba[2] * ba[2];
            uint32_t pa[3];
            for (int k = 0; k < 3; ++k) pa[k] = a[k] == 0 ? p[k] : emit2(b, KDN_SUBTRACT, p[k], emitValue(b, a[k]));
            std::optional<uint32_t> dot;
            for (int k = 0; k < 3; ++k) {
                if (ba[k] == 0) continue;
                uint32_t term = emit2(b, KDN_MULTIPLY, pa[k], emitValue(b, ba[k] / bb));
                dot = dot ? emit2(b, KDN_ADD, *dot, term) : term;
            }
            const uint32_t t = emit2(b, KDN_MINIMUM, emit2(b, KDN_MAXIMUM, dot ? *dot : emitValue(b, 0), emitValue(b, 0)), emitValue(b, 1));
            uint32_t r[3];
            for (int k = 0; k < 3; ++k) r[k] = ba[k] == 0 ? pa[k] : emit2(b, KDN_SUBTRACT, pa[k], emit2(b, KDN_MULTIPLY, t, emitValue(b, ba[k])));
            return emit2(b, KDN_SUBTRACT, lengthOf(b, {r[0], r[1], r[2]}), emitValue(b, a[6]));
        }
        case KDN_PRIM_ELLIPSOID: {  // (length(p / radii) - 1) * min(radii)
            uint32_t s[3];
            for (int k = 0; k < 3; ++k) s[k] = emit2(b, KDN_DIVIDE, p[k], emitValue(b, a[k]));
            const uint32_t l = emit2(b, KDN_SUBTRACT, lengthOf(b, {s[0], s[1], s[2]}), emitValue(b, 1));
            return emit2(b, KDN_MULTIPLY, l, emitValue(b, std::min({a[0], a[1], a[2]})));
        }
    }
    return emitValue(b, 1e30);
}

uint32_t inlineLeaf(Builder& b, const Scene& scene, uint32_t index) {
    const Leaf& leaf = scene.leaves[index];
    b.leaf = index;
    uint32_t local[3];
    startInlining(b, leaf, local);
    const uint32_t result = primitiveSteps(b, leaf, local);
    if (leaf.stretch == 1) return result;
    return emit2(b, KDN_MULTIPLY, result, emitValue(b, leaf.stretch));
}

uint32_t hardJoin(Builder& b, Join join, uint32_t x, uint32_t y) {
    switch (join) {
        case JoinMerge: return emit2(b, KDN_MINIMUM, x, y);
        case JoinCut: return emit2(b, KDN_MAXIMUM, x, emit1(b, KDN_NEGATE, y));
        case JoinClip: return emit2(b, KDN_MAXIMUM, x, y);
        default: return x;
    }
}

uint32_t smoothMin(Builder& b, uint32_t x, uint32_t y, double radius) {
    const uint32_t r = emitValue(b, radius);
    return emitStep(b, KDN_SMOOTH_MINIMUM, x, y, r);
}

// Cut and clip as negated smooth minimums (source/sdf/scene.cpp:200-219).
uint32_t smoothJoin(Builder& b, const Node& node, uint32_t x, uint32_t y) {
    if (node.join == JoinMerge) return smoothMin(b, x, y, node.radius);
    const uint32_t nx = emit1(b, KDN_NEGATE, x);
    const uint32_t ny = node.join == JoinClip ? emit1(b, KDN_NEGATE, y) : y;
    return emit1(b, KDN_NEGATE, smoothMin(b, nx, ny, node.radius));
}

uint32_t joinNode(Builder& b, const Node& node, uint32_t x, uint32_t y) {
    b.leaf = b.program.steps[x].leaf;
    if (node.radius <= 0) return hardJoin(b, node.join, x, y);
    return smoothJoin(b, node, x, y);
}

uint32_t compileNode(Builder& b, const Scene& scene, uint32_t index) {
    const Node& node = scene.nodes[index];
    if (node.join == JoinLeaf) return inlineLeaf(b, scene, node.leaf);
    uint32_t result = compileNode(b, scene, node.children[0]);
    for (size_t i = 1; i < node.children.size(); ++i) {
        const uint32_t other = compileNode(b, scene, node.children[i]);
        result = joinNode(b, node, result, other);
    }
    return result;
}

uint32_t firstLeaf(const Scene& scene, uint32_t index) {
    const Node& node = scene.nodes[index];
    return node.join == JoinLeaf ? node.leaf : firstLeaf(scene, node.children[0]);
}

Program nodeProgram(const Scene& scene, uint32_t node) {
    Builder b;
    b.leaf = firstLeaf(scene, node);
    for (int i = 0; i < 3; ++i) emitStep(b, uint8_t(KDN_POINT_X + i));
    b.program.result = compileNode(b, scene, node);
    return pruneProgram(b.program);
}

// One leaf alone over world space (leafProgram, source/sdf/scene.cpp:
// 284-288): the point's steps, the leaf inlined, pruned.
Program leafAlone(const Scene& scene, uint32_t leaf) {
    Builder b;
    b.leaf = leaf;
    for (int i = 0; i < 3; ++i) emitStep(b, uint8_t(KDN_POINT_X + i));
    b.program.result = inlineLeaf(b, scene, leaf);
    return pruneProgram(b.program);
}

void unionParts(const Scene& scene, uint32_t node, std::vector<uint32_t>& parts) {
    const Node& n = scene.nodes[node];
    const bool split = n.join == JoinMerge && n.radius == 0 && nodeProgram(scene, node).steps.size() >= kFewestUnionSteps;
    if (!split) {
        parts.push_back(node);
        return;
    }
    for (uint32_t child : n.children) unionParts(scene, child, parts);
}

}  // namespace

Program leafProgram(const Scene& scene, uint32_t leaf) { return leafAlone(scene, leaf); }

int operandCount(uint8_t op) {
    switch (op) {
        case KDN_CONSTANT: case KDN_ARGUMENT: case KDN_POINT_X: case KDN_POINT_Y: case KDN_POINT_Z:
        case KDN_NORMAL_X: case KDN_NORMAL_Y: case KDN_NORMAL_Z: return 0;
        case KDN_NEGATE: case KDN_SQUARE_ROOT: case KDN_ABSOLUTE: case KDN_FLOOR: case KDN_SINE:
        case KDN_COSINE: case KDN_SIGN: case KDN_NOT: return 1;
        case KDN_SELECT: case KDN_SMOOTH_MINIMUM: return 3;
        default: return 2;
    }
}

Program pruneProgram(const Program& program) {
    const size_t n = program.steps.size();
    std::vector<uint32_t> live(n, 0);
    live[program.result] = 1;
    for (size_t i = n; i-- > 0;) {
        if (!live[i]) continue;
        const Step& s = program.steps[i];
        const int count = operandCount(s.op);
        const uint32_t ops[3] = {s.a, s.b, s.c};
        for (int k = 0; k < count; ++k) live[ops[k]] = 1;
    }
    Program out;
    std::vector<uint32_t> map(n, 0);
    for (size_t i = 0; i < n; ++i) {
        if (!live[i]) continue;
        Step s = program.steps[i];
        const int count = operandCount(s.op);
        s.a = count > 0 ? map[s.a] : s.a;
        s.b = count > 1 ? map[s.b] : 0;
        s.c = count > 2 ? map[s.c] : 0;
        map[i] = uint32_t(out.steps.size());
        out.steps.push_back(s);
    }
    out.result = map[program.result];
    return out;
}

bool loadScene(const std::string& text, const std::string& baseDir, Scene& scene, std::string& error) {
    (void)baseDir;
    scene = Scene{};
    scene.camera = kdn_camera{{0, -5, 2}, {0, 0, 0}, {0, 0, 1}, 0.8f};
    scene.render = kdn_render_settings{};
    scene.render.width = 1920;
    scene.render.height = 1080;
    scene.render.seed = 1;
    scene.render.samples = 16;
    scene.render.surfaces = 10;
    scene.render.far_ = 1000.0f;
    scene.render.exposure = 1.0f;  // defaultGrade (syntax/evaluator.cpp:2735-2740)
    scene.render.curve = KDN_CURVE_FILMIC;
    try {
        std::vector<Token> tokens = lex(text);
        Loader loader(tokens, scene);
        loader.run();
    } catch (const LoadError& e) {
        error = std::to_string(e.line) + ":" + std::to_string(e.column) + ": " + e.what();
        return false;
    } catch (const std::exception& e) {
        error = std::string("0:0: ") + e.what();
        return false;
    }
    if (scene.objects.empty()) {
        error = "0:0: the scene has no shapes";
        return false;
    }
    compileScene(scene);
    return true;
}

void compileScene(Scene& scene) {
    std::vector<uint32_t> parts;
    for (uint32_t object : scene.objects) unionParts(scene, object, parts);
    scene.programs.clear();
    scene.programMaterial.clear();
    for (uint32_t part : parts) {
        Program prog = nodeProgram(scene, part);
        if (prog.steps.size() > 65535) throw std::runtime_error("an object program is longer than 65535 steps");
        scene.programs.push_back(std::move(prog));
        scene.programMaterial.push_back(scene.leaves[firstLeaf(scene, part)].material);
    }
}

}  // namespace kdn
