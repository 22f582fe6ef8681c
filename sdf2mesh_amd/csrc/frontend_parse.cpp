// frontend_parse.cpp - preprocessing, lexing and parsing of `.sdf3d` WGSL and of GLSL
// fragment shaders into the front end's AST.
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "frontend.h"

namespace s2m::fe {

// ---- preprocessing (shader.rs:159-203) -------------------------------------------------
// A line whose trimmed text ends with ';' and starts with `use` names a module, one that
// starts with `include` names a file spliced in place (recursively, skipping a file that
// includes itself); every other line is kept.  Unreadable files are skipped, as the
// reference logs and continues.
static std::string trim(const std::string& s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

static std::string strip_directive(const std::string& t, const char* kw)
{
    std::string r = t.substr(std::strlen(kw));
    std::string o;
    for (char c : r)
        if (c != '"' && c != ';') o.push_back(c);
    return trim(o);
}

static void expand(const std::string& src, const std::string& path, std::string& out,
                   std::vector<std::string>& uses, int depth)
{
    std::istringstream in(src);
    std::string line;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        const std::string t = trim(line);
        if (!t.empty() && t.back() == ';') {
            if (t.rfind("use", 0) == 0) {
                uses.push_back(strip_directive(t, "use"));
                continue;
            }
            if (t.rfind("include", 0) == 0) {
                const std::string f = strip_directive(t, "include");
                if (f != path && depth < 32) {
                    std::ifstream is(f);
                    if (is) {
                        std::stringstream ss;
                        ss << is.rdbuf();
                        expand(ss.str(), f, out, uses, depth + 1);
                    }
                }
                continue;
            }
        }
        out += line;
        out += '\n';
    }
}

bool preprocess_file(const std::string& path, std::string& out, std::vector<std::string>& uses,
                     Error& err)
{
    std::ifstream is(path);
    if (!is) {
        err = {S2M_E_IO, "cannot read " + path};
        return false;
    }
    std::stringstream ss;
    ss << is.rdbuf();
    out.clear();
    expand(ss.str(), path, out, uses, 0);
    return true;
}

void preprocess_text(const std::string& src, std::string& out, std::vector<std::string>& uses)
{
    out.clear();
    expand(src, "", out, uses, 0);
}

// ---- lexer -----------------------------------------------------------------------------
namespace {

struct Tok {
    enum K { Id, Num, P, End } k;
    std::string s;
    int line;
};

bool lex(const std::string& src, Lang lang, std::vector<Tok>& out, Error& err)
{
    static const char* p3[] = {"<<=", ">>=", nullptr};
    static const char* p2[] = {"->", "+=", "-=", "*=", "/=", "%=", "++", "--", "<=", ">=", "==",
                               "!=", "&&", "||", "<<", ">>", "&=", "|=", "^=", nullptr};
    int line = 1;
    size_t i = 0, n = src.size();
    while (i < n) {
        char c = src[i];
        if (c == '\n') {
            ++line;
            ++i;
            continue;
        }
        if (std::isspace((unsigned char)c)) {
            ++i;
            continue;
        }
        if (c == '/' && i + 1 < n && src[i + 1] == '/') {
            while (i < n && src[i] != '\n') ++i;
            continue;
        }
        if (c == '/' && i + 1 < n && src[i + 1] == '*') {
            i += 2;
            while (i + 1 < n && !(src[i] == '*' && src[i + 1] == '/')) {
                if (src[i] == '\n') ++line;
                ++i;
            }
            i += 2;
            continue;
        }
        if (lang == Lang::Glsl && c == '#') {  // preprocessor directive line (#version ...)
            while (i < n && src[i] != '\n') ++i;
            continue;
        }
        if (std::isalpha((unsigned char)c) || c == '_') {
            size_t j = i;
            while (j < n && (std::isalnum((unsigned char)src[j]) || src[j] == '_')) ++j;
            out.push_back({Tok::Id, src.substr(i, j - i), line});
            i = j;
            continue;
        }
        if (std::isdigit((unsigned char)c) || (c == '.' && i + 1 < n && std::isdigit((unsigned char)src[i + 1]))) {
            size_t j = i;
            if (c == '0' && j + 1 < n && (src[j + 1] == 'x' || src[j + 1] == 'X')) {
                j += 2;
                while (j < n && std::isxdigit((unsigned char)src[j])) ++j;
            } else {
                while (j < n && (std::isdigit((unsigned char)src[j]) || src[j] == '.')) ++j;
                if (j < n && (src[j] == 'e' || src[j] == 'E')) {
                    ++j;
                    if (j < n && (src[j] == '+' || src[j] == '-')) ++j;
                    while (j < n && std::isdigit((unsigned char)src[j])) ++j;
                }
            }
            if (j < n && (src[j] == 'f' || src[j] == 'F' || src[j] == 'u' || src[j] == 'U' ||
                          src[j] == 'i' || src[j] == 'h'))
                ++j;
            out.push_back({Tok::Num, src.substr(i, j - i), line});
            i = j;
            continue;
        }
        bool done = false;
        for (const char** p = p3; *p && !done; ++p)
            if (src.compare(i, 3, *p) == 0) {
                out.push_back({Tok::P, *p, line});
                i += 3;
                done = true;
            }
        for (const char** p = p2; *p && !done; ++p)
            if (src.compare(i, 2, *p) == 0) {
                out.push_back({Tok::P, *p, line});
                i += 2;
                done = true;
            }
        if (done) continue;
        if (std::strchr("(){}[],;:.=+-*/%<>!&|^~?@", c)) {
            out.push_back({Tok::P, std::string(1, c), line});
            ++i;
            continue;
        }
        err = {S2M_E_PARSE, "line " + std::to_string(line) + ": unexpected character '" +
                                std::string(1, c) + "'"};
        return false;
    }
    out.push_back({Tok::End, "", line});
    return true;
}

// ---- parser ----------------------------------------------------------------------------
struct Parser {
    std::vector<Tok> t;
    size_t i = 0;
    Lang lang;
    Error err{0, ""};

    const Tok& peek(size_t k = 0) const { return t[std::min(i + k, t.size() - 1)]; }
    bool is(const char* s, size_t k = 0) const
    {
        const Tok& x = peek(k);
        return (x.k == Tok::P || x.k == Tok::Id) && x.s == s;
    }
    bool accept(const char* s)
    {
        if (is(s)) {
            ++i;
            return true;
        }
        return false;
    }
    bool fail(const std::string& m)
    {
        if (!err.code) err = {S2M_E_PARSE, "line " + std::to_string(peek().line) + ": " + m};
        return false;
    }
    bool expect(const char* s)
    {
        if (accept(s)) return true;
        return fail(std::string("expected '") + s + "', found '" + peek().s + "'");
    }
    bool ident(std::string& out)
    {
        if (peek().k != Tok::Id) return fail("expected an identifier, found '" + peek().s + "'");
        out = t[i++].s;
        return true;
    }

    // ---- types ----
    static bool scalar_of(const std::string& s, Base& b)
    {
        if (s == "f32" || s == "float") b = Base::F32;
        else if (s == "i32" || s == "int") b = Base::I32;
        else if (s == "u32" || s == "uint") b = Base::U32;
        else if (s == "bool") b = Base::Bool;
        else return false;
        return true;
    }
    // WGSL: f32 vec3f vec3<f32>; GLSL: float vec3 ivec3 uvec3 bvec3
    bool type_name(const std::string& s, Type& ty) const
    {
        Base b;
        if (scalar_of(s, b)) {
            ty = {b, 1};
            return true;
        }
        auto vec = [&](const std::string& pre, Base base) {
            if (s.size() == pre.size() + 1 && s.compare(0, pre.size(), pre) == 0 && s.back() >= '2' &&
                s.back() <= '4') {
                ty = {base, (uint8_t)(s.back() - '0')};
                return true;
            }
            return false;
        };
        if (lang == Lang::Wgsl) {
            if (s.size() == 5 && s.compare(0, 3, "vec") == 0 && s[3] >= '2' && s[3] <= '4') {
                const char c = s[4];
                Base bb = c == 'f' ? Base::F32 : c == 'i' ? Base::I32 : c == 'u' ? Base::U32 : Base::Void;
                if (bb == Base::Void) return false;
                ty = {bb, (uint8_t)(s[3] - '0')};
                return true;
            }
            return false;
        }
        return vec("vec", Base::F32) || vec("ivec", Base::I32) || vec("uvec", Base::U32) ||
               vec("bvec", Base::Bool);
    }
    // parse a type at the cursor (WGSL also accepts vecN<T>)
    bool parse_type(Type& ty)
    {
        std::string s;
        if (!ident(s)) return false;
        if (lang == Lang::Wgsl && s.size() == 4 && s.compare(0, 3, "vec") == 0 && s[3] >= '2' &&
            s[3] <= '4' && is("<")) {
            ++i;
            std::string e;
            if (!ident(e)) return false;
            Base b;
            if (!scalar_of(e, b)) return fail("unknown vector element type " + e);
            if (!expect(">")) return false;
            ty = {b, (uint8_t)(s[3] - '0')};
            return true;
        }
        if (!type_name(s, ty)) return fail("unknown type " + s);
        return true;
    }
    bool at_type() const
    {
        if (peek().k != Tok::Id) return false;
        Type ty;
        if (type_name(peek().s, ty)) return true;
        return lang == Lang::Glsl && peek().s == "const";
    }

    // ---- expressions ----
    ExprP make(Expr::Kind k)
    {
        auto e = std::make_shared<Expr>();
        e->kind = k;
        e->line = peek().line;
        return e;
    }
    ExprP bin(const std::string& op, ExprP a, ExprP b)
    {
        auto e = make(Expr::Binary);
        e->op = op;
        e->args = {a, b};
        return e;
    }
    ExprP expr() { return ternary(); }
    ExprP ternary()
    {
        ExprP c = binary(0);
        if (!c) return nullptr;
        if (lang == Lang::Glsl && accept("?")) {
            ExprP a = expr();
            if (!a || !expect(":")) return nullptr;
            ExprP b = ternary();
            if (!b) return nullptr;
            auto e = make(Expr::Call);
            e->name = "select";  // select(false_value, true_value, cond)
            e->args = {b, a, c};
            return e;
        }
        return c;
    }
    int prec(const std::string& op) const
    {
        static const std::map<std::string, int> P = {
            {"||", 1}, {"&&", 2}, {"|", 3}, {"^", 4}, {"&", 5}, {"==", 6}, {"!=", 6},
            {"<", 7},  {">", 7},  {"<=", 7}, {">=", 7}, {"<<", 8}, {">>", 8}, {"+", 9},
            {"-", 9},  {"*", 10}, {"/", 10}, {"%", 10}};
        auto it = P.find(op);
        return it == P.end() ? -1 : it->second;
    }
    ExprP binary(int min_prec)
    {
        ExprP a = unary();
        if (!a) return nullptr;
        for (;;) {
            const Tok& k = peek();
            if (k.k != Tok::P) break;
            const int p = prec(k.s);
            if (p < 0 || p <= min_prec - 1 || p < min_prec) break;
            const std::string op = k.s;
            ++i;
            ExprP b = binary(p + 1);
            if (!b) return nullptr;
            a = bin(op, a, b);
        }
        return a;
    }
    ExprP unary()
    {
        if (is("-") || is("!") || is("~") || is("+")) {
            const std::string op = peek().s;
            ++i;
            ExprP a = unary();
            if (!a) return nullptr;
            if (op == "+") return a;
            auto e = make(Expr::Unary);
            e->op = op;
            e->args = {a};
            return e;
        }
        if (lang == Lang::Glsl && (is("++") || is("--"))) {
            fail("pre-increment inside an expression is not supported");
            return nullptr;
        }
        return postfix();
    }
    ExprP postfix()
    {
        ExprP a = primary();
        if (!a) return nullptr;
        for (;;) {
            if (accept(".")) {
                auto e = make(Expr::Member);
                if (!ident(e->name)) return nullptr;
                e->args = {a};
                a = e;
            } else if (accept("[")) {
                auto e = make(Expr::Index);
                ExprP idx = expr();
                if (!idx || !expect("]")) return nullptr;
                e->args = {a, idx};
                a = e;
            } else {
                break;
            }
        }
        return a;
    }
    ExprP primary()
    {
        const Tok& k = peek();
        if (k.k == Tok::Num) {
            auto e = make(Expr::Lit);
            std::string s = k.s;
            ++i;
            char suf = 0;
            if (!s.empty() && std::strchr("fFuUih", s.back()) && !(s.size() > 1 && s[1] == 'x')) {
                suf = (char)std::tolower(s.back());
                s.pop_back();
            }
            const bool hex = s.size() > 2 && s[0] == '0' && (s[1] == 'x' || s[1] == 'X');
            e->is_float_lit = !hex && (s.find('.') != std::string::npos ||
                                       s.find('e') != std::string::npos ||
                                       s.find('E') != std::string::npos || suf == 'f' || suf == 'h');
            e->suffix = suf == 'h' ? 'f' : suf;
            e->num = hex ? (double)std::strtoull(s.c_str(), nullptr, 16) : std::strtod(s.c_str(), nullptr);
            return e;
        }
        if (k.k == Tok::Id) {
            if (k.s == "true" || k.s == "false") {
                auto e = make(Expr::Lit);
                e->is_bool = true;
                e->num = k.s == "true" ? 1 : 0;
                ++i;
                return e;
            }
            std::string name = k.s;
            ++i;
            // WGSL templated constructor vec3<f32>(...)
            if (lang == Lang::Wgsl && name.size() == 4 && name.compare(0, 3, "vec") == 0 && is("<")) {
                ++i;
                std::string el;
                if (!ident(el) || !expect(">")) return nullptr;
                Base b;
                if (!scalar_of(el, b)) {
                    fail("unknown vector element type " + el);
                    return nullptr;
                }
                name += b == Base::F32 ? "f" : b == Base::I32 ? "i" : "u";
            }
            if (accept("(")) {
                auto e = make(Expr::Call);
                e->name = name;
                if (!accept(")")) {
                    for (;;) {
                        ExprP a = expr();
                        if (!a) return nullptr;
                        e->args.push_back(a);
                        if (accept(")")) break;
                        if (!expect(",")) return nullptr;
                    }
                }
                return e;
            }
            auto e = make(Expr::Ident);
            e->name = name;
            return e;
        }
        if (accept("(")) {
            ExprP a = expr();
            if (!a || !expect(")")) return nullptr;
            return a;
        }
        fail("unexpected '" + k.s + "' in an expression");
        return nullptr;
    }

    // ---- statements ----
    StmtP make_s(Stmt::Kind k)
    {
        auto s = std::make_shared<Stmt>();
        s->kind = k;
        s->line = peek().line;
        return s;
    }
    bool block(std::vector<StmtP>& out)
    {
        if (!expect("{")) return false;
        while (!accept("}")) {
            if (peek().k == Tok::End) return fail("unterminated block");
            if (!statement(out)) return false;
        }
        return true;
    }
    // GLSL allows a single statement where WGSL requires a block
    bool body(std::vector<StmtP>& out)
    {
        if (is("{")) return block(out);
        if (lang == Lang::Glsl) return statement(out);
        return fail("expected '{'");
    }
    // assignment / increment / call without the trailing ';'
    StmtP simple()
    {
        if (lang == Lang::Glsl && (is("++") || is("--"))) {  // ++i
            const std::string op = peek().s;
            ++i;
            ExprP lhs = postfix();
            if (!lhs) return nullptr;
            auto s = make_s(Stmt::Assign);
            s->lhs = lhs;
            s->op = op == "++" ? "+=" : "-=";
            auto one = make(Expr::Lit);
            one->num = 1;
            s->init = one;
            return s;
        }
        ExprP lhs = postfix();
        if (!lhs) return nullptr;
        static const char* aops[] = {"=", "+=", "-=", "*=", "/=", "%=", nullptr};
        for (const char** p = aops; *p; ++p)
            if (accept(*p)) {
                auto s = make_s(Stmt::Assign);
                s->lhs = lhs;
                s->op = *p;
                s->init = expr();
                return s->init ? s : nullptr;
            }
        if (is("++") || is("--")) {
            const std::string op = peek().s;
            ++i;
            auto s = make_s(Stmt::Assign);
            s->lhs = lhs;
            s->op = op == "++" ? "+=" : "-=";
            auto one = make(Expr::Lit);
            one->num = 1;
            s->init = one;
            return s;
        }
        auto s = make_s(Stmt::ExprS);
        s->init = lhs;
        return s;
    }
    // declarations: WGSL let/var/const, GLSL `[const] T a = e, b;`
    bool declaration(std::vector<StmtP>& out)
    {
        if (lang == Lang::Wgsl) {
            auto s = make_s(Stmt::Decl);
            const std::string kw = peek().s;
            ++i;
            s->is_var = kw == "var";
            s->is_const = kw == "const";
            if (!ident(s->name)) return false;
            if (accept(":")) {
                if (!parse_type(s->decl_type)) return false;
                s->has_type = true;
            }
            if (accept("=")) {
                s->init = expr();
                if (!s->init) return false;
            } else if (!s->is_var) {
                return fail("'" + kw + "' needs an initializer");
            } else if (!s->has_type) {
                return fail("'var " + s->name + "' needs a type or an initializer");
            }
            out.push_back(s);
            return true;
        }
        bool is_const = accept("const");
        Type ty;
        if (!parse_type(ty)) return false;
        for (;;) {
            auto s = make_s(Stmt::Decl);
            s->is_var = !is_const;
            s->is_const = false;
            s->has_type = true;
            s->decl_type = ty;
            if (!ident(s->name)) return false;
            if (accept("=")) {
                s->init = expr();
                if (!s->init) return false;
            }
            out.push_back(s);
            if (!accept(",")) break;
        }
        return true;
    }
    bool statement(std::vector<StmtP>& out)
    {
        if (accept(";")) return true;
        if (is("{")) {
            auto s = make_s(Stmt::Block);
            if (!block(s->body)) return false;
            out.push_back(s);
            return true;
        }
        if (accept("return")) {
            auto s = make_s(Stmt::Return);
            if (!is(";")) {
                s->init = expr();
                if (!s->init) return false;
            }
            out.push_back(s);
            return expect(";");
        }
        if (accept("if")) {
            auto s = make_s(Stmt::If);
            if (lang == Lang::Glsl && !expect("(")) return false;
            s->init = expr();
            if (!s->init) return false;
            if (lang == Lang::Glsl && !expect(")")) return false;
            if (!body(s->body)) return false;
            if (accept("else")) {
                if (is("if")) {
                    if (!statement(s->else_body)) return false;
                } else if (!body(s->else_body)) {
                    return false;
                }
            }
            out.push_back(s);
            return true;
        }
        if (accept("for")) {
            auto s = make_s(Stmt::For);
            if (!expect("(")) return false;
            if (!accept(";")) {
                std::vector<StmtP> init;
                if (is("let") || is("var") || is("const") || (lang == Lang::Glsl && at_type())) {
                    if (!declaration(init)) return false;
                } else {
                    StmtP x = simple();
                    if (!x) return false;
                    init.push_back(x);
                }
                if (init.size() != 1) return fail("one declaration expected in a for initializer");
                s->for_init = init[0];
                if (!expect(";")) return false;
            }
            if (!is(";")) {
                s->init = expr();
                if (!s->init) return false;
            }
            if (!expect(";")) return false;
            if (!is(")")) {
                s->for_update = simple();
                if (!s->for_update) return false;
            }
            if (!expect(")")) return false;
            if (!body(s->body)) return false;
            out.push_back(s);
            return true;
        }
        if (accept("while")) {
            auto s = make_s(Stmt::While);
            if (lang == Lang::Glsl && !expect("(")) return false;
            s->init = expr();
            if (!s->init) return false;
            if (lang == Lang::Glsl && !expect(")")) return false;
            if (!body(s->body)) return false;
            out.push_back(s);
            return true;
        }
        if (accept("break")) {
            out.push_back(make_s(Stmt::Break));
            return expect(";");
        }
        if (accept("continue")) {
            out.push_back(make_s(Stmt::Continue));
            return expect(";");
        }
        if ((lang == Lang::Wgsl && (is("let") || is("var") || is("const"))) ||
            (lang == Lang::Glsl && at_type() && peek(1).k == Tok::Id)) {
            if (!declaration(out)) return false;
            return expect(";");
        }
        StmtP x = simple();
        if (!x) return false;
        out.push_back(x);
        return expect(";");
    }

    // ---- top level ----
    bool module(Module& m)
    {
        while (peek().k != Tok::End) {
            if (accept(";")) continue;
            if (lang == Lang::Wgsl) {
                while (accept("@")) {  // attributes
                    std::string a;
                    if (!ident(a)) return false;
                    if (accept("(")) {
                        int d = 1;
                        while (d && peek().k != Tok::End) {
                            if (is("(")) ++d;
                            if (is(")")) --d;
                            ++i;
                        }
                    }
                }
                if (accept("const")) {
                    GlobalConst c;
                    c.line = peek().line;
                    if (!ident(c.name)) return false;
                    if (accept(":")) {
                        if (!parse_type(c.type)) return false;
                        c.has_type = true;
                    }
                    if (!expect("=")) return false;
                    c.init = expr();
                    if (!c.init || !expect(";")) return false;
                    m.consts.push_back(c);
                    continue;
                }
                if (accept("fn")) {
                    Function f;
                    f.line = peek().line;
                    if (!ident(f.name) || !expect("(")) return false;
                    if (!accept(")")) {
                        for (;;) {
                            Param p;
                            if (!ident(p.name) || !expect(":") || !parse_type(p.type)) return false;
                            f.params.push_back(p);
                            if (accept(")")) break;
                            if (!expect(",")) return false;
                        }
                    }
                    if (accept("->")) {
                        if (!parse_type(f.ret)) return false;
                    }
                    if (!block(f.body)) return false;
                    m.func_index[f.name] = m.funcs.size();
                    m.funcs.push_back(std::move(f));
                    continue;
                }
                return fail("unsupported declaration '" + peek().s + "'");
            }
            // GLSL: [precision ...;] [const] T name = e;  |  T name(params) { ... }
            if (accept("precision")) {
                while (!accept(";")) {
                    if (peek().k == Tok::End) return fail("unterminated precision statement");
                    ++i;
                }
                continue;
            }
            bool is_const = accept("const");
            Type ret;
            if (is("void")) {
                ++i;
                ret = {Base::Void, 1};
            } else if (!parse_type(ret)) {
                return false;
            }
            std::string name;
            if (!ident(name)) return false;
            if (accept("(")) {
                Function f;
                f.line = peek().line;
                f.name = name;
                f.ret = ret;
                if (!accept(")")) {
                    if (is("void") && is(")", 1)) {
                        i += 2;
                    } else {
                        for (;;) {
                            accept("in");
                            accept("const");
                            Param p;
                            if (!parse_type(p.type) || !ident(p.name)) return false;
                            f.params.push_back(p);
                            if (accept(")")) break;
                            if (!expect(",")) return false;
                        }
                    }
                }
                if (accept(";")) continue;  // prototype
                if (!block(f.body)) return false;
                m.func_index[f.name] = m.funcs.size();
                m.funcs.push_back(std::move(f));
                continue;
            }
            if (!is_const) return fail("global variables are not supported (use const)");
            GlobalConst c;
            c.line = peek().line;
            c.name = name;
            c.type = ret;
            c.has_type = true;
            if (!expect("=")) return false;
            c.init = expr();
            if (!c.init || !expect(";")) return false;
            m.consts.push_back(c);
        }
        return true;
    }
};

}  // namespace

bool parse(const std::string& src, Lang lang, Module& m, Error& err)
{
    Parser p;
    p.lang = lang;
    if (!lex(src, lang, p.t, err)) return false;
    if (!p.module(m)) {
        err = p.err;
        return false;
    }
    return true;
}

}  // namespace s2m::fe
