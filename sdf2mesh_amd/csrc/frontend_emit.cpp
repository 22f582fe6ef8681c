// frontend_emit.cpp - the two lowerings of a checked program and the front-end driver.
//
//   lower_to_tape  symbolic evaluation of the entry function: the position p, p - c,
//                  constants and distances; library primitives at p - c with constant
//                  operands become tape primitives, min / max / max(-a, b) and the smooth
//                  operators with a constant k become tape operators, user functions are
//                  inlined.  Anything else (loops, swizzled or scaled positions, arithmetic
//                  on distances) is left to the compiled path.
//   emit_hip       HIP C++ for hiprtc: one device function per user function over the
//                  vector types of rtc_prelude.h, every constant written as an exact
//                  hexadecimal fp32 literal, every loop bounded.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <set>
#include <sstream>

#include "frontend.h"

namespace s2m::fe {

// =======================================================================================
// Tape lowering
// =======================================================================================
namespace {

struct TNode {
    s2m_op op{};
    int a = -1, b = -1;  // operator operands (node indices)
};

// Symbolic value.
struct SVal {
    enum K { Bad, Const, Pos, Dist } k = Bad;
    Type type;
    std::vector<double> c;          // Const: components
    float off[3] = {0.0f, 0.0f, 0.0f};  // Pos: p - off
    bool raw = true;                // Pos: p itself (no subtraction applied)
    int node = -1;                  // Dist
    bool neg = false;               // Dist: negated
};

SVal make_const(const Type& t, std::vector<double> c)
{
    SVal v;
    v.k = SVal::Const;
    v.type = t;
    v.c = std::move(c);
    return v;
}

struct Lowerer {
    const Module& m;
    Lang lang;
    std::vector<TNode> nodes;
    std::vector<std::map<std::string, SVal>> scopes;
    std::string why;
    int depth = 0;

    SVal bad(const std::string& w)
    {
        if (why.empty()) why = w;
        return SVal{};
    }

    SVal* lookup(const std::string& n)
    {
        for (auto it = scopes.rbegin(); it != scopes.rend(); ++it) {
            auto f = it->find(n);
            if (f != it->end()) return &f->second;
        }
        return nullptr;
    }

    int add(uint32_t code, const std::vector<float>& f, int a = -1, int b = -1)
    {
        TNode t;
        t.op.code = code;
        for (size_t i = 0; i < f.size() && i < 10; ++i) t.op.f[i] = f[i];
        t.a = a;
        t.b = b;
        nodes.push_back(t);
        return (int)nodes.size() - 1;
    }
    SVal dist(int node)
    {
        SVal v;
        v.k = SVal::Dist;
        v.type = {Base::F32, 1};
        v.node = node;
        return v;
    }

    // a constant as n fp32 components (abstract values round to nearest)
    static std::vector<float> f32s(const SVal& v, int n)
    {
        const auto c = convert_const(v.c, v.type, {Base::F32, (uint8_t)n});
        std::vector<float> r;
        for (double d : c) r.push_back((float)d);
        return r;
    }
    static bool is_f32_const(const SVal& v, int n)
    {
        return v.k == SVal::Const && (v.type.n == n || v.type.n == 1) &&
               (v.type.is_float() || v.type.base == Base::AbsInt);
    }

    // Convert a constant to the type it is bound to (parameter, declaration, return).
    SVal bind(const SVal& v, const Type& t)
    {
        if (v.k != SVal::Const || t.base == Base::Void) return v;
        return make_const(t, convert_const(v.c, v.type, t));
    }

    SVal binary(const std::string& op, const SVal& a, const SVal& b)
    {
        if (a.k == SVal::Bad || b.k == SVal::Bad) return SVal{};
        if (a.k == SVal::Const && b.k == SVal::Const) {
            Type t;
            if (!unify(a.type, b.type, t, lang)) return bad("constant operand types");
            const auto x = convert_const(a.c, a.type, t), y = convert_const(b.c, b.type, t);
            const bool cmp = op == "<" || op == ">" || op == "<=" || op == ">=" || op == "==" || op == "!=";
            if (cmp) {
                if (t.n != 1) return bad("vector comparison");
                const double p = x[0], q = y[0];
                return make_const({Base::Bool, 1}, {(double)(op == "<" ? p < q : op == ">" ? p > q : op == "<=" ? p <= q
                                                             : op == ">=" ? p >= q : op == "==" ? p == q : p != q)});
            }
            if (op == "&&" || op == "||")
                return make_const({Base::Bool, 1}, {(double)(op == "&&" ? (x[0] && y[0]) : (x[0] || y[0]))});
            std::vector<double> r;
            for (int k = 0; k < t.n; ++k) {
                bool ok;
                r.push_back(fold_arith(op, x[k], y[k], {t.base, 1}, ok));
                if (!ok) return bad("constant operation " + op);
            }
            return make_const(t, r);
        }
        if (a.k == SVal::Pos && b.k == SVal::Const && (op == "-" || op == "+") && a.raw &&
            is_f32_const(b, 3)) {
            // p - c and p + c (= p - (-c) exactly in IEEE arithmetic)
            const auto c = f32s(b, 3);
            SVal r = a;
            r.raw = false;
            for (int k = 0; k < 3; ++k) r.off[k] = op == "-" ? c[k] : -c[k];
            return r;
        }
        return bad("arithmetic on the position or on a distance");
    }

    SVal expr(const Expr& e)
    {
        if (e.is_const) return make_const(e.type, e.cval);
        switch (e.kind) {
        case Expr::Lit: return make_const(e.type, e.cval);
        case Expr::Ident: {
            SVal* v = lookup(e.name);
            if (!v) return bad("unknown name " + e.name);
            return *v;
        }
        case Expr::Member: {
            SVal a = expr(*e.args[0]);
            if (a.k == SVal::Pos && e.name == "xyz") return a;
            if (a.k != SVal::Const) return bad("swizzle of the position");
            std::vector<double> c;
            for (char ch : e.name) {
                const int k = ch == 'x' || ch == 'r' ? 0 : ch == 'y' || ch == 'g' ? 1 : ch == 'z' || ch == 'b' ? 2 : 3;
                c.push_back(a.c[a.type.n == 1 ? 0 : k]);
            }
            return make_const({a.type.base, (uint8_t)c.size()}, c);
        }
        case Expr::Index: {
            SVal a = expr(*e.args[0]);
            SVal i = expr(*e.args[1]);
            if (a.k != SVal::Const || i.k != SVal::Const) return bad("indexing");
            const int k = (int)i.c[0];
            if (k < 0 || k >= a.type.n) return bad("index out of range");
            return make_const({a.type.base, 1}, {a.c[k]});
        }
        case Expr::Unary: {
            SVal a = expr(*e.args[0]);
            if (e.op == "-") {
                if (a.k == SVal::Const) {
                    for (double& d : a.c) d = -d;
                    return a;
                }
                if (a.k == SVal::Dist) {
                    a.neg = !a.neg;
                    return a;
                }
                return bad("negated position");
            }
            if (e.op == "!" && a.k == SVal::Const) return make_const(a.type, {a.c[0] ? 0.0 : 1.0});
            return bad("operator " + e.op);
        }
        case Expr::Binary: return binary(e.op, expr(*e.args[0]), expr(*e.args[1]));
        case Expr::Call: return call(e);
        }
        return SVal{};
    }

    bool user_defined(const std::string& f) const { return m.func_index.count(f) != 0; }

    SVal call(const Expr& e)
    {
        const std::string& f = e.name;
        std::vector<SVal> a;
        for (const auto& x : e.args) {
            a.push_back(expr(*x));
            if (a.back().k == SVal::Bad) return SVal{};
        }
        auto all_const = [&] {
            for (const SVal& v : a)
                if (v.k != SVal::Const) return false;
            return true;
        };
        // constructors and conversions of constants
        const bool ctor = (f.size() >= 4 && f.compare(0, 3, "vec") == 0) || f == "ivec2" || f == "ivec3" ||
                          f == "ivec4" || f == "f32" || f == "float" || f == "i32" || f == "int" || f == "u32" ||
                          f == "uint";
        if (ctor && !user_defined(f)) {
            if (!all_const()) return bad("constructor of non-constant values");
            const Type t = e.type;
            std::vector<double> c;
            if (a.empty()) c.assign(t.n, 0.0);
            else if (a.size() == 1 && a[0].type.n == 1)
                c.assign(t.n, convert_const(a[0].c, a[0].type, {t.base, 1})[0]);
            else
                for (const SVal& v : a) {
                    const auto x = convert_const(v.c, v.type, {t.base, v.type.n});
                    c.insert(c.end(), x.begin(), x.end());
                }
            return make_const(t, c);
        }
        if (!user_defined(f)) {
            // library primitives at p - c with constant operands
            auto pos_args = [&](size_t n) { return a.size() == n && a[0].k == SVal::Pos; };
            auto offs = [&](const SVal& p) { return std::vector<float>{p.off[0], p.off[1], p.off[2]}; };
            auto cat = [](std::vector<float> x, const std::vector<float>& y) {
                x.insert(x.end(), y.begin(), y.end());
                return x;
            };
            if (f == "sdf3d_sphere" && pos_args(2) && is_f32_const(a[1], 1))
                return dist(add(S2M_OP_SPHERE, cat(offs(a[0]), f32s(a[1], 1))));
            if (f == "sdf3d_box" && pos_args(2) && is_f32_const(a[1], 3))
                return dist(add(S2M_OP_BOX, cat(offs(a[0]), f32s(a[1], 3))));
            if (f == "sdf3d_torus" && pos_args(2) && is_f32_const(a[1], 2))
                return dist(add(S2M_OP_TORUS, cat(offs(a[0]), f32s(a[1], 2))));
            if (f == "sdf3d_cylinder" && pos_args(3) && is_f32_const(a[1], 1) && is_f32_const(a[2], 1))
                return dist(add(S2M_OP_CYLINDER, cat(cat(offs(a[0]), f32s(a[1], 1)), f32s(a[2], 1))));
            if (f == "sdf3d_capsule" && pos_args(4) && is_f32_const(a[1], 3) && is_f32_const(a[2], 3) &&
                is_f32_const(a[3], 1)) {
                // the tape's capsule takes p itself; p - (+0) is p for every p
                for (int k = 0; k < 3; ++k)
                    if (a[0].off[k] != 0.0f || std::signbit(a[0].off[k])) return bad("capsule at a moved position");
                return dist(add(S2M_OP_CAPSULE, cat(cat(f32s(a[1], 3), f32s(a[2], 3)), f32s(a[3], 1))));
            }
            auto plain = [](const SVal& v) { return v.k == SVal::Dist && !v.neg; };
            if ((f == "min" || f == "max") && a.size() == 2 && a[0].k == SVal::Dist && plain(a[1])) {
                if (f == "min" && !a[0].neg) return dist(add(S2M_OP_UNION, {0.0f}, a[0].node, a[1].node));
                if (f == "max")
                    return dist(add(a[0].neg ? S2M_OP_SUBTRACTION : S2M_OP_INTERSECTION, {0.0f}, a[0].node,
                                    a[1].node));
            }
            static const std::map<std::string, uint32_t> smooth = {
                {"sdf_op_smooth_union", S2M_OP_SMOOTH_UNION},
                {"sdf_op_smooth_intersection", S2M_OP_SMOOTH_INTERSECTION},
                {"sdf_op_smooth_subtraction", S2M_OP_SMOOTH_SUBTRACTION}};
            auto s = smooth.find(f);
            if (s != smooth.end() && a.size() == 3 && plain(a[0]) && plain(a[1]) && is_f32_const(a[2], 1))
                return dist(add(s->second, f32s(a[2], 1), a[0].node, a[1].node));
            return bad("call of " + f + " with these operands");
        }
        const Function& fn = m.funcs[m.func_index.at(f)];
        if (++depth > 64) return bad("call depth");
        std::vector<std::map<std::string, SVal>> saved;
        saved.swap(scopes);
        scopes.emplace_back();
        for (size_t i = 0; i < fn.params.size(); ++i) scopes.back()[fn.params[i].name] = bind(a[i], fn.params[i].type);
        SVal ret;
        const int flow = block(fn.body, ret);
        scopes.swap(saved);
        --depth;
        if (flow != kReturned) return bad(flow == kFail ? why : "function " + f + " does not return");
        return bind(ret, fn.ret);
    }

    enum { kNext, kReturned, kFail };

    int block(const std::vector<StmtP>& body, SVal& ret)
    {
        scopes.emplace_back();
        int flow = kNext;
        for (const StmtP& s : body) {
            flow = stmt(*s, ret);
            if (flow != kNext) break;
        }
        scopes.pop_back();
        return flow;
    }

    int stmt(const Stmt& s, SVal& ret)
    {
        switch (s.kind) {
        case Stmt::Decl: {
            SVal v = s.init ? expr(*s.init) : make_const(s.decl_type, std::vector<double>(s.decl_type.n, 0.0));
            if (v.k == SVal::Bad) return kFail;
            scopes.back()[s.name] = bind(v, s.decl_type);
            return kNext;
        }
        case Stmt::Assign: {
            if (s.lhs->kind != Expr::Ident) {
                bad("assignment to a component");
                return kFail;
            }
            SVal* cur = lookup(s.lhs->name);
            if (!cur) {
                bad("assignment to " + s.lhs->name);
                return kFail;
            }
            SVal v = expr(*s.init);
            if (s.op != "=") v = binary(s.op.substr(0, 1), *cur, v);
            if (v.k == SVal::Bad) return kFail;
            *cur = bind(v, s.lhs->type);
            return kNext;
        }
        case Stmt::Return:
            if (!s.init) {
                bad("return without a value");
                return kFail;
            }
            ret = expr(*s.init);
            return ret.k == SVal::Bad ? kFail : kReturned;
        case Stmt::If: {
            SVal c = expr(*s.init);
            if (c.k != SVal::Const) {
                bad("condition that is not constant");
                return kFail;
            }
            return block(c.c[0] ? s.body : s.else_body, ret);
        }
        case Stmt::Block: return block(s.body, ret);
        case Stmt::ExprS: return kNext;
        default: bad("loop"); return kFail;
        }
    }

    // postfix emission; sizes memoised so shared subtrees cannot make this exponential
    bool emit(int n, std::vector<s2m_op>& out, std::vector<uint64_t>& size)
    {
        std::function<uint64_t(int)> sz = [&](int k) -> uint64_t {
            if (size[k]) return size[k];
            const TNode& t = nodes[k];
            uint64_t s = 1;
            if (t.a >= 0) s += std::min<uint64_t>(sz(t.a) + sz(t.b), 1ull << 40);
            return size[k] = s;
        };
        if (sz(n) > S2M_MAX_TAPE) return false;
        std::function<void(int)> rec = [&](int k) {
            const TNode& t = nodes[k];
            if (t.a >= 0) {
                rec(t.a);
                rec(t.b);
            }
            out.push_back(t.op);
        };
        rec(n);
        return true;
    }
};

}  // namespace

bool lower_to_tape(const Module& m, Lang lang, const std::string& entry, std::vector<s2m_op>& tape,
                   std::string& why)
{
    tape.clear();
    if (m.func_index.count("sdf3d_normal")) {
        why = "user-defined sdf3d_normal";
        return false;
    }
    Lowerer L{m, lang};
    // global constants are folded by the checker wherever they are used
    SVal p;
    p.k = SVal::Pos;
    p.type = {Base::F32, 3};
    Expr call;
    call.kind = Expr::Call;
    call.name = entry;
    call.type = {Base::F32, 1};
    auto arg = std::make_shared<Expr>();
    arg->kind = Expr::Ident;
    arg->name = "\x01p";
    arg->type = p.type;
    call.args.push_back(arg);
    L.scopes.emplace_back();
    L.scopes.back()["\x01p"] = p;
    SVal r = L.call(call);
    if (r.k != SVal::Dist || r.neg) {
        why = L.why.empty() ? "the result is not a library distance" : L.why;
        return false;
    }
    std::vector<uint64_t> size(L.nodes.size(), 0);
    if (!L.emit(r.node, tape, size)) {
        why = "tape longer than the tape limit";
        tape.clear();
        return false;
    }
    // operand stack depth
    uint32_t sp = 0, maxsp = 0;
    for (const s2m_op& o : tape) {
        sp = o.code < S2M_OP_UNION ? sp + 1 : sp - 1;
        maxsp = std::max(maxsp, sp);
    }
    if (maxsp > S2M_MAX_STACK) {
        why = "operand stack deeper than the tape interpreter's";
        tape.clear();
        return false;
    }
    return true;
}

// =======================================================================================
// HIP C++ generation
// =======================================================================================
namespace {

constexpr const char* kLoopBound = "16777216u";  // iterations per loop entry

std::string f32_literal(double v)
{
    const float f = (float)v;
    if (std::isnan(f)) return "__builtin_nanf(\"\")";
    if (std::isinf(f)) return f > 0 ? "__builtin_inff()" : "(-__builtin_inff())";
    char b[64];
    std::snprintf(b, sizeof b, "%af", (double)f);
    return std::string("(") + b + ")";
}

std::string ctype(const Type& t0)
{
    const Type t = concrete(t0);
    const char* s = t.base == Base::F32 ? "float" : t.base == Base::I32 ? "int" : t.base == Base::U32 ? "unsigned"
                    : t.base == Base::Bool ? "bool" : "void";
    if (t.n == 1) return s;
    return std::string("s2m::rtc::V<") + s + ", " + std::to_string(t.n) + ">";
}

std::string scalar_literal(double v, Base b)
{
    switch (b) {
    case Base::F32:
    case Base::AbsFloat: return f32_literal(v);
    case Base::I32:
    case Base::AbsInt: {
        const long long x = (long long)(int32_t)(long long)v;
        return x == -2147483648LL ? "(-2147483647 - 1)" : "(" + std::to_string(x) + ")";
    }
    case Base::U32: return std::to_string((unsigned long long)(uint32_t)(long long)v) + "u";
    case Base::Bool: return v != 0 ? "true" : "false";
    default: return "0";
    }
}

struct Emitter {
    const Module& m;
    Lang lang;
    std::ostringstream o;
    Error err{0, ""};
    int tmp = 0;

    bool fail(int line, const std::string& msg)
    {
        if (!err.code) err = {S2M_E_WGSL, "line " + std::to_string(line) + ": " + msg};
        return false;
    }

    static std::string uname(const std::string& n) { return "u_" + n; }

    std::string literal(const std::vector<double>& c, const Type& from, const Type& to0)
    {
        const Type to = concrete(to0);
        const auto v = convert_const(c, from, to);
        if (to.n == 1) return scalar_literal(v[0], to.base);
        std::string s = ctype(to) + "{{";
        for (int k = 0; k < to.n; ++k) s += (k ? ", " : "") + scalar_literal(v[k], to.base);
        return s + "}}";
    }

    // e as a value of type `want` (Void: its own concrete type)
    std::string ex(const Expr& e, Type want = {})
    {
        if (want.base == Base::Void) want = concrete(e.type);
        if (e.is_const) return literal(e.cval, e.type, want);
        std::string s = raw(e);
        const Type t = concrete(e.type);
        if (t.base != want.base) {
            // GLSL int -> float promotion (the checker admits no other implicit conversion)
            if (t.n == 1) s = "static_cast<" + ctype({want.base, 1}) + ">(" + s + ")";
            else return fail(e.line, "implicit vector conversion"), "0";
        }
        if (t.n == 1 && want.n > 1) s = "s2m::rtc::splat<" + ctype({want.base, 1}) + ", " + std::to_string(want.n) + ">(" + s + ")";
        return s;
    }

    static int comp(char ch)
    {
        return ch == 'x' || ch == 'r' ? 0 : ch == 'y' || ch == 'g' ? 1 : ch == 'z' || ch == 'b' ? 2 : 3;
    }

    std::string raw(const Expr& e)
    {
        switch (e.kind) {
        case Expr::Lit: return literal(e.cval, e.type, e.type);
        case Expr::Ident: return uname(e.name);
        case Expr::Member: {
            const std::string a = "(" + ex(*e.args[0]) + ")";
            if (e.name.size() == 1) return a + ".v[" + std::to_string(comp(e.name[0])) + "]";
            std::string k;
            for (char ch : e.name) k += (k.empty() ? "" : ", ") + std::to_string(comp(ch));
            return "s2m::rtc::sw<" + k + ">" + a;
        }
        case Expr::Index: return "(" + ex(*e.args[0]) + ").v[" + std::to_string((int)e.args[1]->cval[0]) + "]";
        case Expr::Unary: return "(" + e.op + ex(*e.args[0]) + ")";
        case Expr::Binary: {
            const Expr& a = *e.args[0];
            const Expr& b = *e.args[1];
            if (e.op == "&&" || e.op == "||") return "(" + ex(a) + " " + e.op + " " + ex(b) + ")";
            Type t;
            unify(a.type, b.type, t, lang);
            t = concrete(t);
            const std::string x = ex(a, t), y = ex(b, t);
            if (e.op == "%" || (e.op == "/" && !t.is_float()))
                return std::string(e.op == "%" ? "s2m::rtc::b_rem(" : "s2m::rtc::b_div(") + x + ", " + y + ")";
            return "(" + x + " " + e.op + " " + y + ")";
        }
        case Expr::Call: return call(e);
        }
        return "0";
    }

    std::string args(const Expr& e, const std::vector<Type>& ts)
    {
        std::string s;
        for (size_t i = 0; i < e.args.size(); ++i)
            s += (i ? ", " : "") + ex(*e.args[i], ts.size() == 1 ? ts[0] : ts[i]);
        return s;
    }

    std::string call(const Expr& e)
    {
        const std::string& f = e.name;
        const Type rt = concrete(e.type);
        auto uf = m.func_index.find(f);
        if (uf != m.func_index.end()) {
            const Function& fn = m.funcs[uf->second];
            std::vector<Type> ts;
            for (const Param& p : fn.params) ts.push_back(p.type);
            return uname(f) + "(" + args(e, ts) + ")";
        }
        // constructors and conversions
        const bool vctor = (f.size() >= 4 && f.compare(0, 3, "vec") == 0) || f.compare(0, 4, "ivec") == 0;
        if (vctor) {
            const Type el{rt.base, 1};
            std::string s = ctype(rt) + "{{";
            if (e.args.empty()) {
                for (int k = 0; k < rt.n; ++k) s += (k ? ", " : "") + scalar_literal(0, rt.base);
                return s + "}}";
            }
            if (e.args.size() == 1 && e.args[0]->type.n == 1)
                return "s2m::rtc::splat<" + ctype(el) + ", " + std::to_string(rt.n) + ">(" + ex(*e.args[0], el) + ")";
            int k = 0;
            for (const auto& a : e.args) {
                if (a->type.n == 1) {
                    s += (k++ ? ", " : "") + ex(*a, el);
                    continue;
                }
                // a vector argument: bind it once, then take its components
                const Type at = concrete(a->type);
                const std::string v = "(" + ex(*a) + ")";
                if (at.base != rt.base) return fail(e.line, "vector argument of another element type"), "0";
                for (int j = 0; j < at.n; ++j) s += (k++ ? ", " : "") + v + ".v[" + std::to_string(j) + "]";
            }
            return s + "}}";
        }
        if (f == "f32" || f == "float" || f == "i32" || f == "int" || f == "u32" || f == "uint" || f == "bool")
            return "static_cast<" + ctype(rt) + ">(" + ex(*e.args[0]) + ")";
        // library functions
        static const std::map<std::string, std::pair<const char*, std::vector<Type>>> lib = {
            {"sdf3d_sphere", {"lib_sphere", {{Base::F32, 3}, {Base::F32, 1}}}},
            {"sdf3d_box", {"lib_box", {{Base::F32, 3}, {Base::F32, 3}}}},
            {"sdf3d_torus", {"lib_torus", {{Base::F32, 3}, {Base::F32, 2}}}},
            {"sdf3d_cylinder", {"lib_cylinder", {{Base::F32, 3}, {Base::F32, 1}, {Base::F32, 1}}}},
            {"sdf3d_capsule", {"lib_capsule", {{Base::F32, 3}, {Base::F32, 3}, {Base::F32, 3}, {Base::F32, 1}}}},
            {"sdf_op_smooth_union", {"lib_smooth_union", {{Base::F32, 1}, {Base::F32, 1}, {Base::F32, 1}}}},
            {"sdf_op_smooth_intersection", {"lib_smooth_intersection", {{Base::F32, 1}, {Base::F32, 1}, {Base::F32, 1}}}},
            {"sdf_op_smooth_subtraction", {"lib_smooth_subtraction", {{Base::F32, 1}, {Base::F32, 1}, {Base::F32, 1}}}},
        };
        auto l = lib.find(f);
        if (l != lib.end()) return std::string("s2m::rtc::") + l->second.first + "(" + args(e, l->second.second) + ")";
        if (f == "sdf3d_normal") return "s2m_lib_normal(" + args(e, {{Base::F32, 3}, {Base::F32, 1}}) + ")";
        // builtins: arguments in the unified type
        Type t = concrete(e.args.empty() ? rt : e.args[0]->type);
        for (size_t i = 1; i < e.args.size(); ++i) {
            if (f == "select" && i == 2) break;
            Type u;
            unify(t, e.args[i]->type, u, lang);
            t = concrete(u);
        }
        if (f == "select")
            return "((" + ex(*e.args[2]) + ") ? (" + ex(*e.args[1], t) + ") : (" + ex(*e.args[0], t) + "))";
        static const std::map<std::string, const char*> names = {
            {"abs", "b_abs"}, {"min", "b_min"}, {"max", "b_max"}, {"clamp", "b_clamp"}, {"mix", "b_mix"},
            {"step", "b_step"}, {"smoothstep", "b_smoothstep"}, {"sign", "b_sign"}, {"sqrt", "b_sqrt"},
            {"inverseSqrt", "b_inverse_sqrt"}, {"inversesqrt", "b_inverse_sqrt"}, {"sin", "b_sin"},
            {"cos", "b_cos"}, {"tan", "b_tan"}, {"asin", "b_asin"}, {"acos", "b_acos"}, {"sinh", "b_sinh"},
            {"cosh", "b_cosh"}, {"tanh", "b_tanh"}, {"exp", "b_exp"}, {"exp2", "b_exp2"}, {"log", "b_log"},
            {"log2", "b_log2"}, {"pow", "b_pow"}, {"floor", "b_floor"}, {"ceil", "b_ceil"},
            {"trunc", "b_trunc"}, {"round", "b_round"}, {"fract", "b_fract"}, {"mod", "b_mod"},
            {"radians", "b_radians"}, {"degrees", "b_degrees"}, {"atan2", "b_atan2"}, {"length", "b_length"},
            {"distance", "b_distance"}, {"dot", "b_dot"}, {"normalize", "b_normalize"}, {"cross", "b_cross"}};
        std::string name;
        if (f == "atan") name = e.args.size() == 2 ? "b_atan2" : "b_atan";
        else {
            auto n = names.find(f);
            if (n == names.end()) return fail(e.line, "no code for " + f), "0";
            name = n->second;
        }
        return std::string("s2m::rtc::") + name + "(" + args(e, {t}) + ")";
    }

    // ---- statements ----
    std::string lvalue(const Expr& e)
    {
        if (e.kind == Expr::Ident) return uname(e.name);
        if (e.kind == Expr::Member && e.name.size() == 1)
            return "(" + lvalue(*e.args[0]) + ").v[" + std::to_string(comp(e.name[0])) + "]";
        if (e.kind == Expr::Index) return "(" + lvalue(*e.args[0]) + ").v[" + std::to_string((int)e.args[1]->cval[0]) + "]";
        fail(e.line, "unsupported assignment target");
        return "";
    }

    // the value assigned by `lhs op= rhs` / `lhs = rhs`, in the type of lhs
    std::string assigned(const Stmt& s)
    {
        const Type lt = concrete(s.lhs->type);
        if (s.op == "=") return ex(*s.init, lt);
        const std::string op = s.op.substr(0, 1);
        const std::string cur = "(" + ex(*s.lhs) + ")";
        const std::string rhs = ex(*s.init, lt);
        if (op == "%" || (op == "/" && !lt.is_float()))
            return std::string(op == "%" ? "s2m::rtc::b_rem(" : "s2m::rtc::b_div(") + cur + ", " + rhs + ")";
        return "(" + cur + " " + op + " " + rhs + ")";
    }

    void assign(const Stmt& s, const std::string& ind)
    {
        const Expr& l = *s.lhs;
        if (l.kind == Expr::Member && l.name.size() > 1) {
            // multi-component swizzle target: evaluate once, store per component
            const std::string t = "s2m_t" + std::to_string(tmp++);
            o << ind << "{ const auto " << t << " = " << assigned(s) << ";";
            const std::string base = lvalue(*l.args[0]);
            for (size_t k = 0; k < l.name.size(); ++k)
                o << " (" << base << ").v[" << comp(l.name[k]) << "] = " << t << ".v[" << k << "];";
            o << " }\n";
            return;
        }
        o << ind << lvalue(l) << " = " << assigned(s) << ";\n";
    }

    std::string assign_expr(const Stmt& s)
    {
        if (s.lhs->kind == Expr::Member && s.lhs->name.size() > 1) {
            fail(s.line, "swizzle assignment in a loop header");
            return "";
        }
        return lvalue(*s.lhs) + " = " + assigned(s);
    }

    void body(const std::vector<StmtP>& b, const std::string& ind)
    {
        for (const StmtP& s : b) stmt(*s, ind);
    }

    void decl(const Stmt& s, const std::string& ind, bool semicolon = true)
    {
        const Type t = concrete(s.decl_type);
        o << ind << ctype(t) << " " << uname(s.name) << " = ";
        if (s.init) o << ex(*s.init, t);
        else o << literal(std::vector<double>(t.n, 0.0), t, t);
        if (semicolon) o << ";\n";
    }

    void stmt(const Stmt& s, const std::string& ind)
    {
        const std::string in2 = ind + "    ";
        switch (s.kind) {
        case Stmt::Decl: decl(s, ind); break;
        case Stmt::Assign: assign(s, ind); break;
        case Stmt::Return:
            if (s.init) o << ind << "return " << ex(*s.init, cur_ret) << ";\n";
            else o << ind << "return;\n";
            break;
        case Stmt::If:
            o << ind << "if (" << ex(*s.init) << ") {\n";
            body(s.body, in2);
            o << ind << "}";
            if (!s.else_body.empty()) {
                o << " else {\n";
                body(s.else_body, in2);
                o << ind << "}";
            }
            o << "\n";
            break;
        case Stmt::For: {
            // for (init; guard && cond; update) with the guard bounding the trip count
            const std::string g = "s2m_g" + std::to_string(tmp++);
            o << ind << "{\n" << in2 << "unsigned " << g << " = 0;\n";
            if (s.for_init) {
                if (s.for_init->kind == Stmt::Decl) decl(*s.for_init, in2);
                else stmt(*s.for_init, in2);
            }
            o << in2 << "for (; " << g << "++ < " << kLoopBound;
            if (s.init) o << " && (" << ex(*s.init) << ")";
            o << ";";
            if (s.for_update) {  // an assignment, or any expression (GLSL's loop-expression)
                if (s.for_update->kind == Stmt::Assign) o << " " << assign_expr(*s.for_update);
                else o << " (void)(" << ex(*s.for_update->init) << ")";
            }
            o << ") {\n";
            body(s.body, in2 + "    ");
            o << in2 << "}\n" << ind << "}\n";
            break;
        }
        case Stmt::While: {
            const std::string g = "s2m_g" + std::to_string(tmp++);
            o << ind << "{\n" << in2 << "unsigned " << g << " = 0;\n";
            o << in2 << "while (" << g << "++ < " << kLoopBound << " && (" << ex(*s.init) << ")) {\n";
            body(s.body, in2 + "    ");
            o << in2 << "}\n" << ind << "}\n";
            break;
        }
        case Stmt::Break: o << ind << "break;\n"; break;
        case Stmt::Continue: o << ind << "continue;\n"; break;
        case Stmt::Block:
            o << ind << "{\n";
            body(s.body, in2);
            o << ind << "}\n";
            break;
        case Stmt::ExprS: o << ind << "(void)(" << ex(*s.init) << ");\n"; break;
        }
    }

    Type cur_ret;

    std::string signature(const Function& f)
    {
        std::string s = "__device__ " + ctype(f.ret) + " " + uname(f.name) + "(";
        for (size_t i = 0; i < f.params.size(); ++i)
            s += (i ? ", " : "") + ctype(f.params[i].type) + " " + uname(f.params[i].name);
        return s + ")";
    }
};

}  // namespace

bool emit_hip(const Module& m, Lang lang, const std::string& entry, std::string& hip, Error& err)
{
    Emitter E{m, lang};
    std::ostringstream& o = E.o;
    o << "// generated by the sdf2mesh front end\n";
    o << "__device__ float u_" << entry << "(s2m::rtc::vf3);\n";
    o << "__device__ __forceinline__ float s2m_entry(s2m::rtc::vf3 p) { return u_" << entry << "(p); }\n";
    // sdf3d_normal of the library, sdf3d_normal.wgsl:4-9 (unnormalized)
    o << "__device__ __forceinline__ s2m::rtc::vf3 s2m_lib_normal(s2m::rtc::vf3 p, float e)\n{\n"
         "    const float d1 = s2m_entry(s2m::rtc::vf3{{p.v[0] + e, p.v[1] + (-e), p.v[2] + (-e)}});\n"
         "    const float d2 = s2m_entry(s2m::rtc::vf3{{p.v[0] + (-e), p.v[1] + (-e), p.v[2] + e}});\n"
         "    const float d3 = s2m_entry(s2m::rtc::vf3{{p.v[0] + (-e), p.v[1] + e, p.v[2] + (-e)}});\n"
         "    const float d4 = s2m_entry(s2m::rtc::vf3{{p.v[0] + e, p.v[1] + e, p.v[2] + e}});\n"
         "    return s2m::rtc::vf3{{((d1 + (-d2)) + (-d3)) + d4, (((-d1) + (-d2)) + d3) + d4,\n"
         "                         (((-d1) + d2) + (-d3)) + d4}};\n}\n";
    for (const Function& f : m.funcs) {
        if (lang == Lang::Glsl && (f.name == "main" || f.name == "main_1")) continue;
        o << E.signature(f) << ";\n";
    }
    for (const Function& f : m.funcs) {
        if (lang == Lang::Glsl && (f.name == "main" || f.name == "main_1")) continue;
        E.cur_ret = f.ret;
        o << E.signature(f) << "\n{\n";
        E.body(f.body, "    ");
        if (f.ret.base != Base::Void) o << "    return " << ctype(f.ret) << "{};  // not reached\n";
        o << "}\n";
        if (E.err.code) break;
    }
    if (E.err.code) {
        err = E.err;
        return false;
    }
    const bool user_normal = m.func_index.count("sdf3d_normal") != 0;
    o << "\nstruct s2m_user_fn {\n"
         "    __device__ float operator()(float x, float y, float z) const { return s2m_entry(s2m::rtc::vf3{{x, y, z}}); }\n"
         "};\n"
         "extern \"C\" __global__ __launch_bounds__(256) void s2m_c_eval_lattice(const float* __restrict__ cx,\n"
         "    const float* __restrict__ cy, const float* __restrict__ cz, uint32_t nx, uint32_t ny, uint32_t lz0,\n"
         "    uint32_t nzl, uint32_t pitch, float* __restrict__ lat)\n{\n"
         "    s2m::eval_lattice_point(s2m_user_fn{}, cx, cy, cz, nx, ny, lz0, nzl, pitch, lat);\n}\n"
         "struct s2m_user_normal {\n"
         "    __device__ void operator()(float px, float py, float pz, float eps, float* o) const\n    {\n";
    if (user_normal)
        o << "        const s2m::rtc::vf3 g = u_sdf3d_normal(s2m::rtc::vf3{{px, py, pz}}, eps);\n"
             "        s2m::write_unit(o, g.v[0], g.v[1], g.v[2]);\n";
    else
        o << "        s2m::library_normal(s2m_user_fn{}, px, py, pz, eps, o);\n";
    o << "    }\n};\n"
         "extern \"C\" __global__ __launch_bounds__(256) void s2m_c_normals(float eps, float* __restrict__ vtx,\n"
         "    const uint64_t* __restrict__ counters, uint64_t vcap)\n{\n"
         "    s2m::normals_range(s2m_user_normal{}, eps, vtx, counters, vcap);\n}\n";
    hip = o.str();
    return true;
}

// =======================================================================================
// Driver (shader.rs:44-104): preprocessing, parsing, entry resolution, lowering
// =======================================================================================
bool load_source(const std::string& text, bool is_path, Lang lang, std::string& source,
                 std::vector<std::string>& uses, Error& err)
{
    if (lang == Lang::Wgsl) {
        if (is_path) return preprocess_file(text, source, uses, err);
        preprocess_text(text, source, uses);
        return true;
    }
    if (!is_path) {
        source = text;
        return true;
    }
    source.clear();
    FILE* fp = std::fopen(text.c_str(), "rb");
    if (!fp) {
        err = {S2M_E_IO, "cannot read " + text};
        return false;
    }
    char buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, fp)) > 0) source.append(buf, n);
    std::fclose(fp);
    return true;
}

bool translate(const std::string& text, bool is_path, Lang lang, const std::string& sdf_name,
               bool want_hip, Translation& out, Error& err)
{
    std::vector<std::string> uses;
    if (!load_source(text, is_path, lang, out.source, uses, err)) return false;
    Module m;
    if (!parse(out.source, lang, m, err)) return false;
    Visible lib = visible_modules(uses);
    std::string entry = "sdf3d";
    if (lang == Lang::Glsl) {
        // from_glsl_fragment_shader: main is removed, the normal module is added, and a
        // wrapper `sdf3d(p) = <sdf>(p)` is generated unless sdf3d exists (shader.rs:84-98)
        if (!m.func_index.count("main")) {
            err = {S2M_E_SHADER, "Function fn main_1( not found in shader"};
            return false;
        }
        lib = Visible{};
        lib.normal = true;
        const std::string sdf = sdf_name.empty() ? "sdf" : sdf_name;
        if (!m.func_index.count(sdf)) {
            err = {S2M_E_MISSING_SDF, sdf};
            return false;
        }
        if (!m.func_index.count("sdf3d")) entry = sdf;
    }
    if (!check(m, lang, entry, lib, err)) return false;
    // the meshing shader calls sdf3d_normal(p, eps) (dualcontour.wgsl:171)
    auto nf = m.func_index.find("sdf3d_normal");
    if (nf == m.func_index.end()) {
        if (!lib.normal) {
            err = {S2M_E_VALIDATION, "sdf3d_normal is not defined (the program needs `use sdf3d::*;` or "
                                     "`use sdf3d::normal;`)"};
            return false;
        }
    } else {
        const Function& f = m.funcs[nf->second];
        if (f.params.size() != 2 || f.params[0].type != Type{Base::F32, 3} || f.params[1].type != Type{Base::F32, 1} ||
            f.ret != Type{Base::F32, 3}) {
            err = {S2M_E_VALIDATION, "sdf3d_normal must take (vec3<f32>, f32) and return vec3<f32>"};
            return false;
        }
    }
    out.tape.clear();
    out.tape_why.clear();
    lower_to_tape(m, lang, entry, out.tape, out.tape_why);
    if (want_hip || out.tape.empty()) {
        if (!emit_hip(m, lang, entry, out.hip, err)) return false;
    }
    return true;
}

}  // namespace s2m::fe
