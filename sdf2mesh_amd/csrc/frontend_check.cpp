// frontend_check.cpp - name resolution, types and constant folding of the front end's
// AST.
//
// WGSL rules: literals without a suffix are abstract; expressions of abstract operands
// are folded in double precision and converted (rounded to nearest) to f32 / i32 where
// they meet a concrete operand, a typed declaration or a parameter; `let` / `var`
// without a type take the concrete form.  Concrete f32 expressions with constant
// operands are folded in f32 arithmetic, exactly as the device would compute them;
// transcendental calls on concrete values are left to the device (its math library may
// differ from the host's in the last bits).  GLSL literals are f32 / i32 (naga types them).
#include <cmath>
#include <functional>
#include <set>

#include "frontend.h"

namespace s2m::fe {

namespace {

const Type kBool{Base::Bool, 1}, kF32{Base::F32, 1}, kI32{Base::I32, 1},
    kU32{Base::U32, 1}, kAF{Base::AbsFloat, 1}, kAI{Base::AbsInt, 1};

Type vec(Base b, int n) { return {b, (uint8_t)n}; }

// The library functions (sdf3d_primitives.wgsl:7-36, sdf_op.wgsl:7-23,
// sdf3d_normal.wgsl:4-9) and the module that makes each visible.
enum class LibModule { Primitives, Ops, Normal };
struct LibFn {
    const char* name;
    LibModule module;
    std::vector<Type> params;
};
const LibFn kLibrary[] = {
    {"sdf3d_box", LibModule::Primitives, {vec(Base::F32, 3), vec(Base::F32, 3)}},
    {"sdf3d_cylinder", LibModule::Primitives, {vec(Base::F32, 3), kF32, kF32}},
    {"sdf3d_capsule", LibModule::Primitives, {vec(Base::F32, 3), vec(Base::F32, 3), vec(Base::F32, 3), kF32}},
    {"sdf3d_sphere", LibModule::Primitives, {vec(Base::F32, 3), kF32}},
    {"sdf3d_torus", LibModule::Primitives, {vec(Base::F32, 3), vec(Base::F32, 2)}},
    {"sdf_op_smooth_union", LibModule::Ops, {kF32, kF32, kF32}},
    {"sdf_op_smooth_intersection", LibModule::Ops, {kF32, kF32, kF32}},
    {"sdf_op_smooth_subtraction", LibModule::Ops, {kF32, kF32, kF32}},
    {"sdf3d_normal", LibModule::Normal, {vec(Base::F32, 3), kF32}},
};

bool library_visible(const Visible& v, const LibFn& l)
{
    return l.module == LibModule::Primitives ? v.primitives : l.module == LibModule::Ops ? v.ops : v.normal;
}

}  // namespace

Visible visible_modules(const std::vector<std::string>& uses)
{
    Visible v;
    for (const std::string& u : uses) {
        if (u == "sdf::*" || u == "sdf::op") v.ops = true;
        if (u == "sdf3d::normal") v.normal = true;
        if (u == "sdf3d::primitives") v.primitives = true;
        if (u == "sdf3d::*") v.primitives = v.normal = true;
    }
    return v;
}

static std::string tname(const Type& t)
{
    static const char* B[] = {"void", "bool", "i32", "u32", "f32", "abstract-int", "abstract-float"};
    std::string s = B[(int)t.base];
    return t.n == 1 ? s : "vec" + std::to_string(t.n) + "<" + s + ">";
}

// concrete form of an abstract type
Type concrete(Type t)
{
    if (t.base == Base::AbsFloat) t.base = Base::F32;
    else if (t.base == Base::AbsInt) t.base = Base::I32;
    return t;
}

// Can a value of type `from` be used where `to` is expected (with abstract conversion
// and, in GLSL, int -> float promotion)?
bool convertible(const Type& from, const Type& to, Lang lang)
{
    if (from == to) return true;
    if (from.n != to.n) return false;
    if (from.base == Base::AbsFloat) return to.base == Base::F32;
    if (from.base == Base::AbsInt) return to.base == Base::F32 || to.base == Base::I32 || to.base == Base::U32 ||
                                          to.base == Base::AbsFloat;
    if (lang == Lang::Glsl && to.base == Base::F32 && (from.base == Base::I32 || from.base == Base::U32))
        return true;
    return false;
}

// the common type of two operands (binary arithmetic / comparison / builtins)
bool unify(const Type& a, const Type& b, Type& out, Lang lang)
{
    if (a == b) {
        out = a;
        return true;
    }
    Type sa = a, sb = b;
    sa.n = sb.n = 1;
    Type s;
    if (sa == sb) s = sa;
    else if (convertible(sa, sb, lang)) s = sb;
    else if (convertible(sb, sa, lang)) s = sa;
    else return false;
    if (a.n != 1 && b.n != 1 && a.n != b.n) return false;
    out = {s.base, (uint8_t)std::max(a.n, b.n)};
    return true;
}

// A binary operation on constant scalars in the arithmetic of type t: f32 values are
// rounded after every operation as the device rounds them, abstract values fold in
// double precision, i32 / u32 wrap.
double fold_arith(const std::string& op, double a, double b, const Type& t, bool& ok)
{
    ok = true;
    if (t.base == Base::F32) {
        const float x = (float)a, y = (float)b;
        if (op == "+") return (double)(float)(x + y);
        if (op == "-") return (double)(float)(x - y);
        if (op == "*") return (double)(float)(x * y);
        if (op == "/") return (double)(float)(x / y);
        if (op == "%") return (double)std::fmod(x, y);
    } else if (t.base == Base::AbsFloat) {
        if (op == "+") return a + b;
        if (op == "-") return a - b;
        if (op == "*") return a * b;
        if (op == "/") return a / b;
        if (op == "%") return std::fmod(a, b);
    } else {  // integers (i32 / u32 wrap; abstract ints fold exactly)
        const long long x = (long long)a, y = (long long)b;
        long long r;
        if (op == "+") r = x + y;
        else if (op == "-") r = x - y;
        else if (op == "*") r = x * y;
        else if (op == "/") {
            if (y == 0) { ok = false; return 0; }
            r = x / y;
        } else if (op == "%") {
            if (y == 0) { ok = false; return 0; }
            r = x % y;
        } else {
            ok = false;
            return 0;
        }
        if (t.base == Base::I32) r = (long long)(int32_t)(uint32_t)(uint64_t)r;
        if (t.base == Base::U32) r = (long long)(uint32_t)(uint64_t)r;
        return (double)r;
    }
    ok = false;
    return 0;
}

// Constant c of type `from` converted to type `to` (a scalar is broadcast).
std::vector<double> convert_const(const std::vector<double>& c, const Type& from, const Type& to)
{
    std::vector<double> r(to.n);
    for (int k = 0; k < to.n; ++k) {
        const double v = c[from.n == 1 ? 0 : k];
        if (to.base == Base::F32) r[k] = (double)(float)v;
        else if (to.base == Base::I32) r[k] = (double)(int32_t)(long long)v;
        else if (to.base == Base::U32) r[k] = (double)(uint32_t)(long long)v;
        else r[k] = v;
    }
    return r;
}


namespace {

float round_f32(double v) { return (float)v; }

struct Sym {
    Type type;
    bool is_const = false;
    std::vector<double> cval;
    bool mut = false;
};

struct Checker {
    Module& m;
    Lang lang;
    Visible lib;
    Error err{0, ""};
    std::map<std::string, std::set<std::string>> calls;  // call graph (recursion check)
    std::vector<std::map<std::string, Sym>> scopes;
    std::map<std::string, size_t> gconst_index;
    std::set<std::string> gconst_busy;
    const Function* cur = nullptr;

    bool fail(int line, const std::string& msg)
    {
        if (!err.code) err = {S2M_E_VALIDATION, "line " + std::to_string(line) + ": " + msg};
        return false;
    }

    Sym* lookup(const std::string& n)
    {
        for (auto it = scopes.rbegin(); it != scopes.rend(); ++it) {
            auto f = it->find(n);
            if (f != it->end()) return &f->second;
        }
        return nullptr;
    }

    // ---- expressions ----
    bool expr(Expr& e)
    {
        switch (e.kind) {
        case Expr::Lit: {
            if (e.is_bool) {
                e.type = kBool;
            } else if (e.suffix == 'f') {
                e.type = kF32;
            } else if (e.suffix == 'u') {
                e.type = kU32;
            } else if (e.suffix == 'i') {
                e.type = kI32;
            } else if (lang == Lang::Glsl) {
                e.type = e.is_float_lit ? kF32 : kI32;
            } else {
                e.type = e.is_float_lit ? kAF : kAI;
            }
            e.is_const = true;
            e.cval = {e.type.base == Base::F32 ? (double)round_f32(e.num) : e.num};
            return true;
        }
        case Expr::Ident: {
            Sym* s = lookup(e.name);
            if (!s) {
                auto g = gconst_index.find(e.name);
                if (g == gconst_index.end()) return fail(e.line, "unknown identifier '" + e.name + "'");
                if (!global_const(g->second)) return false;
                s = lookup(e.name);
            }
            e.type = s->type;
            e.is_const = s->is_const;
            e.cval = s->cval;
            return true;
        }
        case Expr::Member: {
            Expr& a = *e.args[0];
            if (!expr(a)) return false;
            if (a.type.n == 1) return fail(e.line, "member access on a scalar");
            std::vector<int> idx;
            for (char c : e.name) {
                int k = c == 'x' || c == 'r' ? 0 : c == 'y' || c == 'g' ? 1 : c == 'z' || c == 'b' ? 2
                      : c == 'w' || c == 'a' ? 3 : -1;
                if (k < 0 || k >= a.type.n) return fail(e.line, "bad swizzle ." + e.name);
                idx.push_back(k);
            }
            if (idx.empty() || idx.size() > 4) return fail(e.line, "bad swizzle ." + e.name);
            e.type = {a.type.base, (uint8_t)idx.size()};
            e.is_const = a.is_const;
            if (e.is_const)
                for (int k : idx) e.cval.push_back(a.cval[k]);
            return true;
        }
        case Expr::Index: {
            if (!expr(*e.args[0]) || !expr(*e.args[1])) return false;
            const Expr& a = *e.args[0];
            const Expr& i = *e.args[1];
            if (a.type.n == 1 || !i.type.is_int() || !i.is_const)
                return fail(e.line, "only constant component indices are supported");
            const int k = (int)i.cval[0];
            if (k < 0 || k >= a.type.n) return fail(e.line, "component index out of range");
            e.type = {a.type.base, 1};
            e.is_const = a.is_const;
            if (e.is_const) e.cval = {a.cval[k]};
            return true;
        }
        case Expr::Unary: {
            Expr& a = *e.args[0];
            if (!expr(a)) return false;
            if (e.op == "-") {
                if (!a.type.is_numeric()) return fail(e.line, "negation of " + tname(a.type));
                e.type = a.type;
                e.is_const = a.is_const;
                if (e.is_const)
                    for (double v : a.cval) e.cval.push_back(-v);
                return true;
            }
            if (e.op == "!") {
                if (a.type != kBool) return fail(e.line, "'!' needs a bool");
                e.type = kBool;
                e.is_const = a.is_const;
                if (e.is_const) e.cval = {a.cval[0] ? 0.0 : 1.0};
                return true;
            }
            return fail(e.line, "unsupported operator " + e.op);
        }
        case Expr::Binary: return binary(e);
        case Expr::Call: return call(e);
        }
        return false;
    }

    bool binary(Expr& e)
    {
        Expr& a = *e.args[0];
        Expr& b = *e.args[1];
        if (!expr(a) || !expr(b)) return false;
        const std::string& op = e.op;
        if (op == "&&" || op == "||") {
            if (a.type != kBool || b.type != kBool) return fail(e.line, op + " needs bools");
            e.type = kBool;
            e.is_const = a.is_const && b.is_const;
            if (e.is_const)
                e.cval = {op == "&&" ? (double)(a.cval[0] && b.cval[0]) : (double)(a.cval[0] || b.cval[0])};
            return true;
        }
        Type t;
        if (!unify(a.type, b.type, t, lang))
            return fail(e.line, "operands " + tname(a.type) + " and " + tname(b.type) + " of '" + op + "'");
        const bool cmp = op == "<" || op == ">" || op == "<=" || op == ">=" || op == "==" || op == "!=";
        if (cmp) {
            if (t.n != 1) return fail(e.line, "vector comparisons are not supported");
            e.type = kBool;
            e.cval.clear();
            e.is_const = false;
            if (a.is_const && b.is_const && t.base != Base::Bool) {
                const double x = convert_const(a.cval, a.type, t)[0], y = convert_const(b.cval, b.type, t)[0];
                e.is_const = true;
                e.cval = {(double)(op == "<" ? x < y : op == ">" ? x > y : op == "<=" ? x <= y
                                   : op == ">=" ? x >= y : op == "==" ? x == y : x != y)};
            }
            return true;
        }
        if (!t.is_numeric()) return fail(e.line, "arithmetic on " + tname(t));
        if (op == "%" && t.is_float() && lang == Lang::Glsl) return fail(e.line, "'%' on floats");
        if (op != "+" && op != "-" && op != "*" && op != "/" && op != "%")
            return fail(e.line, "unsupported operator " + op);
        e.type = t;
        e.is_const = a.is_const && b.is_const;
        if (e.is_const) {
            Type ts{t.base, 1};
            const auto x = convert_const(a.cval, a.type, {t.base, t.n});
            const auto y = convert_const(b.cval, b.type, {t.base, t.n});
            for (int k = 0; k < t.n; ++k) {
                bool ok;
                const double v = fold_arith(op, x[k], y[k], ts, ok);
                if (!ok) {
                    e.is_const = false;
                    e.cval.clear();
                    break;
                }
                e.cval.push_back(v);
            }
        }
        return true;
    }

    // builtin function signatures
    bool call(Expr& e)
    {
        for (auto& a : e.args)
            if (!expr(*a)) return false;
        const std::string& f = e.name;
        const size_t n = e.args.size();
        auto A = [&](size_t i) -> Expr& { return *e.args[i]; };
        auto all_const = [&] {
            for (auto& a : e.args)
                if (!a->is_const) return false;
            return true;
        };
        // ---- constructors and conversions ----
        {
            Type ct;
            bool ctor = false;
            if (f == "vec2" || f == "vec3" || f == "vec4") {
                ctor = true;
                Base b = Base::AbsInt;  // inferred from the arguments
                for (auto& a : e.args) {
                    Type u;
                    if (!unify({b, 1}, {a->type.base, 1}, u, lang))
                        return fail(e.line, "mixed component types in " + f);
                    b = u.base;
                }
                if (lang == Lang::Glsl || n == 0) b = Base::F32;
                ct = {b, (uint8_t)(f[3] - '0')};
            } else if (lang == Lang::Wgsl && f.size() == 5 && f.compare(0, 3, "vec") == 0) {
                ctor = true;
                ct = {f[4] == 'f' ? Base::F32 : f[4] == 'i' ? Base::I32 : Base::U32, (uint8_t)(f[3] - '0')};
            } else if (lang == Lang::Glsl && (f == "ivec2" || f == "ivec3" || f == "ivec4")) {
                ctor = true;
                ct = {Base::I32, (uint8_t)(f[4] - '0')};
            }
            if (ctor) return constructor(e, ct);
            Base sb;
            if (f == "f32" || f == "float") sb = Base::F32;
            else if (f == "i32" || f == "int") sb = Base::I32;
            else if (f == "u32" || f == "uint") sb = Base::U32;
            else if (f == "bool") sb = Base::Bool;
            else sb = Base::Void;
            if (sb != Base::Void) {
                if (n != 1 || A(0).type.n != 1) return fail(e.line, f + "() takes one scalar");
                e.type = {sb, 1};
                e.is_const = A(0).is_const;
                if (e.is_const) e.cval = convert_const(A(0).cval, A(0).type, e.type);
                return true;
            }
        }
        // ---- library functions (sdf3d_primitives.wgsl, sdf_op.wgsl, sdf3d_normal.wgsl) ----
        for (const LibFn& l : kLibrary)
            if (f == l.name && library_visible(lib, l) && !m.func_index.count(f)) {
                if (n != l.params.size()) return fail(e.line, f + ": wrong number of arguments");
                for (size_t i = 0; i < n; ++i)
                    if (!convertible(A(i).type, l.params[i], lang))
                        return fail(e.line, f + ": argument " + std::to_string(i + 1) + " is " +
                                                tname(A(i).type) + ", expected " + tname(l.params[i]));
                e.type = f == "sdf3d_normal" ? vec(Base::F32, 3) : kF32;
                return true;
            }
        // ---- user functions ----
        auto uf = m.func_index.find(f);
        if (uf != m.func_index.end()) {
            const Function& fn = m.funcs[uf->second];
            if (cur) calls[cur->name].insert(f);
            if (n != fn.params.size()) return fail(e.line, f + ": wrong number of arguments");
            for (size_t i = 0; i < n; ++i)
                if (!convertible(A(i).type, fn.params[i].type, lang))
                    return fail(e.line, f + ": argument " + std::to_string(i + 1) + " is " +
                                            tname(A(i).type) + ", expected " + tname(fn.params[i].type));
            e.type = fn.ret;
            return true;
        }
        // ---- WGSL / GLSL builtins ----
        auto num_args = [&](size_t k) {
            if (n != k) return fail(e.line, f + " takes " + std::to_string(k) + " arguments");
            return true;
        };
        auto same_float = [&](Type& t) {  // all arguments unify to one float type
            t = A(0).type;
            for (size_t i = 1; i < n; ++i)
                if (!unify(t, A(i).type, t, lang)) return fail(e.line, f + ": argument types");
            if (!t.is_float() && !(t.is_int() && (f == "abs" || f == "min" || f == "max" || f == "clamp")))
                return fail(e.line, f + " needs float arguments");
            return true;
        };
        static const std::set<std::string> unary_f = {
            "abs", "sqrt", "inverseSqrt", "inversesqrt", "sin", "cos", "tan", "asin", "acos",
            "exp", "exp2", "log", "log2", "floor", "ceil", "round", "trunc", "fract", "sign",
            "sinh", "cosh", "tanh", "radians", "degrees"};
        static const std::set<std::string> transcendental = {
            "sin", "cos", "tan", "asin", "acos", "atan", "atan2", "exp", "exp2", "log", "log2",
            "pow", "sinh", "cosh", "tanh", "inverseSqrt", "inversesqrt"};
        Type t;
        if (unary_f.count(f) || (f == "atan" && n == 1)) {
            if (!num_args(1) || !same_float(t)) return false;
            e.type = t;
        } else if (f == "min" || f == "max" || f == "pow" || f == "step" || f == "atan2" ||
                   (f == "atan" && n == 2) || f == "mod") {
            if (!num_args(2) || !same_float(t)) return false;
            e.type = t;
        } else if (f == "clamp" || f == "mix" || f == "smoothstep") {
            if (!num_args(3) || !same_float(t)) return false;
            e.type = t;
        } else if (f == "length") {
            if (!num_args(1) || !A(0).type.is_float()) return fail(e.line, "length needs a float vector");
            e.type = {A(0).type.base, 1};
        } else if (f == "distance" || f == "dot") {
            if (!num_args(2) || !same_float(t)) return false;
            e.type = {t.base, 1};
        } else if (f == "normalize") {
            if (!num_args(1) || !same_float(t)) return false;
            e.type = t;
        } else if (f == "cross") {
            if (!num_args(2) || !same_float(t) || t.n != 3) return fail(e.line, "cross needs vec3");
            e.type = t;
        } else if (f == "select") {
            if (!num_args(3) || A(2).type != kBool || !unify(A(0).type, A(1).type, t, lang))
                return fail(e.line, "select(f, t, bool)");
            e.type = t;
        } else {
            return fail(e.line, "unknown function '" + f + "'");
        }
        // fold abstract calls (in double) and non-transcendental concrete ones (in f32)
        if (all_const() && (e.type.is_abstract() || !transcendental.count(f)) && e.type.base != Base::Bool)
            fold_call(e);
        return true;
    }

    bool constructor(Expr& e, Type ct)
    {
        int comps = 0;
        for (auto& a : e.args) comps += a->type.n;
        if (!(e.args.empty() || (e.args.size() == 1 && e.args[0]->type.n == 1) || comps == ct.n))
            return fail(e.line, "constructor arguments do not add up to " + tname(ct));
        for (auto& a : e.args)
            if (!convertible({a->type.base, 1}, {ct.base, 1}, lang) && a->type.base != ct.base)
                return fail(e.line, "constructor argument " + tname(a->type) + " for " + tname(ct));
        e.type = ct;
        bool c = true;
        for (auto& a : e.args) c = c && a->is_const;
        e.is_const = c;
        if (c) {
            if (e.args.empty()) e.cval.assign(ct.n, 0.0);
            else if (e.args.size() == 1 && e.args[0]->type.n == 1)
                e.cval.assign(ct.n, convert_const(e.args[0]->cval, e.args[0]->type, {ct.base, 1})[0]);
            else
                for (auto& a : e.args) {
                    const auto v = convert_const(a->cval, a->type, {ct.base, a->type.n});
                    e.cval.insert(e.cval.end(), v.begin(), v.end());
                }
        }
        return true;
    }

    void fold_call(Expr& e)
    {
        const Type t = e.type;
        const bool f32 = t.base == Base::F32;
        auto arg = [&](size_t i, int k) {
            const Expr& a = *e.args[i];
            const auto v = convert_const(a.cval, a.type, {t.base, a.type.n});
            return v[a.type.n == 1 ? 0 : k];
        };
        auto R = [&](double v) { return f32 ? (double)(float)v : v; };
        const std::string& f = e.name;
        std::vector<double> out;
        bool ok = true;
        if (f == "length" || f == "dot" || f == "distance") {
            const Expr& a = *e.args[0];
            const int nn = a.type.n;
            double s = 0;
            for (int k = 0; k < nn; ++k) {
                double d = arg(0, k);
                double q = f == "dot" ? arg(1, k) : d;
                if (f == "distance") d = q = R(arg(0, k) - arg(1, k));
                s = k == 0 ? R(d * q) : R(s + R(d * q));
            }
            out = {f == "dot" ? s : (f32 ? (double)std::sqrt((float)s) : std::sqrt(s))};
        } else {
            for (int k = 0; k < t.n; ++k) {
                const double x = arg(0, k);
                double v;
                if (f == "abs") v = std::fabs(x);
                else if (f == "min") v = f32 ? (double)std::fmin((float)x, (float)arg(1, k)) : std::fmin(x, arg(1, k));
                else if (f == "max") v = f32 ? (double)std::fmax((float)x, (float)arg(1, k)) : std::fmax(x, arg(1, k));
                else if (f == "clamp") v = std::fmin(std::fmax(x, arg(1, k)), arg(2, k));
                else if (f == "sqrt") v = f32 ? (double)std::sqrt((float)x) : std::sqrt(x);
                else if (f == "floor") v = std::floor(x);
                else if (f == "ceil") v = std::ceil(x);
                else if (f == "trunc") v = std::trunc(x);
                else if (f == "sign") v = x > 0 ? 1 : x < 0 ? -1 : 0;
                else if (f == "sin") v = std::sin(x);
                else if (f == "cos") v = std::cos(x);
                else if (f == "tan") v = std::tan(x);
                else if (f == "exp") v = std::exp(x);
                else if (f == "log") v = std::log(x);
                else if (f == "pow") v = std::pow(x, arg(1, k));
                else if (f == "atan" && e.args.size() == 1) v = std::atan(x);
                else {
                    ok = false;
                    break;
                }
                out.push_back(R(v));
            }
        }
        if (ok) {
            e.is_const = true;
            e.cval = out;
        }
    }

    // ---- declarations / statements ----
    bool global_const(size_t gi)
    {
        GlobalConst& c = m.consts[gi];
        if (scopes.front().count(c.name)) return true;
        if (gconst_busy.count(c.name)) return fail(c.line, "constant '" + c.name + "' depends on itself");
        gconst_busy.insert(c.name);
        auto saved = scopes;
        scopes.resize(1);  // globals only
        bool ok = expr(*c.init);
        scopes = saved;
        gconst_busy.erase(c.name);
        if (!ok) return false;
        if (!c.init->is_const) return fail(c.line, "'" + c.name + "' is not a constant expression");
        Sym s;
        s.type = c.has_type ? c.type : c.init->type;
        if (c.has_type && !convertible(c.init->type, c.type, lang))
            return fail(c.line, "constant '" + c.name + "' has type " + tname(c.init->type));
        s.is_const = true;
        s.cval = convert_const(c.init->cval, c.init->type, s.type);
        c.type = s.type;
        scopes.front()[c.name] = s;
        return true;
    }

    bool assignable(Expr& lhs)
    {
        if (lhs.kind == Expr::Ident) {
            Sym* s = lookup(lhs.name);
            return s && s->mut;
        }
        if (lhs.kind == Expr::Member || lhs.kind == Expr::Index) return assignable(*lhs.args[0]);
        return false;
    }

    bool stmts(std::vector<StmtP>& body)
    {
        scopes.emplace_back();
        for (auto& s : body)
            if (!stmt(*s)) return false;
        scopes.pop_back();
        return true;
    }

    bool stmt(Stmt& s)
    {
        switch (s.kind) {
        case Stmt::Decl: {
            Sym sym;
            if (!s.has_type && !s.init) return fail(s.line, "'" + s.name + "' needs a type or an initializer");
            if (s.init && !expr(*s.init)) return false;
            if (s.has_type) {
                if (s.init && !convertible(s.init->type, s.decl_type, lang))
                    return fail(s.line, "'" + s.name + "' of type " + tname(s.decl_type) + " initialised with " +
                                            tname(s.init->type));
                sym.type = s.decl_type;
            } else {
                sym.type = s.is_const ? s.init->type : concrete(s.init->type);
                s.decl_type = sym.type;
            }
            sym.mut = s.is_var;
            if (!sym.mut && s.init && s.init->is_const) {
                sym.is_const = true;
                sym.cval = convert_const(s.init->cval, s.init->type, sym.type);
            }
            if (s.is_const && !(s.init && s.init->is_const))
                return fail(s.line, "'" + s.name + "' is not a constant expression");
            scopes.back()[s.name] = sym;
            return true;
        }
        case Stmt::Assign: {
            if (!expr(*s.lhs) || !expr(*s.init)) return false;
            if (!assignable(*s.lhs)) return fail(s.line, "assignment to an immutable value");
            if (s.op == "=") {
                if (!convertible(s.init->type, s.lhs->type, lang))
                    return fail(s.line, "cannot assign " + tname(s.init->type) + " to " + tname(s.lhs->type));
            } else {
                Type t;
                if (!unify(s.lhs->type, s.init->type, t, lang) || t.n != s.lhs->type.n ||
                    !convertible(t, s.lhs->type, lang))
                    return fail(s.line, "compound assignment of " + tname(s.init->type) + " to " +
                                            tname(s.lhs->type));
            }
            return true;
        }
        case Stmt::Return:
            if (!cur) return false;
            if (!s.init) {
                if (cur->ret.base != Base::Void) return fail(s.line, "return needs a value");
                return true;
            }
            if (!expr(*s.init)) return false;
            if (!convertible(s.init->type, cur->ret, lang))
                return fail(s.line, "returns " + tname(s.init->type) + ", expected " + tname(cur->ret));
            return true;
        case Stmt::If:
            if (!expr(*s.init)) return false;
            if (s.init->type != kBool) return fail(s.line, "if condition must be bool");
            return stmts(s.body) && stmts(s.else_body);
        case Stmt::For: {
            scopes.emplace_back();
            bool ok = (!s.for_init || stmt(*s.for_init)) && (!s.init || expr(*s.init)) &&
                      (!s.init || s.init->type == kBool || fail(s.line, "loop condition must be bool")) &&
                      (!s.for_update || stmt(*s.for_update)) && stmts(s.body);
            scopes.pop_back();
            return ok;
        }
        case Stmt::While:
            if (!expr(*s.init)) return false;
            if (s.init->type != kBool) return fail(s.line, "loop condition must be bool");
            return stmts(s.body);
        case Stmt::Break:
        case Stmt::Continue: return true;
        case Stmt::Block: return stmts(s.body);
        case Stmt::ExprS: return expr(*s.init);
        }
        return false;
    }

    bool function(const Function& f)
    {
        cur = &f;
        scopes.resize(1);
        scopes.emplace_back();
        for (const Param& p : f.params) {
            Sym s;
            s.type = p.type;
            s.mut = lang == Lang::Glsl;  // GLSL parameters are local copies
            scopes.back()[p.name] = s;
        }
        bool ok = stmts(const_cast<std::vector<StmtP>&>(f.body));
        scopes.resize(1);
        cur = nullptr;
        return ok;
    }
};

}  // namespace

bool check(Module& m, Lang lang, const std::string& entry, const Visible& lib, Error& err)
{
    Checker c{m, lang, lib};
    for (const LibFn& l : kLibrary)
        if (library_visible(lib, l) && m.func_index.count(l.name)) {
            err = {S2M_E_VALIDATION, std::string("redefinition of library function ") + l.name};
            return false;
        }
    c.scopes.emplace_back();
    for (size_t i = 0; i < m.consts.size(); ++i) c.gconst_index[m.consts[i].name] = i;
    for (size_t i = 0; i < m.consts.size(); ++i)
        if (!c.global_const(i)) {
            err = c.err;
            return false;
        }
    for (const Function& f : m.funcs)
        if (!c.function(f)) {
            err = c.err;
            return false;
        }
    // WGSL and GLSL forbid recursion; a cycle in the call graph is an error
    std::map<std::string, int> state;  // 1 on the DFS stack, 2 done
    std::function<bool(const std::string&)> dfs = [&](const std::string& f) {
        state[f] = 1;
        for (const std::string& g : c.calls[f]) {
            if (state[g] == 1) return false;
            if (state[g] == 0 && !dfs(g)) return false;
        }
        state[f] = 2;
        return true;
    };
    for (const Function& f : m.funcs)
        if (state[f.name] == 0 && !dfs(f.name)) {
            err = {S2M_E_VALIDATION, "recursive call involving '" + f.name + "'"};
            return false;
        }
    auto it = m.func_index.find(entry);
    if (it == m.func_index.end()) {
        err = {S2M_E_MISSING_SDF, "no function '" + entry + "'"};
        return false;
    }
    const Function& f = m.funcs[it->second];
    if (f.params.size() != 1 || f.params[0].type != Type{Base::F32, 3} || f.ret != Type{Base::F32, 1}) {
        err = {S2M_E_VALIDATION, "'" + entry + "' must take one vec3<f32> and return f32"};
        return false;
    }
    return true;
}

}  // namespace s2m::fe
