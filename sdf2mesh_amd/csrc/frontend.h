// frontend.h - SDF source front end: `.sdf3d` WGSL and GLSL fragment shaders.
//
// Source text -> preprocessing (`use` / `include`, shader.rs:159-203) -> one parser for
// the WGSL subset the examples and library use and the GLSL subset of mandelmesh-class
// shaders -> typed AST -> either
//   * a tape (when the program reduces to library primitives at p +- const combined by
//     library operators with constant parameters), or
//   * HIP C++ for hiprtc: `float s2m_user_sdf(float, float, float)`, every other program.
#pragma once
#include <stdint.h>

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/sdf2mesh.h"

namespace s2m::fe {

// ---- types ----------------------------------------------------------------------------
enum class Base : uint8_t { Void, Bool, I32, U32, F32, AbsInt, AbsFloat };
struct Type {
    Base base = Base::Void;
    uint8_t n = 1;  // vector width 1..4
    bool operator==(const Type& o) const { return base == o.base && n == o.n; }
    bool operator!=(const Type& o) const { return !(*this == o); }
    bool is_float() const { return base == Base::F32 || base == Base::AbsFloat; }
    bool is_int() const { return base == Base::I32 || base == Base::U32 || base == Base::AbsInt; }
    bool is_abstract() const { return base == Base::AbsInt || base == Base::AbsFloat; }
    bool is_numeric() const { return is_float() || is_int(); }
};

// ---- AST ------------------------------------------------------------------------------
struct Expr;
using ExprP = std::shared_ptr<Expr>;
struct Expr {
    enum Kind { Lit, Ident, Call, Member, Index, Unary, Binary } kind;
    // Lit
    double num = 0;
    bool is_bool = false, is_float_lit = false;
    char suffix = 0;  // 'f', 'u', 'i' or 0
    // Ident / Call / Member
    std::string name;  // identifier, callee, or member / swizzle
    std::vector<ExprP> args;  // call arguments; Member/Index/Unary: args[0]; Binary: args[0..1]
    std::string op;           // Unary / Binary operator
    int line = 0;
    // filled by the checker
    Type type;
    bool is_const = false;  // constant-foldable
    std::vector<double> cval;  // folded value (n components)
};

struct Stmt;
using StmtP = std::shared_ptr<Stmt>;
struct Stmt {
    enum Kind { Decl, Assign, Return, If, For, While, Break, Continue, Block, ExprS } kind;
    // Decl: name, decl_type (explicit or Void), init (may be null), mut (var) / const
    std::string name;
    Type decl_type;
    bool has_type = false, is_var = false, is_const = false;
    ExprP init;            // Decl init, Return value, If/For/While condition, ExprS
    ExprP lhs;             // Assign target
    std::string op;        // Assign: "=", "+=", ...
    std::vector<StmtP> body, else_body;  // If / For / While / Block
    StmtP for_init, for_update;
    int line = 0;
};

struct Param {
    std::string name;
    Type type;
};

struct Function {
    std::string name;
    std::vector<Param> params;
    Type ret;
    std::vector<StmtP> body;
    int line = 0;
};

struct GlobalConst {
    std::string name;
    Type type;
    bool has_type = false;
    ExprP init;
    int line = 0;
};

struct Module {
    std::vector<GlobalConst> consts;
    std::vector<Function> funcs;
    std::map<std::string, size_t> func_index;
};

enum class Lang { Wgsl, Glsl };

// Front-end error with an s2m_status code.
struct Error {
    int code;
    std::string msg;
};

// Expand `use X;` and `include "file";` lines of an .sdf3d file (shader.rs:159-203): the
// library modules are builtins, so a `use` line only records which ones are visible.
bool preprocess_file(const std::string& path, std::string& out, std::vector<std::string>& uses,
                     Error& err);
void preprocess_text(const std::string& src, std::string& out, std::vector<std::string>& uses);

// Library modules a source can see: `use sdf3d::*` brings the primitives and the normal,
// `use sdf::op` / `use sdf::*` the smooth operators (shader.rs:49-61); a GLSL shader
// gets the normal only (shader.rs:86).
struct Visible {
    bool primitives = false, ops = false, normal = false;
};
Visible visible_modules(const std::vector<std::string>& uses);

bool parse(const std::string& src, Lang lang, Module& m, Error& err);
bool check(Module& m, Lang lang, const std::string& entry, const Visible& lib, Error& err);

// Type rules shared by the checker and the code generator.
Type concrete(Type t);
bool convertible(const Type& from, const Type& to, Lang lang);
bool unify(const Type& a, const Type& b, Type& out, Lang lang);
double fold_arith(const std::string& op, double a, double b, const Type& t, bool& ok);
std::vector<double> convert_const(const std::vector<double>& c, const Type& from, const Type& to);

// Lowerings of a checked module; `entry` is the function of one vec3<f32> returning f32.
// lower_to_tape fails (returning false, with the reason in `why`) for programs that are
// not library primitives and operators with constant operands.
bool lower_to_tape(const Module& m, Lang lang, const std::string& entry, std::vector<s2m_op>& tape,
                   std::string& why);
bool emit_hip(const Module& m, Lang lang, const std::string& entry, std::string& hip, Error& err);

// Source text after `use` / `include` expansion (WGSL) or as read (GLSL).
bool load_source(const std::string& text, bool is_path, Lang lang, std::string& source,
                 std::vector<std::string>& uses, Error& err);

// The whole front end: source text (or a file path) -> checked module -> tape and/or HIP.
struct Translation {
    std::string source;         // after `use` / `include` expansion
    std::vector<s2m_op> tape;   // empty when the program does not lower to a tape
    std::string tape_why;       // why it did not
    std::string hip;            // generated HIP C++ (when requested or needed)
};
bool translate(const std::string& text, bool is_path, Lang lang, const std::string& sdf_name,
               bool want_hip, Translation& out, Error& err);

// hiprtc: HIP C++ -> gfx950 code object (no device needed).
bool compile_code_object(const std::string& hip, std::vector<char>& code, Error& err);

}  // namespace s2m::fe
