// compile.cpp - SDF source programs: the C ABI over the front end, and hiprtc.
//
// The reference turns an .sdf3d file (or a GLSL fragment shader) into a WGSL module that
// wgpu compiles for the GPU (shader.rs:44-104, main.rs:198-227).  Here the front end
// either lowers the program to a tape of library primitives and operators or generates
// HIP C++, which hiprtc compiles for gfx950 with the same floating-point options as the
// library (no contraction, correctly rounded division and square root, denormals kept,
// no packed fp32).  Code objects are cached per process by their source text.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstring>
#include <mutex>
#include <unordered_map>

#include "frontend.h"
#include "internal.h"

namespace {

// rtc_prelude.h and the device headers it needs, as one translation-unit prefix
const char kRtcPrefix[] =
#include "rtc_source.inc"
    ;

std::mutex g_cache_mu;
std::unordered_map<std::string, std::vector<char>> g_cache;

s2m::fe::Lang lang_of(uint32_t lang_flags)
{
    return (lang_flags & 0xffu) == S2M_LANG_GLSL ? s2m::fe::Lang::Glsl : s2m::fe::Lang::Wgsl;
}

}  // namespace

namespace s2m::fe {

bool compile_code_object(const std::string& hip, std::vector<char>& code, Error& err)
{
    {
        std::lock_guard<std::mutex> g(g_cache_mu);
        auto it = g_cache.find(hip);
        if (it != g_cache.end()) {
            code = it->second;
            return true;
        }
    }
    const std::string src = std::string(kRtcPrefix) + "\n" + hip;
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "sdf_program.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        err = {S2M_E_HIP, "hiprtcCreateProgram failed"};
        return false;
    }
    const char* opts[] = {"--offload-arch=gfx950",
                          "-O3",
                          "-std=c++17",
                          "-ffp-contract=off",
                          "-fhip-fp32-correctly-rounded-divide-sqrt",
                          "-fno-gpu-flush-denormals-to-zero",
                          "-Xclang",
                          "-target-feature",
                          "-Xclang",
                          "-packed-fp32-ops"};
    const hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
    if (r != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string log(n + 1, '\0');
        hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        err = {S2M_E_HIP, std::string("hiprtc: ") + hiprtcGetErrorString(r) + "\n" + log.c_str()};
        return false;
    }
    size_t n = 0;
    hiprtcGetCodeSize(prog, &n);
    code.resize(n);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    std::lock_guard<std::mutex> g(g_cache_mu);
    g_cache[hip] = code;
    return true;
}

}  // namespace s2m::fe

extern "C" {

int s2m_program_from_source(s2m_ctx* ctx, const char* source, uint32_t lang_flags, const char* sdf_name,
                            s2m_program** out)
{
    if (!ctx || !source || !out) return s2m::set_error(S2M_E_ARG, "NULL argument");
    *out = nullptr;
    s2m::fe::Translation t;
    s2m::fe::Error err{0, ""};
    const bool force = (lang_flags & S2M_FORCE_COMPILED) != 0;
    if (!s2m::fe::translate(source, (lang_flags & S2M_SOURCE_IS_PATH) != 0, lang_of(lang_flags),
                            sdf_name ? sdf_name : "", force, t, err))
        return s2m::set_error(err.code, err.msg.c_str());
    if (!t.tape.empty() && !force) return s2m_program_from_tape(ctx, t.tape.data(), (uint32_t)t.tape.size(), out);
    std::vector<char> code;
    if (!s2m::fe::compile_code_object(t.hip, code, err)) return s2m::set_error(err.code, err.msg.c_str());
    return s2m::program_from_code_object(ctx, code.data(), code.size(), out);
}

uint64_t s2m_translate(const char* source, uint32_t lang_flags, const char* sdf_name, uint32_t stage, void* buf,
                       uint64_t cap, int* status)
{
    auto finish = [&](int code, const std::string& msg) -> uint64_t {
        if (status) *status = code;
        if (code) s2m::set_error(code, msg.c_str());
        return 0;
    };
    if (!source || stage > S2M_STAGE_CODE) return finish(S2M_E_ARG, "bad argument");
    s2m::fe::Translation t;
    s2m::fe::Error err{0, ""};
    const bool is_path = (lang_flags & S2M_SOURCE_IS_PATH) != 0;
    if (stage == S2M_STAGE_SOURCE) {  // written before validation, as --debug-wgsl is
        std::vector<std::string> uses;
        if (!s2m::fe::load_source(source, is_path, lang_of(lang_flags), t.source, uses, err))
            return finish(err.code, err.msg);
    } else if (!s2m::fe::translate(source, is_path, lang_of(lang_flags), sdf_name ? sdf_name : "",
                                   stage >= S2M_STAGE_HIP, t, err)) {
        return finish(err.code, err.msg);
    }
    std::vector<char> code;
    const void* p = nullptr;
    uint64_t n = 0;
    switch (stage) {
    case S2M_STAGE_SOURCE:
        p = t.source.data();
        n = t.source.size();
        break;
    case S2M_STAGE_TAPE:
        if (t.tape.empty()) return finish(S2M_E_WGSL, "not a tape program: " + t.tape_why);
        p = t.tape.data();
        n = t.tape.size() * sizeof(s2m_op);
        break;
    case S2M_STAGE_HIP:
        p = t.hip.data();
        n = t.hip.size();
        break;
    default:
        if (!s2m::fe::compile_code_object(t.hip, code, err)) return finish(err.code, err.msg);
        p = code.data();
        n = code.size();
    }
    if (buf && cap) std::memcpy(buf, p, (size_t)std::min<uint64_t>(n, cap));
    if (status) *status = S2M_OK;
    return n;
}

}  // extern "C"
