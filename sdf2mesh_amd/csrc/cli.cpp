// cli.cpp - the `sdf2mesh` command line tool over the C ABI (main.rs:95-137, 178-372).
//
//   sdf2mesh (-i|--sdf FILE.sdf3d | --glsl FILE [--glsl-sdf NAME] | --shadertoy-id ID
//            [--shadertoy-sdf NAME]) -0|--mesh OUT.stl|OUT.ply [-r|--resolution N]
//            [-b|--bounds B] [--debug-wgsl FILE] [--debug-png DIR] [--device N] [--timing]
//
// Resolution defaults to 256 and is rounded up to a power of two with a warning, the box
// is the cube of edge B (default 2) around the origin, eps 1e-4 (main.rs:139-175).  The
// mesh is meshed with the reference's per-cell corner coordinates and one-slice lag, so
// the file equals the reference's output for the same program.  Log lines follow the
// reference's (env_logger at info level).  Exit status: 0, 1 for a failed step, 2 for a
// usage error.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sdf2mesh.h"

namespace {

void log_line(const char* level, const char* fmt, ...)
{
    const auto now = std::chrono::system_clock::now().time_since_epoch();
    std::fprintf(stderr, "[%.9f %s sdf2mesh] ", std::chrono::duration<double>(now).count(), level);
    va_list ap;
    va_start(ap, fmt);
    std::vfprintf(stderr, fmt, ap);
    va_end(ap);
    std::fputc('\n', stderr);
}

int usage(const char* msg)
{
    if (msg) std::fprintf(stderr, "error: %s\n\n", msg);
    std::fprintf(stderr,
                 "Usage: sdf2mesh [OPTIONS] --mesh <MESH>\n\n"
                 "Options:\n"
                 "  -i, --sdf <SDF>                      Input SDF file\n"
                 "      --shadertoy-id <SHADERTOY_ID>    Input ShaderToy shader ID\n"
                 "      --shadertoy-sdf <SHADERTOY_SDF>  ShaderToy SDF name [default: sdf]\n"
                 "      --glsl <GLSL>                    Input GLSL fragment shader\n"
                 "      --glsl-sdf <GLSL_SDF>            GLSL SDF name [default: sdf]\n"
                 "  -0, --mesh <MESH>                    Output mesh file (supports STL and PLY output)\n"
                 "      --debug-wgsl <DEBUG_WGSL>        Output the shader source for debugging\n"
                 "      --debug-png <DEBUG_PNG>          Write PNG images for debugging\n"
                 "  -r, --resolution <RESOLUTION>        Grid resolution. Default: 256\n"
                 "  -b, --bounds <BOUNDS>                Size of bounding box. Default: 2\n"
                 "      --device <DEVICE>                HIP device [default: 0]\n"
                 "      --timing                         Print device timings\n"
                 "  -h, --help                           Print help\n");
    return 2;
}

// ---- --debug-png (main.rs:346-349, texture.rs:120-129, png.rs:6-23) -----------------
// PNG with stored (uncompressed) deflate blocks: the pixels equal the reference's, the
// compressed bytes do not (it uses the png crate's default compression).
uint32_t crc32(const uint8_t* p, size_t n, uint32_t c = 0xffffffffu)
{
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t v = i;
            for (int k = 0; k < 8; ++k) v = (v & 1) ? 0xedb88320u ^ (v >> 1) : v >> 1;
            table[i] = v;
        }
        init = true;
    }
    for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xff] ^ (c >> 8);
    return c;
}

void put32(std::vector<uint8_t>& o, uint32_t v)
{
    for (int k = 3; k >= 0; --k) o.push_back((uint8_t)(v >> (8 * k)));
}

void chunk(std::vector<uint8_t>& o, const char* type, const std::vector<uint8_t>& data)
{
    put32(o, (uint32_t)data.size());
    const size_t at = o.size();
    o.insert(o.end(), type, type + 4);
    o.insert(o.end(), data.begin(), data.end());
    put32(o, crc32(o.data() + at, o.size() - at) ^ 0xffffffffu);
}

bool write_png(const std::string& path, uint32_t w, uint32_t h, const std::vector<uint8_t>& rgba)
{
    std::vector<uint8_t> raw;  // filter byte 0 + row
    raw.reserve((size_t)h * (4 * w + 1));
    for (uint32_t y = 0; y < h; ++y) {
        raw.push_back(0);
        raw.insert(raw.end(), rgba.begin() + (size_t)y * w * 4, rgba.begin() + (size_t)(y + 1) * w * 4);
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    uint32_t a = 1, b = 0;  // Adler-32
    for (uint8_t c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    for (size_t off = 0; off < raw.size() || off == 0; off += 65535) {
        const size_t n = std::min<size_t>(65535, raw.size() - off);
        z.push_back(off + n >= raw.size() ? 1 : 0);
        z.push_back((uint8_t)(n & 0xff));
        z.push_back((uint8_t)(n >> 8));
        z.push_back((uint8_t)(~n & 0xff));
        z.push_back((uint8_t)((~n >> 8) & 0xff));
        z.insert(z.end(), raw.begin() + off, raw.begin() + off + n);
        if (n == 0) break;
    }
    put32(z, (b << 16) | a);
    std::vector<uint8_t> o = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<uint8_t> ihdr;
    put32(ihdr, w);
    put32(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8-bit RGBA
    chunk(o, "IHDR", ihdr);
    chunk(o, "IDAT", z);
    chunk(o, "IEND", {});
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(o.data(), 1, o.size(), f) == o.size();
    return std::fclose(f) == 0 && ok;
}

// (f * 127 + 128).clamp(0, 255) as u8
uint8_t to_u8(float f)
{
    const float v = f * 127.0f + 128.0f;
    const float c = std::fmin(std::fmax(v, 0.0f), 255.0f);
    return (uint8_t)c;
}

// The images the reference writes after slice z's readback: the readback lags one slice
// (the buffers are mapped before the slice's work is submitted), so they show slice
// z - 1's textures, and zeros for z = 0.  Texel (x, y) of an active cell: normal texture
// (n, sign bits as a float), position texture (p, 1); inactive cells hold zeros.
bool write_debug_pngs(const std::string& prefix, uint32_t res, const s2m_mesh_out& m)
{
    uint64_t v = 0;
    std::vector<uint8_t> nimg((size_t)res * res * 4), pimg((size_t)res * res * 4);
    for (uint32_t z = 0; z < res; ++z) {
        std::fill(nimg.begin(), nimg.end(), to_u8(0.0f));
        std::fill(pimg.begin(), pimg.end(), to_u8(0.0f));
        if (z > 0) {
            for (; v < m.nvtx && (m.cell_keys[v] >> 32) == z - 1; ++v) {
                const uint32_t x = (uint32_t)(m.cell_keys[v] & 0xffff), y = (uint32_t)((m.cell_keys[v] >> 16) & 0xffff);
                const float* q = m.vtx + 6 * v;
                uint8_t* pn = &nimg[((size_t)y * res + x) * 4];
                uint8_t* pp = &pimg[((size_t)y * res + x) * 4];
                for (int k = 0; k < 3; ++k) {
                    pn[k] = to_u8(q[3 + k]);
                    pp[k] = to_u8(q[k]);
                }
                pn[3] = to_u8((float)m.signs[v]);
                pp[3] = to_u8(1.0f);
            }
        }
        char name[32];
        std::snprintf(name, sizeof name, "%04u", z);
        if (!write_png(prefix + name + "_normal.png", res, res, nimg) ||
            !write_png(prefix + name + "_position.png", res, res, pimg))
            return false;
    }
    return true;
}

int fail_step(const char* what, int rc)
{
    log_line("ERROR", "%s failed (status %d): %s", what, rc, s2m_last_error());
    return 1;
}

}  // namespace

int main(int argc, char** argv)
{
    std::string sdf, glsl, glsl_sdf = "sdf", shadertoy_id, shadertoy_sdf = "sdf", mesh, debug_wgsl,
                                  debug_png;
    uint32_t resolution = 256;
    float bounds = 2.0f;
    int device = 0;
    bool timing = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto value = [&](std::string& out) {
            if (i + 1 >= argc) return false;
            out = argv[++i];
            return true;
        };
        std::string v;
        if (a == "-h" || a == "--help") {
            usage(nullptr);
            return 0;
        } else if (a == "-i" || a == "--sdf") {
            if (!value(sdf)) return usage("--sdf needs a value");
        } else if (a == "--glsl") {
            if (!value(glsl)) return usage("--glsl needs a value");
        } else if (a == "--glsl-sdf") {
            if (!value(glsl_sdf)) return usage("--glsl-sdf needs a value");
        } else if (a == "--shadertoy-id") {
            if (!value(shadertoy_id)) return usage("--shadertoy-id needs a value");
        } else if (a == "--shadertoy-sdf") {
            if (!value(shadertoy_sdf)) return usage("--shadertoy-sdf needs a value");
        } else if (a == "-0" || a == "--mesh") {
            if (!value(mesh)) return usage("--mesh needs a value");
        } else if (a == "--debug-wgsl") {
            if (!value(debug_wgsl)) return usage("--debug-wgsl needs a value");
        } else if (a == "--debug-png") {
            if (!value(debug_png)) return usage("--debug-png needs a value");
        } else if (a == "-r" || a == "--resolution") {
            if (!value(v)) return usage("--resolution needs a value");
            char* end = nullptr;
            const unsigned long r = std::strtoul(v.c_str(), &end, 10);
            if (!*v.c_str() || *end || r == 0 || r > 65535) return usage("invalid --resolution");
            resolution = (uint32_t)r;
        } else if (a == "-b" || a == "--bounds") {
            if (!value(v)) return usage("--bounds needs a value");
            char* end = nullptr;
            bounds = std::strtof(v.c_str(), &end);
            if (!*v.c_str() || *end) return usage("invalid --bounds");
        } else if (a == "--device") {
            if (!value(v)) return usage("--device needs a value");
            device = std::atoi(v.c_str());
        } else if (a == "--timing") {
            timing = true;
        } else {
            return usage(("unexpected argument '" + a + "'").c_str());
        }
    }
    if (mesh.empty()) return usage("the following required arguments were not provided: --mesh <MESH>");

    // main.rs:141-148: a resolution that is not a power of two is rounded up, with a warning
    const uint32_t rounded = s2m_round_resolution(resolution);
    if (rounded != resolution)
        log_line("WARN", "Resolution should be a power of 2 (actual resolution : %u)", rounded);

    // the SDF source (main.rs:198-221): ShaderToy first, then --sdf, then --glsl
    std::string source;
    uint32_t lang = S2M_LANG_SDF3D | S2M_SOURCE_IS_PATH;
    const char* sdf_name = nullptr;
    if (!shadertoy_id.empty()) {
        log_line("INFO", "Reading SDF from ShaderToy (shader ID %s)", shadertoy_id.c_str());
        log_line("ERROR", "ShaderToy requests need network access, which this tool does not use "
                          "(RequestError)");
        return 1;
    } else if (!sdf.empty()) {
        log_line("INFO", "Reading SDF from %s...", sdf.c_str());
        source = sdf;
    } else if (!glsl.empty()) {
        log_line("INFO", "Reading SDF from GLSL fragment shader %s...", glsl.c_str());
        source = glsl;
        lang = S2M_LANG_GLSL | S2M_SOURCE_IS_PATH;
        sdf_name = glsl_sdf.c_str();
    } else {
        log_line("ERROR", "no SDF given (--sdf, --glsl or --shadertoy-id)");
        return 1;
    }

    if (!debug_wgsl.empty()) {  // main.rs:223-225: written before validation
        int st = 0;
        const uint64_t n = s2m_translate(source.c_str(), lang, sdf_name, S2M_STAGE_SOURCE, nullptr, 0, &st);
        if (st) return fail_step("reading the SDF source", st);
        std::vector<char> text(n);
        s2m_translate(source.c_str(), lang, sdf_name, S2M_STAGE_SOURCE, text.data(), n, &st);
        FILE* f = std::fopen(debug_wgsl.c_str(), "wb");
        if (!f || std::fwrite(text.data(), 1, n, f) != n) {
            if (f) std::fclose(f);
            log_line("ERROR", "could not write %s", debug_wgsl.c_str());
            return 1;
        }
        std::fclose(f);
    }

    s2m_ctx* ctx = s2m_create(device);
    if (!ctx) return fail_step("device set-up", S2M_E_HIP);
    s2m_program* prog = nullptr;
    int rc = s2m_program_from_source(ctx, source.c_str(), lang, sdf_name, &prog);
    if (rc) {
        s2m_destroy(ctx);
        return fail_step("building the SDF program", rc);
    }
    s2m_params params;
    s2m_params_reference(&params, resolution, bounds);
    if (timing) s2m_set_profiling(ctx, 1);
    s2m_mesh_out out;
    rc = s2m_mesh(ctx, prog, &params, debug_png.empty() ? 0 : 1, &out);
    if (rc) {
        s2m_program_free(prog);
        s2m_destroy(ctx);
        return fail_step("meshing", rc);
    }
    log_line("INFO", "Mesh has %llu vertices.", (unsigned long long)out.nvtx);
    if (timing) {
        s2m_timing t;
        s2m_last_timing(ctx, &t);
        log_line("INFO", "device %.3f ms (prune %.3f, eval %.3f, contour %.3f, vertices+normals %.3f, "
                         "quads %.3f), copy-out %.3f ms",
                 t.total_ms, t.prune_ms, t.eval_ms, t.contour_ms, t.normal_ms, t.quad_ms, t.d2h_ms);
    }
    if (!debug_png.empty() && !write_debug_pngs(debug_png, params.res[0], out))
        log_line("ERROR", "could not write the debug PNG files %s*", debug_png.c_str());
    rc = s2m_write_to_file(&out, mesh.c_str());
    if (rc) log_line("ERROR", "Could not write mesh to %s!", s2m_last_error());
    else log_line("INFO", "Mesh written to %s", mesh.c_str());
    s2m_mesh_free(&out);
    s2m_program_free(prog);
    s2m_destroy(ctx);
    return rc ? 1 : 0;
}
