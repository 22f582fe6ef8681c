// writer.cpp - STL / PLY output with the reference's exact text (mesh.rs:8-210).
//
// Floats are printed the way Rust's `{}` Display prints an f32: the shortest decimal
// digit string that reads back as the same f32, laid out in fixed notation (no
// exponent), "-0" for negative zero, "NaN" / "inf" / "-inf".  Triangles are formatted
// in parallel chunks and written in order.
#include <algorithm>
#include <cctype>
#include <charconv>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sdf2mesh.h"

namespace {

// Shortest round-trip digits via std::to_chars (scientific), then the fixed layout of
// Rust's flt2dec::digits_to_dec_str with no minimum fraction digits.
int format_f32(float x, char* out)
{
    char* o = out;
    if (x != x) {
        std::memcpy(out, "NaN", 4);
        return 3;
    }
    if (std::signbit(x)) *o++ = '-';
    const float ax = std::fabs(x);
    if (ax == INFINITY) {
        std::memcpy(o, "inf", 4);
        return (int)(o - out) + 3;
    }
    if (ax == 0.0f) {
        *o++ = '0';
        *o = 0;
        return (int)(o - out);
    }
    char sci[48];
    auto r = std::to_chars(sci, sci + sizeof sci, ax, std::chars_format::scientific);
    *r.ptr = 0;
    // sci = d[.ddd]e[+-]XX
    char digits[16];
    int nd = 0;
    const char* s = sci;
    for (; *s && *s != 'e'; ++s)
        if (*s != '.') digits[nd++] = *s;
    int e10 = std::atoi(s + 1);
    while (nd > 1 && digits[nd - 1] == '0') --nd;
    // to_chars breaks an exact tie between the two nearest nd-digit strings towards an
    // even last digit; Rust's shortest mode (flt2dec Dragon) rounds such a tie up.  A
    // tie means the exact value has nd+1 significant digits, the last being 5.
    {
        char ex[160];
        auto rx = std::to_chars(ex, ex + sizeof ex, ax, std::chars_format::scientific, 120);
        *rx.ptr = 0;
        char X[130];
        int nx = 0;
        const char* q = ex;
        for (; *q && *q != 'e'; ++q)
            if (*q != '.') X[nx++] = *q;
        const int ex10 = std::atoi(q + 1);
        while (nx > 1 && X[nx - 1] == '0') --nx;
        if (ex10 == e10 && nx == nd + 1 && X[nd] == '5' && std::memcmp(X, digits, nd) == 0) {
            char up[16];
            std::memcpy(up, digits, nd);
            int i = nd - 1;
            while (i >= 0 && up[i] == '9') up[i--] = '0';
            int nu = nd, eu = e10;
            if (i >= 0) {
                ++up[i];
            } else {  // 99..9 -> 100..0
                std::memmove(up + 1, up, nd);
                up[0] = '1';
                ++nu;
                ++eu;
            }
            char chk[48];
            int n = std::snprintf(chk, sizeof chk, "%.*se%d", nu, up, eu - (nu - 1));
            (void)n;
            if (std::strtof(chk, nullptr) == ax) {
                std::memcpy(digits, up, nu);
                nd = nu;
                e10 = eu;
                while (nd > 1 && digits[nd - 1] == '0') --nd;
            }
        }
    }
    const int exp = e10 + 1;  // value = 0.d1d2..dn * 10^exp
    if (exp <= 0) {
        *o++ = '0';
        *o++ = '.';
        for (int i = 0; i < -exp; ++i) *o++ = '0';
        std::memcpy(o, digits, nd);
        o += nd;
    } else if (exp < nd) {
        std::memcpy(o, digits, exp);
        o += exp;
        *o++ = '.';
        std::memcpy(o, digits + exp, nd - exp);
        o += nd - exp;
    } else {
        std::memcpy(o, digits, nd);
        o += nd;
        for (int i = nd; i < exp; ++i) *o++ = '0';
    }
    *o = 0;
    return (int)(o - out);
}

inline void put(std::string& s, const char* lit) { s.append(lit); }
inline void putf(std::string& s, float x)
{
    char b[64];
    s.append(b, (size_t)format_f32(x, b));
}

// STLWriter::write_triangle (mesh.rs:19-41) for triangle t.
void stl_triangle(std::string& s, const float* vtx, const uint32_t* tri, uint64_t t)
{
    const float* v0 = vtx + 6 * (uint64_t)tri[3 * t];
    const float* v1 = vtx + 6 * (uint64_t)tri[3 * t + 1];
    const float* v2 = vtx + 6 * (uint64_t)tri[3 * t + 2];
    // Triangle::normal (lib.rs:182-184): (v2 - v0).cross(v1 - v0)
    const float ax = v2[0] - v0[0], ay = v2[1] - v0[1], az = v2[2] - v0[2];
    const float bx = v1[0] - v0[0], by = v1[1] - v0[1], bz = v1[2] - v0[2];
    const float n[3] = {ay * bz - az * by, az * bx - ax * bz, ax * by - ay * bx};
    put(s, "facet normal ");
    putf(s, n[0]);
    put(s, " ");
    putf(s, n[1]);
    put(s, " ");
    putf(s, n[2]);
    put(s, "\n\touter loop\n");
    for (const float* v : {v0, v1, v2}) {
        put(s, "\t\tvertex ");
        putf(s, v[0]);
        put(s, " ");
        putf(s, v[1]);
        put(s, " ");
        putf(s, v[2]);
        put(s, "\n");
    }
    put(s, "\tendloop\nendfacet\n");
}

// Format items [0, n) with fn(string&, i) on worker threads and write them in order.
template <class F>
bool write_parallel(FILE* f, uint64_t n, F fn)
{
    const uint64_t chunk = 16384;
    unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    uint64_t nchunks = (n + chunk - 1) / chunk;
    for (uint64_t c0 = 0; c0 < nchunks; c0 += hw) {
        const uint64_t c1 = std::min<uint64_t>(nchunks, c0 + hw);
        std::vector<std::string> out(c1 - c0);
        std::vector<std::thread> th;
        for (uint64_t c = c0; c < c1; ++c)
            th.emplace_back([&, c] {
                std::string& s = out[c - c0];
                s.reserve(chunk * 200);
                const uint64_t e = std::min(n, (c + 1) * chunk);
                for (uint64_t i = c * chunk; i < e; ++i) fn(s, i);
            });
        for (auto& t : th) t.join();
        for (auto& s : out)
            if (fwrite(s.data(), 1, s.size(), f) != s.size()) return false;
    }
    return true;
}

bool valid_mesh(const s2m_mesh_out* m)
{
    if (!m) return false;
    if (m->ntri && (!m->tri || !m->vtx)) return false;
    for (uint64_t i = 0; i < 3 * m->ntri; ++i)
        if (m->tri[i] >= m->nvtx) return false;
    return true;
}

}  // namespace

extern "C" {

int s2m_format_f32(float x, char* buf) { return format_f32(x, buf); }

int s2m_write_stl(const s2m_mesh_out* mesh, const char* path)
{
    if (!valid_mesh(mesh) || !path) return S2M_E_ARG;
    FILE* f = std::fopen(path, "wb");
    if (!f) return S2M_E_IO;
    bool ok = std::fputs("solid\n", f) >= 0;
    ok = ok && write_parallel(f, mesh->ntri, [&](std::string& s, uint64_t t) {
             stl_triangle(s, mesh->vtx, mesh->tri, t);
         });
    ok = ok && std::fputs("endsolid\n", f) >= 0;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? S2M_OK : S2M_E_IO;
}

int s2m_write_ply(const s2m_mesh_out* mesh, const char* path)
{
    if (!valid_mesh(mesh) || !path) return S2M_E_ARG;
    FILE* f = std::fopen(path, "wb");
    if (!f) return S2M_E_IO;
    bool ok = std::fprintf(f,
                           "ply\nformat ascii 1.0\ncomment written by rust-sdf\n"
                           "element vertex %llu\nproperty float x\nproperty float y\n"
                           "property float z\nproperty float nx\nproperty float ny\n"
                           "property float nz\nelement face %llu\n"
                           "property list uchar int vertex_index\nend_header\n",
                           (unsigned long long)mesh->nvtx, (unsigned long long)mesh->ntri) > 0;
    ok = ok && write_parallel(f, mesh->nvtx, [&](std::string& s, uint64_t i) {
             const float* v = mesh->vtx + 6 * i;
             for (int k = 0; k < 6; ++k) {
                 putf(s, v[k]);
                 s.push_back(k == 5 ? '\n' : ' ');
             }
         });
    ok = ok && write_parallel(f, mesh->ntri, [&](std::string& s, uint64_t t) {
             char b[64];
             int n = std::snprintf(b, sizeof b, "3 %u %u %u\n", mesh->tri[3 * t],
                                   mesh->tri[3 * t + 1], mesh->tri[3 * t + 2]);
             s.append(b, (size_t)n);
         });
    ok = (std::fclose(f) == 0) && ok;
    return ok ? S2M_OK : S2M_E_IO;
}

int s2m_write_to_file(const s2m_mesh_out* mesh, const char* path)
{
    if (!path) return S2M_E_ARG;
    // mesh.rs:182-196: extension compared case-insensitively; others are logged only
    const char* dot = std::strrchr(path, '.');
    const char* slash = std::strrchr(path, '/');
    std::string ext;
    if (dot && (!slash || dot > slash))
        for (const char* c = dot + 1; *c; ++c) ext.push_back((char)std::toupper((unsigned char)*c));
    if (ext == "STL") return s2m_write_stl(mesh, path);
    if (ext == "PLY") return s2m_write_ply(mesh, path);
    std::fprintf(stderr, "[sdf2mesh] Unknown file extension: %s\n", ext.c_str());
    return S2M_OK;
}

}  // extern "C"
