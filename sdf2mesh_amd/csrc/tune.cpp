// tune.cpp - the autotuner's table of chosen shapes and its JSON cache.
//
// The meshing calls consult a process-wide table keyed by (device, program class, grid)
// for the shapes they would otherwise pick by rule: the z-passes of a single tile-sparse
// job streamed to host memory, the cell planes per dense slab, the cells per pass of a
// batch.  s2m_tune (api.cpp) fills it by timing candidates; s2m_tuning_save /
// s2m_tuning_load keep it across processes as a JSON file:
//
//   {"entries": [{"key": "gfx950:256|fold|512x512x512|7", "stream_passes": 4,
//                 "dense_slab": 0, "batch_cells": 0, "ms": 0.95}, ...]}
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "internal.h"

namespace {

struct Entry {
    s2m::TuneChoice c;
    double ms;
};

std::mutex g_mu;
std::map<std::string, Entry>& table()
{
    static std::map<std::string, Entry>* t = new std::map<std::string, Entry>();
    return *t;
}
thread_local const s2m::TuneChoice* g_override = nullptr;

// Minimal reader for the flat JSON the writer below produces (and hand edits of it):
// objects of string and number fields inside an "entries" array.
struct Reader {
    const char* p;
    const char* end;
    void ws()
    {
        while (p < end && std::isspace((unsigned char)*p)) ++p;
    }
    bool eat(char c)
    {
        ws();
        if (p < end && *p == c) {
            ++p;
            return true;
        }
        return false;
    }
    bool str(std::string& out)
    {
        ws();
        if (p >= end || *p != '"') return false;
        ++p;
        out.clear();
        while (p < end && *p != '"') {
            if (*p == '\\' && p + 1 < end) ++p;
            out.push_back(*p++);
        }
        return p < end && *p++ == '"';
    }
    bool num(double& out)
    {
        ws();
        char* e = nullptr;
        out = std::strtod(p, &e);
        if (e == p) return false;
        p = e;
        return true;
    }
};

}  // namespace

namespace s2m {

bool tune_lookup(const std::string& key, TuneChoice* out)
{
    if (g_override) {
        *out = *g_override;
        return true;
    }
    std::lock_guard<std::mutex> g(g_mu);
    auto it = table().find(key);
    if (it == table().end()) return false;
    *out = it->second.c;
    return true;
}

void tune_store(const std::string& key, const TuneChoice& c, double ms)
{
    std::lock_guard<std::mutex> g(g_mu);
    table()[key] = Entry{c, ms};
}

bool tune_cached(const std::string& key, TuneChoice* out, double* ms)
{
    std::lock_guard<std::mutex> g(g_mu);
    auto it = table().find(key);
    if (it == table().end()) return false;
    *out = it->second.c;
    *ms = it->second.ms;
    return true;
}

void tune_override(const TuneChoice* c) { g_override = c; }

uint32_t tune_pick(const double* ms, uint32_t n)
{
    // the first candidate holds until another is faster by more than 2 %: a choice that
    // depends only on the timings, in candidate order, and ignores differences within
    // the noise of repeated runs
    uint32_t best = 0;
    for (uint32_t i = 1; i < n; ++i)
        if (ms[i] < ms[best] * 0.98) best = i;
    return best;
}

}  // namespace s2m

extern "C" {

int s2m_tuning_save(const char* path)
{
    if (!path) return s2m::set_error(S2M_E_ARG, "path is NULL");
    FILE* f = std::fopen(path, "w");
    if (!f) return s2m::set_error(S2M_E_IO, "cannot write the tuning cache");
    std::lock_guard<std::mutex> g(g_mu);
    std::fprintf(f, "{\"entries\": [");
    bool first = true;
    for (const auto& kv : table()) {
        std::fprintf(f, "%s\n  {\"key\": \"%s\", \"stream_passes\": %u, \"dense_slab\": %u, "
                        "\"batch_cells\": %llu, \"ms\": %.6g}",
                     first ? "" : ",", kv.first.c_str(), kv.second.c.stream_passes, kv.second.c.dense_slab,
                     (unsigned long long)kv.second.c.batch_cells, kv.second.ms);
        first = false;
    }
    std::fprintf(f, "\n]}\n");
    return std::fclose(f) == 0 ? S2M_OK : s2m::set_error(S2M_E_IO, "cannot write the tuning cache");
}

int s2m_tuning_load(const char* path)
{
    if (!path) return s2m::set_error(S2M_E_ARG, "path is NULL");
    FILE* f = std::fopen(path, "rb");
    if (!f) return s2m::set_error(S2M_E_IO, "cannot read the tuning cache");
    std::string text;
    char buf[4096];
    for (size_t n; (n = std::fread(buf, 1, sizeof buf, f)) > 0;) text.append(buf, n);
    std::fclose(f);
    Reader r{text.data(), text.data() + text.size()};
    std::string name;
    auto bad = [] { return s2m::set_error(S2M_E_PARSE, "malformed tuning cache"); };
    if (!r.eat('{') || !r.str(name) || name != "entries" || !r.eat(':') || !r.eat('[')) return bad();
    std::map<std::string, Entry> got;
    if (!r.eat(']')) {
        do {
            if (!r.eat('{')) return bad();
            std::string key;
            Entry e{};
            do {
                std::string field;
                if (!r.str(field) || !r.eat(':')) return bad();
                if (field == "key") {
                    if (!r.str(key)) return bad();
                } else {
                    double v = 0;
                    if (!r.num(v) || v < 0) return bad();
                    if (field == "stream_passes") e.c.stream_passes = (uint32_t)v;
                    else if (field == "dense_slab") e.c.dense_slab = (uint32_t)v;
                    else if (field == "batch_cells") e.c.batch_cells = (uint64_t)v;
                    else if (field == "ms") e.ms = v;
                }
            } while (r.eat(','));
            if (!r.eat('}') || key.empty()) return bad();
            got[key] = e;
        } while (r.eat(','));
        if (!r.eat(']')) return bad();
    }
    if (!r.eat('}')) return bad();
    std::lock_guard<std::mutex> g(g_mu);
    for (auto& kv : got) table()[kv.first] = kv.second;
    return S2M_OK;
}

void s2m_tuning_clear(void)
{
    std::lock_guard<std::mutex> g(g_mu);
    table().clear();
}

uint32_t s2m_tune_pick(const double* ms, uint32_t n) { return ms && n ? s2m::tune_pick(ms, n) : 0u; }

}  // extern "C"
