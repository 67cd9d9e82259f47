// Autotuner: the search over cell size, Verlet skin and launch shape, the workload signature,
// and the JSON cache that persists results.  The engine it drives (applying a configuration,
// snapshots, timed device-driven runs) is in abi.hip.  The reference has no counterpart: its
// cell size, skin and launch shapes are whatever the caller passes.
#include "internal.hpp"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

using namespace dpmd;

namespace {

// ---- a small JSON reader / writer for the cache file ---------------------------------
struct JVal {
    enum Type { Null, Bool, Num, Str, Arr, Obj } type = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal* get(const std::string& k) const {
        for (const auto& kv : obj) if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

struct JParser {
    const std::string& t;
    size_t i = 0;
    explicit JParser(const std::string& text) : t(text) {}
    [[noreturn]] void fail(const char* what) const {
        throw ArgError(std::string("tune cache: malformed JSON (") + what + ") at offset " + std::to_string(i));
    }
    void ws() { while (i < t.size() && (t[i] == ' ' || t[i] == '\n' || t[i] == '\r' || t[i] == '\t')) ++i; }
    bool eat(char ch) { ws(); if (i < t.size() && t[i] == ch) { ++i; return true; } return false; }
    std::string string() {
        if (!eat('"')) fail("expected a string");
        std::string out;
        while (i < t.size() && t[i] != '"') {
            char ch = t[i++];
            if (ch == '\\') {
                if (i >= t.size()) fail("bad escape");
                const char e = t[i++];
                switch (e) {
                    case '"': case '\\': case '/': out += e; break;
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    default: fail("unsupported escape");
                }
            } else {
                out += ch;
            }
        }
        if (i >= t.size()) fail("unterminated string");
        ++i;
        return out;
    }
    JVal value() {
        ws();
        if (i >= t.size()) fail("unexpected end");
        JVal v;
        const char ch = t[i];
        if (ch == '{') {
            ++i;
            v.type = JVal::Obj;
            if (eat('}')) return v;
            do {
                std::string k = string();
                if (!eat(':')) fail("expected ':'");
                v.obj.emplace_back(std::move(k), value());
            } while (eat(','));
            if (!eat('}')) fail("expected '}'");
        } else if (ch == '[') {
            ++i;
            v.type = JVal::Arr;
            if (eat(']')) return v;
            do v.arr.push_back(value()); while (eat(','));
            if (!eat(']')) fail("expected ']'");
        } else if (ch == '"') {
            v.type = JVal::Str;
            v.str = string();
        } else if (t.compare(i, 4, "true") == 0) {
            v.type = JVal::Bool; v.b = true; i += 4;
        } else if (t.compare(i, 5, "false") == 0) {
            v.type = JVal::Bool; i += 5;
        } else if (t.compare(i, 4, "null") == 0) {
            i += 4;
        } else {
            const char* p = t.c_str() + i;
            char* end = nullptr;
            v.num = std::strtod(p, &end);
            if (end == p) fail("expected a value");
            v.type = JVal::Num;
            i += (size_t)(end - p);
        }
        return v;
    }
};

const char* kFormat = "dpmd_amd autotune cache 1";

std::string quoted(const std::string& s) {
    std::string out = "\"";
    for (char ch : s) {
        if (ch == '"' || ch == '\\') { out += '\\'; out += ch; }
        else if (ch == '\n') out += "\\n";
        else out += ch;
    }
    return out + "\"";
}

std::string num(double v) {
    char buf[40];
    std::snprintf(buf, sizeof(buf), "%.17g", v);
    return buf;
}

struct Entry { dpmd_tune_config cfg; double rate; };

// Reads every entry of a cache file; a missing file is an empty cache.
std::vector<std::pair<std::string, Entry>> read_cache(const std::string& path) {
    std::vector<std::pair<std::string, Entry>> out;
    std::ifstream f(path);
    if (!f) return out;
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    JParser p(text);
    const JVal root = p.value();
    p.ws();
    if (p.i != text.size()) p.fail("trailing characters");
    if (root.type != JVal::Obj) throw ArgError("tune cache: the top level is not an object");
    const JVal* fmt = root.get("format");
    if (!fmt || fmt->type != JVal::Str || fmt->str != kFormat)
        throw ArgError("tune cache: " + path + " is not a dpmd_amd autotune cache");
    const JVal* ents = root.get("entries");
    if (!ents || ents->type != JVal::Obj) throw ArgError("tune cache: no entries object");
    for (const auto& kv : ents->obj) {
        const JVal& e = kv.second;
        auto field = [&](const char* k) {
            const JVal* v = e.get(k);
            if (!v || v->type != JVal::Num) throw ArgError("tune cache: entry lacks numeric '" + std::string(k) + "'");
            return v->num;
        };
        Entry en;
        en.cfg.skin = field("skin");
        en.cfg.cell_factor = field("cell_factor");
        en.cfg.block = (int)field("block");
        en.cfg.max_batch = (int)field("max_batch");
        en.cfg.fused_vv = (int)field("fused_vv");
        en.rate = field("rate");
        out.emplace_back(kv.first, en);
    }
    return out;
}

void write_cache(const std::string& path, const std::vector<std::pair<std::string, Entry>>& ents) {
    std::string s = "{\n  \"format\": " + quoted(kFormat) + ",\n  \"entries\": {";
    for (size_t k = 0; k < ents.size(); ++k) {
        const Entry& e = ents[k].second;
        s += (k ? ",\n    " : "\n    ") + quoted(ents[k].first) + ": {\"skin\": " + num(e.cfg.skin) +
             ", \"cell_factor\": " + num(e.cfg.cell_factor) + ", \"block\": " + std::to_string(e.cfg.block) +
             ", \"max_batch\": " + std::to_string(e.cfg.max_batch) + ", \"fused_vv\": " +
             std::to_string(e.cfg.fused_vv) + ", \"rate\": " + num(e.rate) + "}";
    }
    s += ents.empty() ? "}\n}\n" : "\n  }\n}\n";
    const std::string tmp = path + ".tmp";
    {
        std::ofstream f(tmp, std::ios::trunc);
        if (!f) throw ArgError("tune cache: cannot write " + tmp);
        f << s;
        if (!f) throw ArgError("tune cache: write failed for " + tmp);
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) throw ArgError("tune cache: cannot rename " + tmp + " to " + path);
}

// ---- workload signature ---------------------------------------------------------------
std::string signature(dpmd_ctx* c, int integrator, int mode) {
    need(c->kind != DPMD_RIGID_BUMPY && c->method == DPMD_NEIGHBOR_CELL, "the autotuner handles disk systems with cell lists");
    const int S = c->S;
    auto rad = dl(c, c->rad);
    auto Lx = dl(c, c->Lx), Ly = dl(c, c->Ly);
    auto bnd = dl(c, c->boundary);
    int nmin = c->N, nmax = 0, walls = 0;
    double rmin = 1e300, rmax = 0.0, phi = 0.0;
    for (int s = 0; s < S; ++s) {
        const int n = c->h_off[s + 1] - c->h_off[s];
        nmin = std::min(nmin, n); nmax = std::max(nmax, n);
        walls += bnd[s] != 0 ? 1 : 0;
        double area = 0.0;
        for (int i = c->h_off[s]; i < c->h_off[s + 1]; ++i) {
            rmin = std::min(rmin, rad[i]); rmax = std::max(rmax, rad[i]);
            area += 3.14159265358979323846 * rad[i] * rad[i];
        }
        phi += area / (Lx[s] * Ly[s]);
    }
    phi /= S;
    hipDeviceProp_t prop;
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipGetDeviceProperties(&prop, dev));
    std::string arch = prop.gcnArchName;
    arch = arch.substr(0, arch.find(':'));
    char buf[512];
    std::snprintf(buf, sizeof(buf), "disk|S=%d|N=%d|n=%d..%d|pbc=%d|wall=%d|phi=%.2f|ratio=%.2f|%s|mode=%d|%s|%s", S,
                  c->N, nmin, nmax, S - walls, walls, phi, rmax / rmin, integrator == DPMD_TUNE_FIRE ? "fire" : "nve",
                  mode, arch.c_str(), dpmd_version());
    return buf;
}

bool same(const dpmd_tune_config& a, const dpmd_tune_config& b) {
    return a.skin == b.skin && a.cell_factor == b.cell_factor && a.block == b.block && a.max_batch == b.max_batch &&
           a.fused_vv == b.fused_vv;
}

}  // namespace

extern "C" {

int dpmd_cell_grid(double Lx, double Ly, double r_max, double cutoff, double cell_factor, int* nx, int* ny) {
    return guard(nullptr, [&] {
        need(Lx > 0.0 && Ly > 0.0 && r_max > 0.0 && cutoff >= 0.0 && cell_factor >= 1.0,
             "cell grid: need Lx, Ly, r_max > 0, cutoff >= 0, cell_factor >= 1");
        const double side = cell_factor * (2.0 * r_max + cutoff);
        const double fx = std::floor(Lx / side), fy = std::floor(Ly / side);
        need(fx >= 3.0 && fy >= 3.0,
             "cell grid: a box of " + num(Lx) + " x " + num(Ly) + " holds fewer than 3 x 3 cells of side " + num(side) +
                 " (cell_factor x (2 r_max + cutoff)); the periodic 3x3 stencil needs at least 3 per axis");
        need(fx <= 1e8 && fy <= 1e8, "cell grid: too many cells");
        const int x = (int)fx, y = (int)fy;
        // both stencil preconditions, stated directly
        need(x >= 3 && y >= 3 && Lx / x >= 2.0 * r_max + cutoff && Ly / y >= 2.0 * r_max + cutoff,
             "cell grid: stencil preconditions violated");
        if (nx) *nx = x;
        if (ny) *ny = y;
    });
}

int dpmd_tune_get(dpmd_ctx* c, dpmd_tune_config* t) {
    return guard(c, [&] {
        need(t != nullptr, "tune_get: config is NULL");
        tune_get(c, t);
    });
}

int dpmd_tune_apply(dpmd_ctx* c, const dpmd_tune_config* t) {
    return guard(c, [&] {
        need(t != nullptr, "tune_apply: config is NULL");
        tune_apply(c, *t);
    });
}

int dpmd_tune_signature(dpmd_ctx* c, int integrator, int mode, char* buf, int cap) {
    return guard(c, [&] {
        need(buf != nullptr && cap > 0, "tune_signature: no buffer");
        const std::string s = signature(c, integrator, mode);
        need((int)s.size() < cap, "tune_signature: buffer too small");
        std::memcpy(buf, s.c_str(), s.size() + 1);
    });
}

int dpmd_tune_cache_load(const char* path, const char* sig, dpmd_tune_config* cfg, double* rate, int* found) {
    return guard(nullptr, [&] {
        need(path && sig && found, "tune_cache_load: NULL argument");
        *found = 0;
        for (const auto& kv : read_cache(path)) {
            if (kv.first != sig) continue;
            if (cfg) *cfg = kv.second.cfg;
            if (rate) *rate = kv.second.rate;
            *found = 1;
        }
    });
}

int dpmd_tune_cache_store(const char* path, const char* sig, const dpmd_tune_config* cfg, double rate) {
    return guard(nullptr, [&] {
        need(path && sig && cfg, "tune_cache_store: NULL argument");
        auto ents = read_cache(path);
        bool put = false;
        for (auto& kv : ents)
            if (kv.first == sig) { kv.second = Entry{*cfg, rate}; put = true; }
        if (!put) ents.emplace_back(sig, Entry{*cfg, rate});
        write_cache(path, ents);
    });
}

int dpmd_autotune(dpmd_ctx* c, int integrator, const double* dt, int mode, int steps, const char* cache_path,
                  dpmd_tune_config* best_out, double* best_rate_out, dpmd_tune_trial* trials, int trial_cap,
                  int* n_trials) {
    return guard(c, [&] {
        need(integrator == DPMD_TUNE_FIRE || integrator == DPMD_TUNE_NVE, "autotune: unknown integrator");
        need(dt != nullptr, "autotune: dt is NULL");
        need(steps >= 8, "autotune: steps must be at least 8");
        need(mode == DPMD_FORCE_PBC || mode == DPMD_FORCE_WALL || mode == DPMD_FORCE_PER_SYSTEM, "autotune: bad mode");
        need_state(c->nbr_ready, "neighbors are not initialised (call init_neighbors)");
        const std::string sig = signature(c, integrator, mode);
        int ntr = 0;
        if (cache_path) {
            for (const auto& kv : read_cache(cache_path)) {
                if (kv.first != sig) continue;
                tune_apply(c, kv.second.cfg);
                if (best_out) *best_out = kv.second.cfg;
                if (best_rate_out) *best_rate_out = kv.second.rate;
                if (n_trials) *n_trials = 0;
                return;
            }
        }
        const int N = c->N;
        dpmd::DVec<double> snap;
        snap.resize(4 * (int64_t)N);
        tune_snapshot(c, snap.p);
        const std::vector<int> si0 = dl(c, c->static_index);
        // ends every path: the starting particle order and state under configuration `cfg`
        auto finish = [&](const dpmd_tune_config& cfg) {
            tune_restore(c, snap.p, mode);
            tune_restore_order(c, si0);
            tune_apply(c, cfg);
            tune_restore(c, snap.p, mode);
        };
        const int warm = std::min(steps, 32);
        auto trial = [&](const dpmd_tune_config& cfg) -> double {
            tune_apply(c, cfg);
            tune_restore(c, snap.p, mode);
            (void)tune_run(c, integrator, dt, mode, warm);
            tune_restore(c, snap.p, mode);
            HIP_OK(hipStreamSynchronize(c->st));
            const auto t0 = std::chrono::steady_clock::now();
            const int64_t work = tune_run(c, integrator, dt, mode, steps);
            HIP_OK(hipStreamSynchronize(c->st));
            const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            const double rate = (double)work / sec;
            if (trials && ntr < trial_cap) trials[ntr] = dpmd_tune_trial{cfg, rate, sec};
            ++ntr;
            return rate;
        };
        dpmd_tune_config best;
        tune_get(c, &best);
        // start from the context's skin and options, on the candidate grid at or below its own
        best.cell_factor = best.cell_factor >= 1.5 ? 1.5 : best.cell_factor >= 1.25 ? 1.25 : 1.0;
        double best_rate = 0.0;
        try {
            best_rate = trial(best);
            // One knob at a time, in the order of expected effect; a change must beat the
            // best so far by 1 % to count (timing noise).
            const double skins[] = {0.1, 0.15, 0.2, 0.3, 0.4, 0.6};
            const double factors[] = {1.0, 1.25, 1.5};
            const int blocks[] = {64, 128, 256, 512};
            const int batches[] = {16, 32, 64, 128};
            auto consider = [&](dpmd_tune_config cand) {
                if (same(cand, best)) return;
                try {
                    tune_apply(c, cand);
                } catch (const ArgError&) {
                    return;   // a grid that breaks the stencil preconditions for some box
                }
                const double r = trial(cand);
                if (r > best_rate * 1.01) { best = cand; best_rate = r; }
            };
            const dpmd_tune_config base0 = best;
            for (double v : skins) { dpmd_tune_config t = base0; t.skin = v; consider(t); }
            const dpmd_tune_config base1 = best;
            for (double v : factors) { dpmd_tune_config t = base1; t.cell_factor = v; consider(t); }
            const dpmd_tune_config base2 = best;
            for (int v : blocks) { dpmd_tune_config t = base2; t.block = v; consider(t); }
            if (integrator == DPMD_TUNE_NVE) {
                const dpmd_tune_config base3 = best;
                for (int v : {0, 1}) { dpmd_tune_config t = base3; t.fused_vv = v; consider(t); }
            }
            const dpmd_tune_config base4 = best;
            for (int v : batches) { dpmd_tune_config t = base4; t.max_batch = v; consider(t); }
        } catch (...) {
            finish(best);
            throw;
        }
        finish(best);
        HIP_OK(hipStreamSynchronize(c->st));
        if (cache_path) {
            auto ents = read_cache(cache_path);
            bool put = false;
            for (auto& kv : ents)
                if (kv.first == sig) { kv.second = Entry{best, best_rate}; put = true; }
            if (!put) ents.emplace_back(sig, Entry{best, best_rate});
            write_cache(cache_path, ents);
        }
        if (best_out) *best_out = best;
        if (best_rate_out) *best_rate_out = best_rate;
        if (n_trials) *n_trials = ntr;
    });
}

}  // extern "C"
