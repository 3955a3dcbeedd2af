// nb_tune.hip -- launch-configuration autotuner for the all-pairs kernel.
//
// For a context it times every candidate (targets per thread x source
// segments) on the context's own particles with HIP events and keeps the
// fastest.  Results are remembered per (precision mode, ceil(log2 N)) and can
// be saved to / loaded from a small JSON file, so later contexts of a similar
// size start from the tuned configuration.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "nb_ctx.h"

namespace nb {

struct TuneEntry {
    int mode;
    int log2n;
    nb_tuning t;
    double ms;
};

static std::mutex g_mu;
static std::vector<TuneEntry> g_table;

static int log2_bucket(int64_t n) {
    int b = 0;
    while (((int64_t)1 << b) < n && b < 62) ++b;
    return b;
}

void tuner_lookup(int mode, int64_t n, nb_tuning *t) {
    std::lock_guard<std::mutex> lk(g_mu);
    const int b = log2_bucket(n);
    for (const TuneEntry &e : g_table)
        if (e.mode == mode && e.log2n == b) {
            *t = e.t;
            return;
        }
}

static void tuner_store(int mode, int64_t n, const nb_tuning &t, double ms) {
    std::lock_guard<std::mutex> lk(g_mu);
    const int b = log2_bucket(n);
    for (TuneEntry &e : g_table)
        if (e.mode == mode && e.log2n == b) {
            e.t = t;
            e.ms = ms;
            return;
        }
    g_table.push_back(TuneEntry{mode, b, t, ms});
}

static int time_force(nb_ctx *c, int reps, float *ms_out) {
    hipEvent_t e0, e1;
    NB_HIP(hipEventCreate(&e0));
    NB_HIP(hipEventCreate(&e1));
    int rc = ctx_run_force(c, false, false);  // warm-up
    if (rc == NB_OK) {
        (void)hipEventRecord(e0, c->stream);
        for (int r = 0; r < reps && rc == NB_OK; ++r) rc = ctx_run_force(c, false, false);
        (void)hipEventRecord(e1, c->stream);
        hipError_t e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess && rc == NB_OK) {
            set_error("autotune timing failed: %s", hipGetErrorString(e));
            rc = NB_ERR_HIP;
        }
        *ms_out = ms / (float)reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int autotune_ctx(nb_ctx *c) {
    const int64_t n = c->n;
    if (n == 0) return NB_OK;
    const nb_tuning keep = c->tun;
    const int keep_interval = c->sort.interval;
    // Sort once so every candidate sees the Morton order it will run on.
    if (c->mode != NB_MODE_FP64) NB_TRY(morton_reorder(c));
    c->sort.interval = 0;
    const int64_t ntiles = (n + BS - 1) / BS;
    const int segs[8] = {1, 2, 4, 8, 16, 32, 64, 128};
    nb_tuning best = keep;
    float best_ms = 1e30f;
    // A single evaluation is timed per candidate when it is long.
    // candidates (targets per thread, far path); the weighted and FP32
    // matrix-core far paths exist in the SPLIT modes only, the bf16 one in
    // the SPLIT and FP32 modes
    const int cand[9][2] = {{1, 0}, {2, 0}, {4, 0}, {1, 4}, {1, 3}, {2, 3}, {4, 3}, {1, 1}, {1, 2}};
    const bool split = c->mode == NB_MODE_SPLIT || c->mode == NB_MODE_SPLIT_CENTRAL;
    const int ncand = split ? 9 : (c->mode == NB_MODE_FP32 ? 4 : 3);
    for (int ti = 0; ti < ncand; ++ti) {
        const int tpt = cand[ti][0];
        const int mf = cand[ti][1];
        if (c->mode == NB_MODE_FP64 && tpt != 1) continue;
        const int64_t nblk = (n + (int64_t)BS * tpt - 1) / ((int64_t)BS * tpt);
        for (int si = 0; si < 8; ++si) {
            const int nseg = segs[si];
            if (nseg > ntiles) break;
            if (nblk * nseg > (1 << 20)) break;
            if ((double)n * nseg * 32.0 > 4.0e9) break;  // partial-sum buffer bound
            if (nseg > 1 && nblk * (nseg / 2) >= 65536) break;  // the grid was already full
            if (nblk * nseg < 64 && nseg < ntiles && si + 1 < 8) continue;  // far too few blocks
            c->tun.targets_per_thread = tpt;
            c->tun.j_segments = nseg;
            c->tun.far_path = mf;
            const int reps = (double)n * (double)n > 1e11 ? 1 : 3;
            float ms = 0.f;
            int rc = time_force(c, reps, &ms);
            if (rc != NB_OK) {
                c->tun = keep;
                c->sort.interval = keep_interval;
                return rc;
            }
            if (ms < best_ms) {
                best_ms = ms;
                best = c->tun;
            }
        }
    }
    best.sort_interval = c->mode == NB_MODE_FP64 ? 0 : (keep.sort_interval > 0 ? keep.sort_interval : 32);
    c->tun = best;
    c->sort.interval = best.sort_interval;
    ctx_drop_graph(c);
    tuner_store(c->mode, n, best, best_ms);
    // Leave Acc consistent with the chosen configuration.
    NB_TRY(ctx_run_force(c, false, false));
    NB_HIP(hipStreamSynchronize(c->stream));
    return NB_OK;
}

}  // namespace nb

// ---------------------------------------------------------------------------
// Persistence of the tuning table (plain JSON, one object per entry).
// ---------------------------------------------------------------------------
extern "C" int nb_tuner_save(const char *path) {
    nb::clear_error();
    if (!path) {
        nb::set_error("null path");
        return NB_ERR_INVALID;
    }
    FILE *f = fopen(path, "w");
    if (!f) {
        nb::set_error("cannot open %s for writing", path);
        return NB_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(nb::g_mu);
    fprintf(f, "{\"version\": 1, \"entries\": [");
    for (size_t k = 0; k < nb::g_table.size(); ++k) {
        const nb::TuneEntry &e = nb::g_table[k];
        fprintf(f,
                "%s\n  {\"mode\": %d, \"log2n\": %d, \"targets_per_thread\": %d, "
                "\"j_segments\": %d, \"sort_interval\": %d, \"far_path\": %d, \"ms\": %.6g}",
                k ? "," : "", e.mode, e.log2n, e.t.targets_per_thread, e.t.j_segments,
                e.t.sort_interval, e.t.far_path, e.ms);
    }
    fprintf(f, "\n]}\n");
    fclose(f);
    return NB_OK;
}

static bool json_int(const std::string &obj, const char *key, long *out) {
    std::string k = std::string("\"") + key + "\"";
    size_t p = obj.find(k);
    if (p == std::string::npos) return false;
    p = obj.find(':', p + k.size());
    if (p == std::string::npos) return false;
    char *end = nullptr;
    long v = strtol(obj.c_str() + p + 1, &end, 10);
    if (end == obj.c_str() + p + 1) return false;
    *out = v;
    return true;
}

static bool json_double(const std::string &obj, const char *key, double *out) {
    std::string k = std::string("\"") + key + "\"";
    size_t p = obj.find(k);
    if (p == std::string::npos) return false;
    p = obj.find(':', p + k.size());
    if (p == std::string::npos) return false;
    char *end = nullptr;
    double v = strtod(obj.c_str() + p + 1, &end);
    if (end == obj.c_str() + p + 1) return false;
    *out = v;
    return true;
}

extern "C" int nb_tuner_load(const char *path) {
    nb::clear_error();
    if (!path) {
        nb::set_error("null path");
        return NB_ERR_INVALID;
    }
    FILE *f = fopen(path, "r");
    if (!f) {
        nb::set_error("cannot open %s", path);
        return NB_ERR_INVALID;
    }
    std::string text;
    char buf[4096];
    size_t got;
    while ((got = fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, got);
    fclose(f);
    std::vector<nb::TuneEntry> entries;
    size_t pos = text.find('[');
    if (pos == std::string::npos) {
        nb::set_error("%s: no entries array", path);
        return NB_ERR_INVALID;
    }
    while (true) {
        size_t a = text.find('{', pos);
        if (a == std::string::npos) break;
        size_t b = text.find('}', a);
        if (b == std::string::npos) break;
        std::string obj = text.substr(a, b - a + 1);
        long mode, log2n, tpt, nseg, si = 0, mf = 0;
        if (!json_int(obj, "mode", &mode) || !json_int(obj, "log2n", &log2n) ||
            !json_int(obj, "targets_per_thread", &tpt) || !json_int(obj, "j_segments", &nseg)) {
            nb::set_error("%s: malformed entry", path);
            return NB_ERR_INVALID;
        }
        json_int(obj, "sort_interval", &si);
        json_int(obj, "far_path", &mf);
        if ((tpt != 1 && tpt != 2 && tpt != 4) || nseg < 1 || nseg > 256 || mode < 0 || mode > 3 ||
            si < 0 || mf < 0 || mf > 4 || ((mf == 1 || mf == 2 || mf == 4) && tpt != 1) ||
            (mode == NB_MODE_FP32 && mf != 0 && mf != 4)) {
            nb::set_error("%s: entry out of range", path);
            return NB_ERR_INVALID;
        }
        double ms = 0.0;   // the measured time, kept so that a saved table still carries it
        json_double(obj, "ms", &ms);
        nb::TuneEntry e{(int)mode, (int)log2n, {(int32_t)tpt, (int32_t)nseg, (int32_t)si, (int32_t)mf},
                        ms};
        entries.push_back(e);
        pos = b + 1;
    }
    std::lock_guard<std::mutex> lk(nb::g_mu);
    for (const nb::TuneEntry &e : entries) {
        bool found = false;
        for (nb::TuneEntry &g : nb::g_table)
            if (g.mode == e.mode && g.log2n == e.log2n) {
                g = e;
                found = true;
            }
        if (!found) nb::g_table.push_back(e);
    }
    return NB_OK;
}

extern "C" int nb_tuner_clear(void) {
    nb::clear_error();
    std::lock_guard<std::mutex> lk(nb::g_mu);
    nb::g_table.clear();
    return NB_OK;
}
