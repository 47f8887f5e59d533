// autotune.cpp — launch-configuration search for a render workload.
//
// The reference has no tuner (it has no compute work).  The knobs are those
// of fnd_tuning: traversal block size and LDS stack depth (the instantiated
// pairs), persistent waves per CU, the lane-refill threshold, wavefront
// queue capacity and pixel tile.  Candidates are the library default plus a
// fixed pseudo-random sample of the product space; successive halving
// (Jamieson & Talwalkar, AISTATS 2016) keeps the fastest third each round
// and doubles the timing renders of the survivors.  Timing renders are the
// caller's jobs at 1 spp with cache training off, into scratch framebuffers.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <set>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "ctx.hpp"

using namespace fnd;

namespace {

struct Candidate {
    fnd_tuning t{};
    double ms = 0.0;  // mean of the last round it took part in
};

// Instantiated (block, LDS stack) pairs of the traversal kernel.
const uint32_t kPairs[][2] = {{64, 16}, {64, 32}, {128, 8}, {128, 12}, {128, 16}, {128, 32}, {256, 8}, {256, 12}, {256, 16}};
const uint32_t kWaves[] = {12, 16, 20, 24, 32};
const uint32_t kRefill[] = {8, 12, 16, 24, 32, 48};
const uint32_t kQueue[] = {1u << 20, 1u << 21, 1u << 22};
const uint32_t kTile[] = {4, 8, 16};

uint32_t pcg_next(uint64_t& state)
{
    const uint64_t old = state;
    state = old * 6364136223846793005ULL + 1442695040888963407ULL;
    const uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = uint32_t(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

std::vector<Candidate> candidates(const fnd_tuning& def, uint32_t count)
{
    std::vector<Candidate> out;
    std::set<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t>> seen;
    auto add = [&](const fnd_tuning& t) {
        const auto k = std::make_tuple(t.block_size, t.stack_lds, t.waves_per_cu, t.refill, t.queue_capacity, t.tile);
        if (seen.insert(k).second) out.push_back(Candidate{t, 0.0});
    };
    add(def);
    uint64_t state = 0x7E57AB1Eull;
    for (int guard = 0; out.size() < count && guard < 10000; ++guard) {
        fnd_tuning t{};
        const uint32_t p = pcg_next(state) % uint32_t(std::size(kPairs));
        t.block_size = kPairs[p][0];
        t.stack_lds = kPairs[p][1];
        t.waves_per_cu = kWaves[pcg_next(state) % uint32_t(std::size(kWaves))];
        t.refill = kRefill[pcg_next(state) % uint32_t(std::size(kRefill))];
        t.queue_capacity = kQueue[pcg_next(state) % uint32_t(std::size(kQueue))];
        t.tile = kTile[pcg_next(state) % uint32_t(std::size(kTile))];
        add(t);
    }
    return out;
}

int ilog2(uint64_t x)
{
    int r = 0;
    while (x > 1) { x >>= 1; ++r; }
    return r;
}

bool read_cache(const char* path, const std::string& key, fnd_tuning& t, double& ms)
{
    if (!path) return false;
    std::ifstream in(path);
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string k;
        fnd_tuning c{};
        double m = 0;
        if (ls >> k >> c.block_size >> c.stack_lds >> c.waves_per_cu >> c.refill >> c.queue_capacity >> c.tile >> m &&
            k == key) {
            t = c;
            ms = m;
            return true;
        }
    }
    return false;
}

bool write_cache(const char* path, const std::string& key, const fnd_tuning& t, double ms)
{
    std::vector<std::string> keep;
    {
        std::ifstream in(path);
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string k;
            if ((ls >> k) && k != key) keep.push_back(line);
        }
    }
    std::ofstream out(path, std::ios::trunc);
    if (!out) return false;
    for (const auto& l : keep) out << l << "\n";
    out << key << " " << t.block_size << " " << t.stack_lds << " " << t.waves_per_cu << " " << t.refill << " "
        << t.queue_capacity << " " << t.tile << " " << ms << "\n";
    return bool(out);
}

}  // namespace

extern "C" fnd_status fnd_pt_autotune(fnd_pt_ctx* ctx, const fnd_render_job* jobs, uint32_t num_jobs,
                                      const char* cache_path, uint32_t flags, fnd_autotune_result* out)
{
    return guard(ctx, [&]() -> fnd_status {
        if (!ctx) return FND_E_INVALID_ARG;
        if (!jobs || !num_jobs || !out) return set_error(ctx, FND_E_INVALID_ARG, "autotune needs jobs and a result");
        if (flags & ~FND_AUTOTUNE_FORCE) return set_error(ctx, FND_E_INVALID_ARG, "unknown autotune flags 0x%x", flags);
        std::memset(out, 0, sizeof(*out));

        // workload key and the timing jobs (1 spp, no training, scratch framebuffers)
        uint64_t tris = 0, pixels = 0;
        uint32_t depth = 0;
        std::vector<fnd_render_job> tj(jobs, jobs + num_jobs);
        std::vector<fnd_handle> fbs;
        auto release = [&]() {
            for (fnd_handle h : fbs) (void)fnd_pt_framebuffer_destroy(ctx, h);
            fbs.clear();
        };
        std::set<fnd_handle> scenes;
        for (uint32_t i = 0; i < num_jobs; ++i) {
            fnd_scene_info info{};
            fnd_status st = fnd_pt_scene_info(ctx, jobs[i].scene, &info);
            if (st != FND_OK) {
                release();
                return st;
            }
            if (scenes.insert(jobs[i].scene).second) tris += info.num_triangles;
            Framebuffer* f = find_framebuffer(ctx, jobs[i].framebuffer);
            if (!f) {
                release();
                return set_error(ctx, FND_E_NOT_FOUND, "job %u: unknown framebuffer handle", i);
            }
            pixels += uint64_t(f->width) * f->height;
            depth = std::max(depth, jobs[i].max_depth);
            fnd_handle h = FND_INVALID_HANDLE;
            st = fnd_pt_framebuffer_create(ctx, f->width, f->height, 0, &h);
            if (st != FND_OK) {
                release();
                return st;
            }
            fbs.push_back(h);
            tj[i].framebuffer = h;
            tj[i].spp = 1;
            tj[i].flags &= ~FND_JOB_TRAIN_CACHE;
        }
        char key[128];
        std::snprintf(key, sizeof key, "%s/cu%u/tri%d/px%d/d%u/j%d", ctx->arch, ctx->num_cus, ilog2(tris),
                      ilog2(pixels), depth, ilog2(num_jobs));
        std::snprintf(out->key, sizeof out->key, "%s", key);

        fnd_tuning def{};
        {
            fnd_tuning zero{};
            const fnd_tuning keep = ctx->tuning;
            (void)fnd_pt_set_tuning(ctx, &zero);  // zeros select the library defaults
            def = ctx->tuning;
            ctx->tuning = keep;
        }

        if (!(flags & FND_AUTOTUNE_FORCE)) {
            fnd_tuning t{};
            double ms = 0;
            if (read_cache(cache_path, key, t, ms) && fnd_pt_set_tuning(ctx, &t) == FND_OK) {
                release();
                out->tuning = ctx->tuning;
                out->best_ms = ms;
                out->from_cache = 1;
                return FND_OK;
            }
        }

        // one timing measurement: a warm-up render (buffers sized, caches warm)
        // then `reps` renders; the mean device time of those
        auto time_one = [&](Candidate& c, uint32_t reps) -> fnd_status {
            fnd_status st = fnd_pt_set_tuning(ctx, &c.t);
            if (st != FND_OK) return st;
            double sum = 0;
            for (uint32_t r = 0; r <= reps; ++r) {
                for (uint32_t i = 0; i < num_jobs; ++i) tj[i].sample_offset = jobs[i].sample_offset + r;
                st = fnd_pt_render(ctx, tj.data(), num_jobs);
                if (st != FND_OK) return st;
                out->renders += 1;
                if (r > 0) sum += ctx->stats.last_call_ms;
            }
            c.ms = sum / reps;
            return FND_OK;
        };

        const fnd_tuning keep = ctx->tuning;
        std::vector<Candidate> pool = candidates(def, 27);
        out->candidates = uint32_t(pool.size());
        uint32_t reps = 1;
        bool first = true;
        fnd_status st = FND_OK;
        while (st == FND_OK) {
            for (auto& c : pool) {
                st = time_one(c, reps);
                if (st != FND_OK) break;
            }
            if (st != FND_OK) break;
            std::sort(pool.begin(), pool.end(), [](const Candidate& a, const Candidate& b) { return a.ms < b.ms; });
            if (first) {
                out->median_ms = pool[pool.size() / 2].ms;
                out->median = pool[pool.size() / 2].t;
                for (const auto& c : pool)
                    if (c.t.block_size == def.block_size && c.t.stack_lds == def.stack_lds &&
                        c.t.waves_per_cu == def.waves_per_cu && c.t.refill == def.refill &&
                        c.t.queue_capacity == def.queue_capacity && c.t.tile == def.tile)
                        out->default_ms = c.ms;
                first = false;
            }
            if (pool.size() == 1) break;
            pool.resize((pool.size() + 2) / 3);
            reps *= 2;
        }
        release();
        if (st != FND_OK) {
            ctx->tuning = keep;
            return st;
        }
        st = fnd_pt_set_tuning(ctx, &pool[0].t);
        if (st != FND_OK) return st;
        out->tuning = ctx->tuning;
        out->best_ms = pool[0].ms;
        if (cache_path && !write_cache(cache_path, key, pool[0].t, pool[0].ms))
            return set_error(ctx, FND_E_INVALID_ARG, "cannot write autotune cache %s", cache_path);
        return FND_OK;
    });
}
