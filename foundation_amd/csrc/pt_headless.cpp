// pt_headless.cpp — headless driver of the path-tracing pass through the C++
// wrapper: generates one of the benchmark scenes, renders `frames` frames of
// `spp` samples with PathTracerPass, and writes the mean radiance.
//
// usage: pt_headless <kind 1..4> <width> <height> <frames> <spp> <max_depth>
//                    <rr_depth> <seed> <cache 0|1> <out.f32|out.ppm>
// A ".f32" output is width*height*3 little-endian floats; ".ppm" is the
// R8G8B8A8 image without alpha.  Prints one JSON line of statistics.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "PathTracerPass.hpp"

namespace {

// A Core::Allocator-shaped host allocator (Allocate / Reallocate /
// Deallocate) that counts its live bytes, to show the borrowed-allocator path.
struct CountingAllocator {
    size_t live = 0, peak = 0, calls = 0;
    void* Allocate(size_t size, size_t alignment)
    {
        ++calls;
        const size_t a = alignment < sizeof(void*) ? sizeof(void*) : alignment;
        void* raw = std::malloc(size + a + sizeof(size_t) * 2);
        if (!raw) return nullptr;
        uintptr_t p = (reinterpret_cast<uintptr_t>(raw) + sizeof(size_t) * 2 + a - 1) & ~uintptr_t(a - 1);
        reinterpret_cast<size_t*>(p)[-1] = size;
        reinterpret_cast<void**>(p)[-2] = raw;
        live += size;
        if (live > peak) peak = live;
        return reinterpret_cast<void*>(p);
    }
    void Deallocate(void* p)
    {
        if (!p) return;
        live -= reinterpret_cast<size_t*>(p)[-1];
        std::free(reinterpret_cast<void**>(p)[-2]);
    }
    void* Reallocate(void* p, size_t size, size_t alignment)
    {
        void* q = Allocate(size, alignment);
        if (q && p) {
            const size_t old = reinterpret_cast<size_t*>(p)[-1];
            std::memcpy(q, p, old < size ? old : size);
            Deallocate(p);
        }
        return q;
    }
};

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 11) {
        std::fprintf(stderr,
                     "usage: %s <kind 1..4> <width> <height> <frames> <spp> <max_depth> <rr_depth> <seed> "
                     "<cache 0|1> <out.f32|out.ppm>\n",
                     argv[0]);
        return 2;
    }
    const uint32_t kind = uint32_t(std::atoi(argv[1]));
    fnd::PathTracerPass::Desc d;
    d.width = uint32_t(std::atoi(argv[2]));
    d.height = uint32_t(std::atoi(argv[3]));
    const int frames = std::atoi(argv[4]);
    d.spp = uint32_t(std::atoi(argv[5]));
    d.max_depth = uint32_t(std::atoi(argv[6]));
    d.rr_depth = uint32_t(std::atoi(argv[7]));
    d.seed = std::strtoull(argv[8], nullptr, 10);
    d.use_cache = d.train_cache = std::atoi(argv[9]) != 0;
    const std::string out = argv[10];

    fnd_gen_params gp{};
    if (fnd_gen_default_params(kind, &gp) != FND_OK) {
        std::fprintf(stderr, "unknown scene kind %u\n", kind);
        return 2;
    }
    fnd_generated_scene* gen = nullptr;
    if (fnd_gen_scene(&gp, &gen) != FND_OK) return 1;
    fnd_scene_desc sd{};
    fnd_camera cam{};
    fnd_gen_scene_get(gen, &sd, &cam);

    CountingAllocator host;
    double ms = 0.0;
    uint64_t rays = 0;
    {
        fnd::PathTracerPass pass(d, &host);
        const fnd_handle scene = pass.CreateScene(sd);
        if (d.use_cache) pass.ConfigureCache(nullptr);
        pass.SetView(scene, cam);
        const auto t0 = std::chrono::steady_clock::now();
        for (int f = 0; f < frames; ++f) pass.Execute();
        ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        rays = pass.Stats().render_rays;
        std::vector<float> rgb;
        pass.ReadRGB(rgb);
        FILE* fp = std::fopen(out.c_str(), "wb");
        if (!fp) return 1;
        if (out.size() > 4 && out.compare(out.size() - 4, 4, ".ppm") == 0) {
            std::vector<uint8_t> rgba;
            pass.ReadRGBA8(rgba);
            std::fprintf(fp, "P6\n%u %u\n255\n", d.width, d.height);
            for (size_t i = 0; i < rgba.size(); i += 4) std::fwrite(&rgba[i], 1, 3, fp);
        } else {
            std::fwrite(rgb.data(), sizeof(float), rgb.size(), fp);
        }
        std::fclose(fp);
        pass.DestroyScene(scene);
    }
    fnd_gen_scene_free(gen);
    std::printf("{\"frames\": %d, \"spp\": %u, \"wall_ms\": %.3f, \"render_rays\": %llu, \"host_alloc_calls\": %zu, "
                "\"host_peak_bytes\": %zu, \"host_live_bytes_after\": %zu}\n",
                frames, d.spp, ms, (unsigned long long)rays, host.calls, host.peak, host.live);
    return host.live == 0 ? 0 : 3;
}
