// A C++ host program over the C ABI, as the reference's editor would drive it
// (EditorLayer::OnRender -> RT::RecordRenderCmdBuf, Editor/src/EditorLayer.cpp:612):
// create a context, load the 256 x 256 box scene, build the acceleration
// structure, render and accumulate frames into a device image, read it back.
//
//   hipcc -O2 -I include snake-vk_amd/examples/render_room.cpp -L snake-vk_amd -lsnkpt \
//         -Wl,-rpath,'$ORIGIN/..' -o snake-vk_amd/examples/render_room
//   render_room [frames] [out.pfm]
//
// Prints the image's mean radiance and, with a file name, writes the running
// mean as a PFM (RGB float, bottom row first).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "snkpt.h"

namespace {

int check(snkpt_ctx* ctx, snkpt_status s, const char* what) {
    if (s == SNKPT_OK) return 0;
    std::fprintf(stderr, "%s failed (%d): %s\n", what, (int)s, ctx ? snkpt_last_error(ctx) : "");
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    const int frames = argc > 1 ? std::atoi(argv[1]) : 16;
    const char* out_path = argc > 2 ? argv[2] : nullptr;
    snkpt_ctx* ctx = nullptr;
    if (check(nullptr, snkpt_create(0, &ctx), "snkpt_create")) return 1;
    snkpt_scene* scene = nullptr;
    if (check(ctx, snkpt_scene_generate(1, 0, &scene), "snkpt_scene_generate")) return 1;
    if (check(ctx, snkpt_scene_load(ctx, scene), "snkpt_scene_load")) return 1;
    if (check(ctx, snkpt_build_accel(ctx, nullptr), "snkpt_build_accel")) return 1;
    const uint32_t w = scene->width, h = scene->height;

    float* image = nullptr;
    if (hipMalloc(&image, (size_t)w * h * 4 * sizeof(float)) != hipSuccess) return 1;
    snkpt_settings s{};
    s.mode = SNKPT_MODE_PT;
    s.spp = 1;
    s.max_depth = scene->max_depth;
    s.rr_start = scene->rr_start;
    s.seed = 7;
    for (int f = 0; f < frames; ++f) {
        s.frame = (uint32_t)f;
        s.flags = SNKPT_FLAG_ACCUMULATE | (f == 0 ? SNKPT_FLAG_RESET : 0u);
        if (check(ctx, snkpt_render(ctx, &scene->camera, &s, image, w, h, SNKPT_FORMAT_RGBA32F, nullptr),
                  "snkpt_render"))
            return 1;
    }
    std::vector<float> px((size_t)w * h * 4);
    if (hipMemcpy(px.data(), image, px.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return 1;
    double sum = 0.0;
    for (size_t i = 0; i < (size_t)w * h; ++i) sum += (double)px[4 * i] + px[4 * i + 1] + px[4 * i + 2];
    std::printf("render_room: %u x %u, %d frames, mean radiance %.6f\n", w, h, frames, sum / (3.0 * w * h));
    if (out_path) {
        FILE* fp = std::fopen(out_path, "wb");
        if (!fp) return 1;
        std::fprintf(fp, "PF\n%u %u\n-1.0\n", w, h);
        for (uint32_t y = h; y-- > 0;)
            for (uint32_t x = 0; x < w; ++x) std::fwrite(&px[4 * ((size_t)y * w + x)], sizeof(float), 3, fp);
        std::fclose(fp);
    }
    (void)hipFree(image);
    snkpt_scene_free(scene);
    snkpt_destroy(ctx);
    return 0;
}
