// A GPU consumer of the C++ mirror's ShadowTask (include/rt/ray_query.hpp),
// the way an integration of the reference's ShadowTask would drive it:
// switchToScene, resize (the ballot buffer of shadow_task.cpp:146-165), then
// one pass per frame over caller-made shadow rays, at two image sizes.
// Built with g++ (no HIP compiler), linked against librt_hip.so.
//
// Input file (argv[1]): u32 ntris, samples, nsizes, then per size u32 width,
// height; ntris * 9 floats (p0 p1 p2); then per size the rays
// [samples][height][width] as 8 floats each (org, tmax, dir, prev_id bits).
// Output file (argv[2]): per size u64 outputBytes(), then that many bytes of
// ballot words.
#include <cstdio>
#include <cstring>
#include <vector>

#include "rt/ray_query.hpp"

template <typename T>
static bool read_n(FILE *f, T *dst, size_t n) {
    return n == 0 || std::fread(dst, sizeof(T), n, f) == n;
}

int main(int argc, char **argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: shadow_consumer in.bin out.bin\n");
        return 2;
    }
    FILE *in = std::fopen(argv[1], "rb");
    if (!in) return 2;
    uint32_t hdr[3];
    if (!read_n(in, hdr, 3) || hdr[2] == 0 || hdr[2] > 8) return 2;
    const uint32_t nt = hdr[0], samples = hdr[1], nsizes = hdr[2];
    std::vector<uint32_t> dims(2 * nsizes);
    if (!read_n(in, dims.data(), dims.size())) return 2;
    std::vector<float> tris(static_cast<size_t>(nt) * 9);
    if (!read_n(in, tris.data(), tris.size())) return 2;
    std::vector<std::vector<float>> rays(nsizes);
    for (uint32_t k = 0; k < nsizes; ++k) {
        rays[k].resize(static_cast<size_t>(samples) * dims[2 * k] * dims[2 * k + 1] * 8);
        if (!read_n(in, rays[k].data(), rays[k].size())) return 2;
    }
    std::fclose(in);

    igx::rt::hip::Context ctx(0);
    igx::rt::hip::Scene scene(ctx, tris.data(), nt);
    igx::rt::hip::ShadowTask shadow(ctx, samples);
    shadow.switchToScene(&scene);
    FILE *f = std::fopen(argv[2], "wb");
    if (!f) return 2;
    for (uint32_t k = 0; k < nsizes; ++k) {
        shadow.resize(dims[2 * k], dims[2 * k + 1]);
        void *d_rays = nullptr;
        const size_t rb = rays[k].size() * sizeof(float);
        if (hipMalloc(&d_rays, rb) != hipSuccess ||
            hipMemcpy(d_rays, rays[k].data(), rb, hipMemcpyHostToDevice) != hipSuccess)
            return 3;
        shadow.prepareCommandList(static_cast<const rt_ray *>(d_rays), nullptr);
        const uint64_t bytes = shadow.outputBytes();
        std::vector<unsigned char> out(bytes);
        if (hipMemcpy(out.data(), shadow.getOutput(), bytes, hipMemcpyDeviceToHost) != hipSuccess) return 3;
        (void)hipFree(d_rays);
        if (std::fwrite(&bytes, sizeof bytes, 1, f) != 1 || std::fwrite(out.data(), 1, bytes, f) != bytes) return 2;
    }
    std::fclose(f);
    uint32_t code = 0;
    igx::rt::hip::check(rt_check_device_error(ctx.get(), nullptr, &code), ctx.get());
    if (code != 0) return 4;
    std::printf("shadow ok\n");
    return 0;
}
