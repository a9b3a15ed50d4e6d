// A GPU consumer of the C++ mirror (include/rt/ray_query.hpp), the way an
// integration of the reference's RaygenTask would drive it: a scene made from
// the reference's 48-byte Triangle records (normals included) plus analytic
// objects, then one raygen pass.  Built with g++ (no HIP compiler), linked
// against librt_hip.so.
//
// Input file (argv[1]): u32 ntris, nspheres, ncubes, nplanes, width, height;
// ntris * 48 B Triangle records; nspheres * 4 floats; ncubes * 6 floats;
// nplanes * 4 floats; the camera as 14 floats (eye, p0, p1, p2, jitter).
// Output file (argv[2]): dir_object then uv_normal, float4 [height][width]
// each; uv_normal starts zeroed (raygen writes it on hits only).
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
        std::fprintf(stderr, "usage: gpu_consumer in.bin out.bin\n");
        return 2;
    }
    FILE *in = std::fopen(argv[1], "rb");
    if (!in) return 2;
    uint32_t hdr[6];
    if (!read_n(in, hdr, 6)) return 2;
    const uint32_t nt = hdr[0], ns = hdr[1], nc = hdr[2], np = hdr[3], w = hdr[4], h = hdr[5];
    std::vector<unsigned char> tris(static_cast<size_t>(nt) * 48);
    std::vector<float> sph(static_cast<size_t>(ns) * 4), cub(static_cast<size_t>(nc) * 6),
        pln(static_cast<size_t>(np) * 4), cam(14);
    if (!read_n(in, tris.data(), tris.size()) || !read_n(in, sph.data(), sph.size()) ||
        !read_n(in, cub.data(), cub.size()) || !read_n(in, pln.data(), pln.size()) || !read_n(in, cam.data(), 14))
        return 2;
    std::fclose(in);

    igx::rt::hip::Context ctx(0);
    auto scene = igx::rt::hip::Scene::fromPacked(ctx, tris.data(), nt);
    scene->setAnalytic(sph.data(), ns, cub.data(), nc, pln.data(), np);
    igx::rt::hip::RaygenTask raygen(ctx);
    raygen.switchToScene(scene.get());
    raygen.resize(w, h);
    rt_camera c;
    std::memcpy(c.eye, cam.data(), 3 * sizeof(float));
    std::memcpy(c.p0, cam.data() + 3, 3 * sizeof(float));
    std::memcpy(c.p1, cam.data() + 6, 3 * sizeof(float));
    std::memcpy(c.p2, cam.data() + 9, 3 * sizeof(float));
    std::memcpy(c.jitter, cam.data() + 12, 2 * sizeof(float));
    raygen.setCamera(c);
    const size_t bytes = static_cast<size_t>(w) * h * 4 * sizeof(float);
    if (hipMemset(raygen.getTexture(1), 0, bytes) != hipSuccess) return 3;
    raygen.prepareCommandList(nullptr);  // normals: the scene's own
    std::vector<float> out(static_cast<size_t>(w) * h * 8);
    if (hipMemcpy(out.data(), raygen.getTexture(0), bytes, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(out.data() + static_cast<size_t>(w) * h * 4, raygen.getTexture(1), bytes, hipMemcpyDeviceToHost) !=
            hipSuccess)
        return 3;
    uint32_t code = 0;
    igx::rt::hip::check(rt_check_device_error(ctx.get(), nullptr, &code), ctx.get());
    if (code != 0) return 4;
    FILE *f = std::fopen(argv[2], "wb");
    if (!f || std::fwrite(out.data(), sizeof(float), out.size(), f) != out.size()) return 2;
    std::fclose(f);
    std::printf("raygen %ux%u ok\n", w, h);
    return 0;
}
