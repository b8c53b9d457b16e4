// Minimal C++ host of the drop-in path: builds a small Cornell-style room, renders it through
// GpuPathRenderer (CPURenderer's role over the C ABI) and writes the tone-mapped image as PFM.
//   render_demo <out.pfm> [width height frames [spatial_denoise 0/1 [accumulated.pfm]]]
// The last line of stdout is a JSON record with the frame statistics and the camera record used.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gpu_renderer.h"

using sparkle_amd::GpuPathRenderer;
using sparkle_amd::RenderConfig;
using sparkle_amd::SceneDesc;

struct QuadMesh {
    std::vector<float> pos, nrm, tan;
    std::vector<uint32_t> idx;
};

static QuadMesh quad(const float p[4][3], const float n[3], const float t[3]) {
    QuadMesh q;
    for (int v = 0; v < 4; ++v) {
        q.pos.insert(q.pos.end(), p[v], p[v] + 3);
        q.nrm.insert(q.nrm.end(), n, n + 3);
        q.tan.insert(q.tan.end(), {t[0], t[1], t[2], 1.f});
    }
    q.idx = {0, 1, 2, 0, 2, 3};
    return q;
}

int main(int argc, char **argv) {
    const char *out = argc > 1 ? argv[1] : "render_demo.pfm";
    const uint32_t W = argc > 3 ? (uint32_t)atoi(argv[2]) : 128, H = argc > 3 ? (uint32_t)atoi(argv[3]) : 128;
    const int frames = argc > 4 ? atoi(argv[4]) : 16;

    // room [-1,1]^2 x [0,2], light quad below the ceiling (radiance 15), z up
    const float floor_p[4][3] = {{-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 1, 0}};
    const float ceil_p[4][3] = {{-1, -1, 2}, {-1, 1, 2}, {1, 1, 2}, {1, -1, 2}};
    const float back_p[4][3] = {{-1, 1, 0}, {1, 1, 0}, {1, 1, 2}, {-1, 1, 2}};
    const float left_p[4][3] = {{-1, -1, 0}, {-1, 1, 0}, {-1, 1, 2}, {-1, -1, 2}};
    const float right_p[4][3] = {{1, -1, 0}, {1, -1, 2}, {1, 1, 2}, {1, 1, 0}};
    const float light_p[4][3] = {{-0.5f, -0.5f, 1.99f}, {-0.5f, 0.5f, 1.99f}, {0.5f, 0.5f, 1.99f}, {0.5f, -0.5f, 1.99f}};
    const float up[3] = {0, 0, 1}, down[3] = {0, 0, -1}, toward[3] = {0, -1, 0}, px[3] = {1, 0, 0}, nx[3] = {-1, 0, 0};
    const float ty[3] = {0, 1, 0}, tx[3] = {1, 0, 0};
    std::vector<QuadMesh> quads = {quad(floor_p, up, tx),     quad(ceil_p, down, tx),   quad(back_p, toward, tx),
                                   quad(left_p, px, ty),      quad(right_p, nx, ty),    quad(light_p, down, tx)};
    SceneDesc scene;
    const float albedo[6][3] = {{0.7f, 0.7f, 0.7f}, {0.7f, 0.7f, 0.7f}, {0.7f, 0.7f, 0.7f},
                                {0.7f, 0.2f, 0.2f}, {0.2f, 0.6f, 0.2f}, {1.f, 1.f, 1.f}};
    for (size_t k = 0; k < quads.size(); ++k) {
        spk_mesh_desc m;
        std::memset(&m, 0, sizeof(m));
        m.positions = quads[k].pos.data();
        m.num_vertices = 4;
        m.indices = quads[k].idx.data();
        m.num_triangles = 2;
        m.normals = quads[k].nrm.data();
        m.tangents = quads[k].tan.data();
        scene.meshes.push_back(m);
        spk_material_desc mat;
        std::memset(&mat, 0, sizeof(mat));
        mat.type = SPK_MATERIAL_PBR;
        for (int c = 0; c < 3; ++c) mat.base_color[c] = albedo[k][c];
        if (k == 5)
            for (int c = 0; c < 3; ++c) mat.emissive[c] = 15.f;
        mat.metallic = 0.f;
        mat.roughness = 1.f;
        scene.materials.push_back(mat);
        spk_instance_desc ins;
        std::memset(&ins, 0, sizeof(ins));
        ins.mesh = (uint32_t)k;
        ins.material = (uint32_t)k;
        ins.world_from_local[0] = ins.world_from_local[5] = ins.world_from_local[10] = 1.f;
        scene.instances.push_back(ins);
    }
    // camera: CameraComponent defaults (35 mm, 24 mm sensor, f/22) looking along +y from outside the room
    spk_camera_desc cam;
    std::memset(&cam, 0, sizeof(cam));
    const float pos[3] = {0.f, -3.4f, 1.f}, right[3] = {1, 0, 0}, upv[3] = {0, 0, 1}, front[3] = {0, 1, 0};
    const float vfov = 2.f * std::atan(0.024f / (2.f * 0.035f)), focus = 3.4f;
    const float h = std::tan(vfov / 2.f), height = 2.f * h * focus, width = (float)W / (float)H * height;
    for (int a = 0; a < 3; ++a) {
        cam.position[a] = pos[a];
        cam.right[a] = right[a];
        cam.up[a] = upv[a];
        cam.front[a] = front[a];
        cam.max_u[a] = right[a] * width;
        cam.max_v[a] = upv[a] * height;
        cam.lower_left[a] = pos[a] + front[a] * focus - cam.max_u[a] * .5f - cam.max_v[a] * .5f;
    }
    cam.lens_radius = 0.035f / 22.f * 0.5f;
    cam.far_plane = 1000.f;

    RenderConfig cfg;
    cfg.width = W;
    cfg.height = H;
    cfg.max_sample_per_pixel = (uint32_t)frames;
    cfg.spatial_denoise = argc > 5 && atoi(argv[5]) != 0;
    GpuPathRenderer r(0);
    int rc = r.InitRenderResources(scene, cfg, cam);
    if (rc != SPK_OK) {
        std::fprintf(stderr, "init failed (%d): %s\n", rc, r.LastError().c_str());
        return 1;
    }
    double ms = 0.0;
    uint64_t rays = 0;
    for (int f = 0; f < frames; ++f) {
        if ((rc = r.Render()) != SPK_OK) {
            std::fprintf(stderr, "frame %d failed (%d): %s\n", f, rc, r.LastError().c_str());
            return 1;
        }
        ms += r.LastFrameMilliseconds();
        rays += r.LastFrameRays();
    }
    if (argc > 6) {
        // the accumulated frame buffer before tone mapping, row j = 0 (v = 0) first
        std::vector<float> acc;
        if ((rc = r.ReadbackFrameBuffer(acc)) != SPK_OK) return 1;
        FILE *fa = std::fopen(argv[6], "wb");
        if (!fa) return 1;
        std::fprintf(fa, "PF\n%u %u\n-1.0\n", W, H);
        for (size_t k = 0; k < (size_t)W * H; ++k) std::fwrite(&acc[4 * k], 4, 3, fa);
        std::fclose(fa);
    }
    std::vector<float> img;
    if ((rc = r.ReadbackImage(img)) != SPK_OK) return 1;
    FILE *fp = std::fopen(out, "wb");
    if (!fp) return 1;
    std::fprintf(fp, "PF\n%u %u\n-1.0\n", W, H);
    for (uint32_t j = 0; j < H; ++j)
        for (uint32_t i = 0; i < W; ++i) std::fwrite(&img[4 * ((size_t)(H - 1 - j) * W + i)], 4, 3, fp);
    std::fclose(fp);
    double mean = 0.0;
    for (size_t k = 0; k < img.size(); k += 4) mean += img[k] + img[k + 1] + img[k + 2];
    mean /= 3.0 * W * H;
    auto vec = [](const float *v) {
        char b[96];
        std::snprintf(b, sizeof(b), "[%.9g, %.9g, %.9g]", v[0], v[1], v[2]);
        return std::string(b);
    };
    std::printf("{\"frames\": %d, \"width\": %u, \"height\": %u, \"device_ms\": %.3f, \"rays\": %llu, "
                "\"mean\": %.6f, \"cumulated\": %u, \"camera\": {\"position\": %s, \"right\": %s, \"up\": %s, "
                "\"front\": %s, \"lower_left\": %s, \"max_u\": %s, \"max_v\": %s, \"lens_radius\": %.9g, "
                "\"far_plane\": %.9g}}\n",
                frames, W, H, ms, (unsigned long long)rays, mean, r.CumulatedSampleCount(), vec(cam.position).c_str(),
                vec(cam.right).c_str(), vec(cam.up).c_str(), vec(cam.front).c_str(), vec(cam.lower_left).c_str(),
                vec(cam.max_u).c_str(), vec(cam.max_v).c_str(), cam.lens_radius, cam.far_plane);
    return 0;
}
