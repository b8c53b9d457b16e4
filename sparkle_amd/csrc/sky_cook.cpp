// Equirectangular sky map -> six cube-map faces, the host-side cook of a textured sky light
// (SkyLightCookJob::Execute, libraries/source/scene/component/light/SkyLight.cpp:83-135). The path
// tracer samples the faces per missed ray (spk_set_sky_cubemap); the cook runs once per map.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/spk_pt.h"

namespace {

// float -> IEEE binary16 -> float, round to nearest even (the cube faces are RGBAFloat16)
float round_half(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = x & 0x80000000u;
    const uint32_t ax = x & 0x7FFFFFFFu;
    uint32_t h;  // magnitude bits of the half
    if (ax >= 0x7F800000u) {
        h = ax > 0x7F800000u ? 0x7E00u : 0x7C00u;  // NaN stays NaN, inf stays inf
    } else if (ax >= 0x477FF000u) {
        h = 0x7C00u;  // rounds past 65504: inf
    } else if (ax >= 0x38800000u) {  // normal half
        const uint32_t mant = ax & 0x7FFFFFu;
        const uint32_t exp = (ax >> 23) - 112u;
        h = (exp << 10) | (mant >> 13);
        const uint32_t rest = mant & 0x1FFFu;
        if (rest > 0x1000u || (rest == 0x1000u && (h & 1u))) ++h;  // carries into the exponent correctly
    } else if (ax >= 0x33000000u) {  // subnormal half (or rounds up to the smallest)
        const uint32_t e = ax >> 23;
        const uint32_t mant = (ax & 0x7FFFFFu) | 0x800000u;
        const uint32_t shift = 126u - e;  // 14..24
        h = mant >> shift;
        const uint32_t rest = mant & ((1u << shift) - 1u);
        const uint32_t halfway = 1u << (shift - 1u);
        if (rest > halfway || (rest == halfway && (h & 1u))) ++h;
    } else {
        h = 0u;
    }
    // back to float
    uint32_t out;
    const uint32_t he = (h >> 10) & 0x1Fu, hm = h & 0x3FFu;
    if (he == 0x1Fu) {
        out = sign | 0x7F800000u | (hm << 13);
    } else if (he != 0u) {
        out = sign | ((he + 112u) << 23) | (hm << 13);
    } else if (hm != 0u) {
        float v = std::ldexp(static_cast<float>(hm), -24);
        std::memcpy(&out, &v, 4);
        out |= sign;
    } else {
        out = sign;
    }
    float r;
    std::memcpy(&r, &out, 4);
    return r;
}

inline unsigned wrap_mod(int a, unsigned b) { return (static_cast<unsigned>(a % static_cast<int>(b)) + b) % b; }

// Image::Sample (io/Image.h:112-152) of an RGBA float image: uv scaled by (w - 1, h - 1), the cell
// split by truncation, bilinear with wrapping
void sample_rgb(const float *img, uint32_t w, uint32_t h, float u, float v, float out[3]) {
    const float px = u * static_cast<float>(w - 1), py = v * static_cast<float>(h - 1);
    const int iu = static_cast<int>(px), iv = static_cast<int>(py);
    const float fu = px - static_cast<float>(iu), fv = py - static_cast<float>(iv);
    const float *s00 = img + 4 * ((size_t)wrap_mod(iv, h) * w + wrap_mod(iu, w));
    const float *s10 = img + 4 * ((size_t)wrap_mod(iv, h) * w + wrap_mod(iu + 1, w));
    const float *s01 = img + 4 * ((size_t)wrap_mod(iv + 1, h) * w + wrap_mod(iu, w));
    const float *s11 = img + 4 * ((size_t)wrap_mod(iv + 1, h) * w + wrap_mod(iu + 1, w));
    for (int c = 0; c < 3; ++c) {
        const float l0 = s00[c] + (s10[c] - s00[c]) * fu;
        const float l1 = s01[c] + (s11[c] - s01[c]) * fu;
        out[c] = l0 + (l1 - l0) * fv;
    }
}

// Image2DCube::TextureCoordinateToDirection (io/Image.h:287-315), normalised
void face_direction(int face, float u, float v, float d[3]) {
    switch (face) {
        case 0: d[0] = 1.f, d[1] = -v, d[2] = -u; break;   // +X
        case 1: d[0] = -1.f, d[1] = -v, d[2] = u; break;   // -X
        case 2: d[0] = u, d[1] = 1.f, d[2] = v; break;     // +Y
        case 3: d[0] = u, d[1] = -1.f, d[2] = -v; break;   // -Y
        case 4: d[0] = u, d[1] = -v, d[2] = 1.f; break;    // +Z
        default: d[0] = -u, d[1] = -v, d[2] = -1.f; break; // -Z
    }
    const float n = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const float s = std::sqrt(n);
    d[0] = d[0] / s, d[1] = d[1] / s, d[2] = d[2] / s;
}

}  // namespace

extern "C" int spk_sky_cook(const float *equirect, uint32_t width, uint32_t height, uint32_t face_size,
                            float *faces) {
    if (!equirect || !faces || width < 2 || height < 2 || face_size == 0 || face_size > 16384)
        return SPK_ERR_INVALID_ARGUMENT;
    const float inv_2pi = 0.15915494309189533577f, inv_pi = 0.31830988618379067154f;  // InvPi * 0.5, InvPi
    const uint64_t rows = 6ull * face_size;
    const unsigned nt = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    auto work = [&](unsigned t) {
        for (uint64_t r = t; r < rows; r += nt) {
            const int face = static_cast<int>(r / face_size);
            const uint32_t i = static_cast<uint32_t>(r % face_size);  // texel column (SetPixel(i, j))
            const float u = (static_cast<float>(i) + 0.5f) / static_cast<float>(face_size) * 2.f - 1.f;
            for (uint32_t j = 0; j < face_size; ++j) {
                const float v = (static_cast<float>(j) + 0.5f) / static_cast<float>(face_size) * 2.f - 1.f;
                float d[3];
                face_direction(face, u, v, d);
                // CartesianToEquirectangular (Utilities.h:510-518)
                const float eu = std::atan2(d[1], d[0]) * inv_2pi + 0.5f;
                const float ev = std::asin(-d[2]) * inv_pi + 0.5f;
                float c[3];
                sample_rgb(equirect, width, height, eu, ev, c);
                float *o = faces + 4 * (((size_t)face * face_size + j) * face_size + i);
                o[0] = round_half(c[0]);
                o[1] = round_half(c[1]);
                o[2] = round_half(c[2]);
                o[3] = 1.f;
            }
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
    return SPK_OK;
}
