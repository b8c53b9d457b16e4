// Device-side per-path sampling: xoshiro128+ per-pixel streams, camera rays, and the BxDFs of the
// reference's SamplePixel loop, written in the reference's order of float operations.
//   RNG           libraries/include/core/math/Sampler.h:16-53
//   camera ray    libraries/source/renderer/renderer/CPURenderer.cpp:220-238, 398-417
//   PBR surface   libraries/include/renderer/proxy/PbrMaterialRenderProxy.h:15-51
//   BxDFs         libraries/include/scene/material/BxDF.h:32-141
//   VNDF sample   libraries/include/core/math/Sampler.h:133-165
//   helpers       libraries/include/core/math/Utilities.h:201-451
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "spk_layout.h"

namespace spk {

struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 operator-(F3 a) { return F3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ F3 operator*(F3 a, float s) { return F3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ F3 operator*(float s, F3 a) { return F3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ F3 operator/(F3 a, float s) { return F3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ F3 cmul(F3 a, F3 b) { return F3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
    return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float sqnorm(F3 a) { return dot(a, a); }
__device__ __forceinline__ F3 normalized(F3 a) {
    const float n = sqnorm(a);
    if (n > 0.f) {
        const float s = sqrtf(n);
        return F3{a.x / s, a.y / s, a.z / s};
    }
    return a;
}
__device__ __forceinline__ F3 ld3(const float *p) { return F3{p[0], p[1], p[2]}; }

#define SPK_PI 3.14159265358979323846f
#define SPK_EPS 1e-6f
#define SPK_TOLERANCE 1e-4f
#define SPK_OUTPUT_LIMIT 6.f

// xoshiro128+ 1.0 (Blackman & Vigna) seeded by SplitMix64; RandomUnit() = float(u32) / 2^32
struct Rng {
    uint32_t s0, s1, s2, s3;
};
__device__ __forceinline__ Rng rng_seed(uint64_t seed) {
    uint64_t x = seed;
    uint32_t s[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        x += 0x9e3779b97f4a7c15ULL;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        z = z ^ (z >> 31);
        s[k] = (uint32_t)z;
    }
    return Rng{s[0], s[1], s[2], s[3]};
}
__device__ __forceinline__ uint32_t rng_next(Rng &r) {
    const uint32_t result = r.s0 + r.s3;
    const uint32_t t = r.s1 << 9;
    r.s2 ^= r.s0;
    r.s3 ^= r.s1;
    r.s1 ^= r.s2;
    r.s0 ^= r.s3;
    r.s2 ^= t;
    r.s3 = (r.s3 << 11) | (r.s3 >> 21);
    return result;
}
__device__ __forceinline__ float rng_unit(Rng &r) { return (float)rng_next(r) * 2.3283064365386963e-10f; }
__device__ __forceinline__ void rng_disk(Rng &r, float &x, float &y) {
    const float rnd = rng_unit(r);
    const float theta = rng_unit(r) * 2.0f * SPK_PI;
    const float rr = sqrtf(rnd);
    x = rr * cosf(theta);
    y = rr * sinf(theta);
}

__device__ __forceinline__ void local_axes(F3 n, F3 major, F3 &u, F3 &v, F3 &w) {
    u = normalized(cross(n, major));
    v = cross(n, u);
    w = n;
}

__device__ __forceinline__ float schlick1(float c, float r0) { return r0 + (1.f - r0) * powf((1.f - c), 5.f); }
__device__ __forceinline__ F3 schlick3(float c, F3 r0) {
    const float p = powf((1.f - c), 5.f);
    return r0 + (f3(1.f, 1.f, 1.f) - r0) * p;
}

__device__ __forceinline__ float smith_ggx_correlated(float cos_o, float cos_i, float roughness) {
    const float a = roughness * roughness;
    const float a2 = a * a;
    const float ggx_i = cos_o * sqrtf((-cos_i * a2 + cos_i) * cos_i + a2);
    const float ggx_o = cos_i * sqrtf((-cos_o * a2 + cos_o) * cos_o + a2);
    return 2.f * cos_o * cos_i / (ggx_o + ggx_i);
}
__device__ __forceinline__ float geometry_schlick_ggx(float cos_theta, float roughness) {
    const float a = roughness * roughness;
    const float a2 = a * a;
    const float tan2_v = (1.f - cos_theta * cos_theta) / (cos_theta * cos_theta);
    return 2.f / (1.f + sqrtf(1.f + a2 * tan2_v));
}

__device__ __forceinline__ F3 sample_vndf(Rng &rng, F3 w_o, float roughness) {
    const float a = roughness * roughness;
    const float u1 = rng_unit(rng);
    const float u2 = rng_unit(rng);
    const F3 v = normalized(f3(w_o.x * a, w_o.y * a, w_o.z));
    const F3 t1 = (v.z < 0.999f) ? normalized(cross(v, f3(0.f, 0.f, 1.f))) : f3(1.f, 0.f, 0.f);
    const F3 t2 = cross(v, t1);
    const float r = sqrtf(u1);
    const float phi = 2.f * SPK_PI * u2;
    const float p1 = r * cosf(phi);
    float p2 = r * sinf(phi);
    const float s = 0.5f * (1.f + v.z);
    p2 = (1.f - s) * sqrtf(1.f - p1 * p1) + s * p2;
    const F3 n = p1 * t1 + p2 * t2 + sqrtf(fmaxf(0.f, 1.f - p1 * p1 - p2 * p2)) * v;
    return normalized(f3(a * n.x, a * n.y, fmaxf(0.f, n.z)));
}

struct Sampled {
    bool valid;
    bool diffuse;  // the Lambertian fallback lobe was taken
    F3 throughput, wi;
};

__device__ __forceinline__ Sampled specular_sample(Rng &rng, F3 wo, F3 base, float roughness, float metallic) {
    Sampled s;
    s.valid = false;
    s.diffuse = false;
    const F3 wm = sample_vndf(rng, wo, roughness);
    const F3 wi = wm * (dot(wo, wm) * 2.f) - wo;
    const float cos_o = fmaxf(fabsf(wo.z), 0.f);
    const float cos_i = fmaxf(fabsf(wi.z), 0.f);
    const float cos_m = fmaxf(fabsf(wm.z), 0.f);
    const float local_cos_i = dot(wi, wm);
    F3 fres = f3(0.f, 0.f, 0.f);
    if (cos_o > SPK_EPS && cos_i > SPK_EPS && cos_m > SPK_EPS) {
        if (rng_unit(rng) < metallic) {
            fres = schlick3(local_cos_i, base);
            s.valid = true;
        } else {
            const float f = schlick1(local_cos_i, 0.04f);
            if (rng_unit(rng) < f) {
                fres = f3(1.f, 1.f, 1.f);
                s.valid = true;
            }
        }
    }
    if (!s.valid) return s;
    const float occlusion = smith_ggx_correlated(cos_o, cos_i, roughness);
    const float normalizer = geometry_schlick_ggx(cos_o, roughness) + SPK_EPS;
    s.throughput = fres * occlusion / normalizer;
    s.wi = wi;
    return s;
}

__device__ __forceinline__ Sampled lambert_sample(Rng &rng, F3 base) {
    Sampled s;
    s.valid = true;
    s.diffuse = true;
    float x, y;
    rng_disk(rng, x, y);
    const float z = sqrtf(fmaxf(0.f, 1.0f - (x * x + y * y)));
    s.wi = f3(x, y, z);
    s.throughput = base;
    return s;
}

__device__ __forceinline__ float fr_dielectric(float cos_i, float eta_i, float eta_t) {
    if (!(cos_i > 0.f)) {
        const float t = eta_i;
        eta_i = eta_t;
        eta_t = t;
        cos_i = -cos_i;
    }
    const float sin_i = sqrtf(1 - cos_i * cos_i);
    const float ref_idx = eta_i / eta_t;
    const float sin_t = ref_idx * sin_i;
    if (sin_t >= 1.f) return 1.f;
    float r0 = (1.f - ref_idx) / (1.f + ref_idx);
    r0 = r0 * r0;
    return schlick1(cos_i, r0);
}

__device__ __forceinline__ Sampled dielectric_sample(Rng &rng, F3 wo, F3 base, float eta) {
    Sampled s;
    s.valid = true;
    s.diffuse = false;
    const float fres = fr_dielectric(wo.z, 1.0f, eta);
    if (rng_unit(rng) < fres) {
        s.wi = f3(-wo.x, -wo.y, wo.z);
    } else {
        const bool entering = wo.z > 0;
        const float e = entering ? 1.0f / eta : eta / 1.0f;
        const float cos_i = wo.z;
        const float sin_i2 = 1.f - cos_i * cos_i;
        const float sin_t2 = e * e * sin_i2;
        const float cos_t = sqrtf(1 - sin_t2) * (cos_i > 0 ? -1.f : 1.f);
        s.wi = normalized(f3(-e * wo.x, -e * wo.y, cos_t));
    }
    s.throughput = base;
    return s;
}

// ---- textures: Image::Sample (io/Image.h:112-152) ----
__device__ __forceinline__ uint32_t wrap_mod(int a, uint32_t b) {
    return ((uint32_t)(a % (int)b) + b) % b;  // Utilities.h:526-529
}

__device__ __forceinline__ float srgb_to_linear(float c) {
    // Utilities.h:416-425, per channel
    const float cutoff = c < 0.04045f ? 1.f : 0.f;
    const float higher = powf((c + 0.055f) / 1.055f, 2.4f);
    const float lower = c / 12.92f;
    return higher * (1.f - cutoff) + lower * cutoff;
}

// bilinear with wrapping; uv scales by (width - 1, height - 1) (Image.h:21, 123) and the integer
// part is a truncation (Utilities.h:124-128)
__device__ __forceinline__ F3 tex_sample(const TexRecord *__restrict__ table, const float4 *__restrict__ texels,
                                         int index, float u, float v) {
    const TexRecord t = table[index];
    const float px = u * (float)(t.width - 1u), py = v * (float)(t.height - 1u);
    const int iu = (int)px, iv = (int)py;
    const float fu = px - (float)iu, fv = py - (float)iv;
    const uint32_t x0 = wrap_mod(iu, t.width), x1 = wrap_mod(iu + 1, t.width);
    const uint32_t y0 = wrap_mod(iv, t.height), y1 = wrap_mod(iv + 1, t.height);
    const float4 *T = texels + t.offset;
    const float4 s00 = T[(size_t)y0 * t.width + x0], s10 = T[(size_t)y0 * t.width + x1];
    const float4 s01 = T[(size_t)y1 * t.width + x0], s11 = T[(size_t)y1 * t.width + x1];
    const F3 l0 = f3(s00.x + (s10.x - s00.x) * fu, s00.y + (s10.y - s00.y) * fu, s00.z + (s10.z - s00.z) * fu);
    const F3 l1 = f3(s01.x + (s11.x - s01.x) * fu, s01.y + (s11.y - s01.y) * fu, s01.z + (s11.z - s01.z) * fu);
    F3 r = f3(l0.x + (l1.x - l0.x) * fv, l0.y + (l1.y - l0.y) * fv, l0.z + (l1.z - l0.z) * fv);
    if (t.srgb) r = f3(srgb_to_linear(r.x), srgb_to_linear(r.y), srgb_to_linear(r.z));
    return r;
}

// MaterialRenderProxy::SampleSurface for the PBR and dielectric proxies; w_i in world space. base,
// metallic and roughness are the material getters' values at the hit's uv (MaterialRenderProxy.h:117-144).
__device__ __forceinline__ F3 sample_surface(Rng &rng, const MaterialGpu &mat, F3 base, float metallic,
                                             float roughness, F3 ray_dir, F3 normal, F3 tangent, F3 &w_i,
                                             bool &diffuse) {
    const F3 wo = -ray_dir;
    F3 u, v, w;
    local_axes(normal, tangent, u, v, w);
    const F3 lwo = f3(dot(wo, u), dot(wo, v), dot(wo, w));
    Sampled s;
    if (mat.type == 1u) {
        s = dielectric_sample(rng, lwo, base, mat.eta);
    } else {
        s = specular_sample(rng, lwo, base, roughness, metallic);
        if (!s.valid) s = lambert_sample(rng, base);
    }
    F3 u2, v2, w2;
    local_axes(normal, tangent, u2, v2, w2);
    w_i = normalized(u2 * s.wi.x + v2 * s.wi.y + w2 * s.wi.z);
    diffuse = s.diffuse;
    return s.throughput;
}

}  // namespace spk
