// Device-side closest-hit traversal of the quantised wide BVH, shared by the closest-hit kernel
// and the path tracer's extend stage.
//
// Replaces SceneRenderProxy::Intersect<false> -> TLAS/BLAS bvh::v2 traversal ->
// MeshRenderProxy::IntersectTriangle (SceneRenderProxy.cpp:82-126, MeshRenderProxy.cpp:94-194)
// for flattened world-space geometry, with two deliberate differences:
//   * the triangle test is the watertight FP32 test of Woop, Benthin and Wald (JCGT 2013)
//     instead of Moller-Trumbore with an absolute |det| < 1e-6 cutoff, which rejects every cell of
//     a 1024^2 heightfield over the unit square;
//   * exact-t ties go to the smaller global triangle id rather than to the first leaf visited,
//     so the result does not depend on tree shape or visit order.
// The arithmetic of the triangle test and of the final hit distance is written operation for
// operation like the CPU oracle and compiled without FMA contraction, so hit ids and t are
// bit-identical to it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "spk_layout.h"

namespace spk {

struct WtSetup {
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

__device__ __forceinline__ float vsel(int k, float x, float y, float z) { return k == 0 ? x : (k == 1 ? y : z); }

__device__ __forceinline__ WtSetup wt_setup(float dx, float dy, float dz) {
    WtSetup r;
    const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    int kz = 0;
    if (ay > ax) kz = 1;
    if (az > fmaxf(ax, ay)) kz = 2;
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    const float dkz = vsel(kz, dx, dy, dz);
    if (dkz < 0.f) {
        const int t = kx;
        kx = ky;
        ky = t;
    }
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    r.Sx = vsel(kx, dx, dy, dz) / dkz;
    r.Sy = vsel(ky, dx, dy, dz) / dkz;
    r.Sz = 1.0f / dkz;
    return r;
}

// Ray translated and sheared so that the test per triangle needs only selects, no divisions
// until a hit is accepted.
struct TriTest {
    bool hit;
    float t, U, V, W, det;
};

__device__ TriTest tri_wt_precise(WtSetup r, float ox, float oy, float oz, const float4 a, const float4 b,
                                  const float4 c);

#define SPK_WT_ERR (8.f * 5.9604644775390625e-08f)     /* 8u, u = 2^-24 */
#define SPK_WT_ERR_ABS (2.f * 5.9604644775390625e-08f) /* 2u */

// The FP32 edge functions are trusted only when a forward error bound certifies their signs; an
// uncertified triangle (ray within rounding distance of an edge, or the triangle seen edge-on) is
// re-evaluated in double-float arithmetic. This takes the place of the paper's FP64 fallback and
// removes the false hits that exact-zero edge values would otherwise produce.
__device__ __forceinline__ TriTest tri_wt(const WtSetup &r, float ox, float oy, float oz, const float4 a,
                                          const float4 b, const float4 c) {
    TriTest res;
    res.hit = false;
    const float Ax0 = a.x - ox, Ay0 = a.y - oy, Az0 = a.z - oz;
    const float Bx0 = b.x - ox, By0 = b.y - oy, Bz0 = b.z - oz;
    const float Cx0 = c.x - ox, Cy0 = c.y - oy, Cz0 = c.z - oz;
    const float Akz = vsel(r.kz, Ax0, Ay0, Az0), Bkz = vsel(r.kz, Bx0, By0, Bz0), Ckz = vsel(r.kz, Cx0, Cy0, Cz0);
    const float Akx = vsel(r.kx, Ax0, Ay0, Az0), Aky = vsel(r.ky, Ax0, Ay0, Az0);
    const float Bkx = vsel(r.kx, Bx0, By0, Bz0), Bky = vsel(r.ky, Bx0, By0, Bz0);
    const float Ckx = vsel(r.kx, Cx0, Cy0, Cz0), Cky = vsel(r.ky, Cx0, Cy0, Cz0);
    const float sAx = r.Sx * Akz, sAy = r.Sy * Akz;
    const float sBx = r.Sx * Bkz, sBy = r.Sy * Bkz;
    const float sCx = r.Sx * Ckz, sCy = r.Sy * Ckz;
    const float Ax = Akx - sAx;
    const float Ay = Aky - sAy;
    const float Bx = Bkx - sBx;
    const float By = Bky - sBy;
    const float Cx = Ckx - sCx;
    const float Cy = Cky - sCy;
    const float U = Cx * By - Cy * Bx;
    const float V = Ax * Cy - Ay * Cx;
    const float W = Bx * Ay - By * Ax;
    const float mAx = fabsf(Akx) + fabsf(sAx), mAy = fabsf(Aky) + fabsf(sAy);
    const float mBx = fabsf(Bkx) + fabsf(sBx), mBy = fabsf(Bky) + fabsf(sBy);
    const float mCx = fabsf(Ckx) + fabsf(sCx), mCy = fabsf(Cky) + fabsf(sCy);
    const float eU = SPK_WT_ERR * (mCx * mBy + mCy * mBx) + SPK_WT_ERR_ABS * fabsf(U);
    const float eV = SPK_WT_ERR * (mAx * mCy + mAy * mCx) + SPK_WT_ERR_ABS * fabsf(V);
    const float eW = SPK_WT_ERR * (mBx * mAy + mBy * mAx) + SPK_WT_ERR_ABS * fabsf(W);
    // reject quickly when two certified values already disagree in sign
    const bool cU = fabsf(U) > eU, cV = fabsf(V) > eV, cW = fabsf(W) > eW;
    if (!(cU && cV && cW)) {
        const bool mixed = (cU && cV && ((U < 0.f) != (V < 0.f))) || (cU && cW && ((U < 0.f) != (W < 0.f))) ||
                           (cV && cW && ((V < 0.f) != (W < 0.f)));
        if (mixed) return res;
        return tri_wt_precise(r, ox, oy, oz, a, b, c);
    }
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return res;
    const float det = U + V + W;
    if (det == 0.f) return res;
    const float Az = r.Sz * Akz;
    const float Bz = r.Sz * Bkz;
    const float Cz = r.Sz * Ckz;
    const float T = U * Az + V * Bz + W * Cz;
    if ((det < 0.f && T > 0.f) || (det > 0.f && T < 0.f)) return res;
    res.hit = true;
    res.t = T / det;
    res.U = U;
    res.V = V;
    res.W = W;
    res.det = det;
    return res;
}

// Hit distance of the selected triangle, n.(v0 - o) / n.d, in double-float arithmetic (pairs
// hi + lo from TwoSum and FMA-based TwoProd, Dekker 1971): edges and v0 - o are exact, the normal
// and both dot products carry ~44 bits, and only the final division rounds to FP32.
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ void two_prod(float a, float b, float &p, float &e) {
    p = a * b;
    e = __builtin_fmaf(a, b, -p);
}
struct DD {
    float h, l;
};
__device__ __forceinline__ DD dd_diff(float a, float b) {
    DD r;
    two_sum(a, -b, r.h, r.l);
    return r;
}
__device__ __forceinline__ DD dd_mul(DD a, DD b) {
    float p, e;
    two_prod(a.h, b.h, p, e);
    e = e + (a.h * b.l + a.l * b.h);
    DD r;
    r.h = p + e;
    r.l = e - (r.h - p);
    return r;
}
__device__ __forceinline__ DD dd_add(DD a, DD b) {
    float s, e;
    two_sum(a.h, b.h, s, e);
    e = e + (a.l + b.l);
    DD r;
    two_sum(s, e, r.h, r.l);
    return r;
}
__device__ __forceinline__ DD dd_neg(DD a) { return DD{-a.h, -a.l}; }
__device__ __forceinline__ DD dd_sel(int k, DD x, DD y, DD z) { return k == 0 ? x : (k == 1 ? y : z); }

// tri_wt evaluated in double-float arithmetic: exact translation, shear and edge functions carried
// to ~44 bits, signs decided on the rounded high parts. Taken only for uncertified triangles.
__device__ __noinline__ TriTest tri_wt_precise(WtSetup r, float ox, float oy, float oz, const float4 a,
                                               const float4 b, const float4 c) {
    TriTest res;
    res.hit = false;
    const DD A0 = dd_diff(a.x, ox), A1 = dd_diff(a.y, oy), A2 = dd_diff(a.z, oz);
    const DD B0 = dd_diff(b.x, ox), B1 = dd_diff(b.y, oy), B2 = dd_diff(b.z, oz);
    const DD C0 = dd_diff(c.x, ox), C1 = dd_diff(c.y, oy), C2 = dd_diff(c.z, oz);
    const DD Akz = dd_sel(r.kz, A0, A1, A2), Bkz = dd_sel(r.kz, B0, B1, B2), Ckz = dd_sel(r.kz, C0, C1, C2);
    const DD Sx{r.Sx, 0.f}, Sy{r.Sy, 0.f}, Sz{r.Sz, 0.f};
    const DD Ax = dd_add(dd_sel(r.kx, A0, A1, A2), dd_neg(dd_mul(Sx, Akz)));
    const DD Ay = dd_add(dd_sel(r.ky, A0, A1, A2), dd_neg(dd_mul(Sy, Akz)));
    const DD Bx = dd_add(dd_sel(r.kx, B0, B1, B2), dd_neg(dd_mul(Sx, Bkz)));
    const DD By = dd_add(dd_sel(r.ky, B0, B1, B2), dd_neg(dd_mul(Sy, Bkz)));
    const DD Cx = dd_add(dd_sel(r.kx, C0, C1, C2), dd_neg(dd_mul(Sx, Ckz)));
    const DD Cy = dd_add(dd_sel(r.ky, C0, C1, C2), dd_neg(dd_mul(Sy, Ckz)));
    const DD U = dd_add(dd_mul(Cx, By), dd_neg(dd_mul(Cy, Bx)));
    const DD V = dd_add(dd_mul(Ax, Cy), dd_neg(dd_mul(Ay, Cx)));
    const DD W = dd_add(dd_mul(Bx, Ay), dd_neg(dd_mul(By, Ax)));
    if ((U.h < 0.f || V.h < 0.f || W.h < 0.f) && (U.h > 0.f || V.h > 0.f || W.h > 0.f)) return res;
    const DD det = dd_add(dd_add(U, V), W);
    if (det.h == 0.f) return res;
    const DD T =
        dd_add(dd_add(dd_mul(U, dd_mul(Sz, Akz)), dd_mul(V, dd_mul(Sz, Bkz))), dd_mul(W, dd_mul(Sz, Ckz)));
    if ((det.h < 0.f && T.h > 0.f) || (det.h > 0.f && T.h < 0.f)) return res;
    const float detf = det.h + det.l;
    res.hit = true;
    res.t = (T.h + T.l) / detf;
    res.U = U.h;
    res.V = V.h;
    res.W = W.h;
    res.det = detf;
    return res;
}
__device__ __forceinline__ float refine_t(float ox, float oy, float oz, float dx, float dy, float dz, const float4 v0,
                                          const float4 v1, const float4 v2, float t_fallback) {
    const DD e1x = dd_diff(v1.x, v0.x), e1y = dd_diff(v1.y, v0.y), e1z = dd_diff(v1.z, v0.z);
    const DD e2x = dd_diff(v2.x, v0.x), e2y = dd_diff(v2.y, v0.y), e2z = dd_diff(v2.z, v0.z);
    const DD ax = dd_diff(v0.x, ox), ay = dd_diff(v0.y, oy), az = dd_diff(v0.z, oz);
    const DD nx = dd_add(dd_mul(e1y, e2z), dd_neg(dd_mul(e1z, e2y)));
    const DD ny = dd_add(dd_mul(e1z, e2x), dd_neg(dd_mul(e1x, e2z)));
    const DD nz = dd_add(dd_mul(e1x, e2y), dd_neg(dd_mul(e1y, e2x)));
    const DD num = dd_add(dd_add(dd_mul(nx, ax), dd_mul(ny, ay)), dd_mul(nz, az));
    const DD den = dd_add(dd_add(dd_mul(nx, DD{dx, 0.f}), dd_mul(ny, DD{dy, 0.f})), dd_mul(nz, DD{dz, 0.f}));
    const float numf = num.h + num.l, denf = den.h + den.l;
    if (denf == 0.f) return t_fallback;
    const float t = numf / denf;
    if (!(t > 0.f) || !isfinite(t)) return t_fallback;
    return t;
}

__device__ __forceinline__ float exp2i(int e) { return __int_as_float((e + 127) << 23); }
__device__ __forceinline__ float ubyte(uint32_t w, int k) {
    return (float)((w >> (8 * k)) & 0xffu);  // v_cvt_f32_ubyte{k}
}

// Slack factor of the slab test: covers the rounding of t = q*scale + (p - o)*invd at any
// distance; the builder adds one quantisation step of slack in space on each side of every box.
#define SPK_BOX_SLACK 1.00000095367431640625f /* 1 + 2^-20 */

struct RayIn {
    float ox, oy, oz, tmin, dx, dy, dz, tmax;
};

struct BestHit {
    float t;           // Woop t of the best hit so far (selection and culling)
    uint32_t prim;     // global triangle id, 0xFFFFFFFF when none
    uint32_t tri;      // leaf-order index of the best triangle
    float U, V, W, det;
    uint32_t inst;     // two-level scenes: instance record of the best hit
};

// ---- two-level scenes: a TLAS over instances, one BLAS per mesh in object space ----
// Instance record (64 bytes, TLAS leaf order): rows 0-2 = object_from_world (3x4, float bits),
// row 3 = (BLAS root node, first global triangle id of the instance, instance index, 0).
// A TLAS leaf slot holds one instance; the ray enters its BLAS transformed to object space with an
// unnormalised direction, so the ray parameter t, box culling and the tie rule carry over.
#define SPK_INST_ENTRY 0x80000000u  // stack entry flag: pending instances {leaf base, bits}

__device__ __forceinline__ void object_ray(const uint4 r0, const uint4 r1, const uint4 r2, float wox, float woy,
                                           float woz, float wdx, float wdy, float wdz, float &ox, float &oy,
                                           float &oz, float &dx, float &dy, float &dz) {
    const float a00 = __uint_as_float(r0.x), a01 = __uint_as_float(r0.y), a02 = __uint_as_float(r0.z),
                a03 = __uint_as_float(r0.w);
    const float a10 = __uint_as_float(r1.x), a11 = __uint_as_float(r1.y), a12 = __uint_as_float(r1.z),
                a13 = __uint_as_float(r1.w);
    const float a20 = __uint_as_float(r2.x), a21 = __uint_as_float(r2.y), a22 = __uint_as_float(r2.z),
                a23 = __uint_as_float(r2.w);
    ox = ((a00 * wox + a01 * woy) + a02 * woz) + a03;
    oy = ((a10 * wox + a11 * woy) + a12 * woz) + a13;
    oz = ((a20 * wox + a21 * woy) + a22 * woz) + a23;
    dx = (a00 * wdx + a01 * wdy) + a02 * wdz;
    dy = (a10 * wdx + a11 * wdy) + a12 * wdz;
    dz = (a20 * wdx + a21 * wdy) + a22 * wdz;
}

struct TraverseCounters {
    uint32_t nodes;
    uint32_t tris;
};

__device__ __forceinline__ bool ray_valid(const RayIn &r) {
    const bool finite = isfinite(r.ox) && isfinite(r.oy) && isfinite(r.oz) && isfinite(r.dx) && isfinite(r.dy) &&
                        isfinite(r.dz) && !isnan(r.tmin) && !isnan(r.tmax);
    return finite && (r.dx != 0.f || r.dy != 0.f || r.dz != 0.f);
}

__device__ __forceinline__ float safe_inv(float d) {
    const float dd = fabsf(d) < 1e-20f ? copysignf(1e-20f, d) : d;
    return __builtin_amdgcn_rcpf(dd);
}

// Closest-hit traversal state of one ray over a wide BVH. STACK entries per lane live in LDS at
// stack[k * stride]; each entry is a node group {child_base, imask << 8 | pending-hit bits}.
// step() pops a group when the current one is exhausted, visits one node (all child boxes) and
// tests that node's leaf triangles; traverse() runs steps until the ray is done, and the
// persistent closest-hit kernel interleaves steps with refilling finished lanes.
// TWO: the tree rooted at `root` is a TLAS whose leaf slots are instances (records in `insts`).
// Reaching TLAS leaves pushes the node's remaining group and one entry for its other instances,
// then enters the first instance's BLAS with the object-space ray; when the stack is back at its
// depth on entry the BLAS is done and the world ray is restored.
template <int WIDTH, int STACK, bool COUNT, bool ANYHIT = false, bool TWO = false>
struct Traversal {
    float ox, oy, oz, ix, iy, iz, box_tmin;
    WtSetup wt;
    uint32_t oct;
    // current node group: slot s of the group's node was hit iff bit (s ^ g_oct) of g_hits is set
    uint32_t g_base, g_imask, g_hits, g_oct;
    int sp;
    BestHit best;
    TraverseCounters cnt;
    // two-level state: the world ray, the level (0 TLAS, 1 BLAS), the stack depth on entering the
    // BLAS, the instance's first global triangle id and its record
    float wox, woy, woz, wdx, wdy, wdz;
    uint32_t level, blas_sp, id_base, cur;
    const uint4 *insts;

    __device__ __forceinline__ void set_ray(float px, float py, float pz, float dx, float dy, float dz) {
        ox = px, oy = py, oz = pz;
        wt = wt_setup(dx, dy, dz);
        ix = safe_inv(dx), iy = safe_inv(dy), iz = safe_inv(dz);
        oct = (dx < 0.f ? 1u : 0u) | (dy < 0.f ? 2u : 0u) | (dz < 0.f ? 4u : 0u);
    }

    __device__ __forceinline__ void enter(uint32_t slot) {
        const uint4 *ip = insts + (size_t)slot * 4u;
        const uint4 r0 = ip[0], r1 = ip[1], r2 = ip[2], r3 = ip[3];
        float px, py, pz, dx, dy, dz;
        object_ray(r0, r1, r2, wox, woy, woz, wdx, wdy, wdz, px, py, pz, dx, dy, dz);
        set_ray(px, py, pz, dx, dy, dz);
        level = 1u;
        blas_sp = (uint32_t)sp;
        id_base = r3.y;
        cur = slot;
        g_base = r3.x;
        g_imask = 1u;
        g_hits = WIDTH == 8 ? 1u << oct : 1u;
        g_oct = WIDTH == 8 ? oct : 0u;
    }

    __device__ __forceinline__ void leave() {
        set_ray(wox, woy, woz, wdx, wdy, wdz);
        level = 0u;
    }

    __device__ __forceinline__ void init(const RayIn &ray, uint32_t root, const uint4 *inst_records = nullptr) {
        set_ray(ray.ox, ray.oy, ray.oz, ray.dx, ray.dy, ray.dz);
        box_tmin = fmaxf(ray.tmin, 0.f);
        if (TWO) {
            wox = ray.ox, woy = ray.oy, woz = ray.oz, wdx = ray.dx, wdy = ray.dy, wdz = ray.dz;
            level = 0u;
            blas_sp = 0u;
            id_base = 0u;
            cur = 0u;
            insts = inst_records;
        }
        g_base = root;
        g_imask = 1u;
        g_hits = WIDTH == 8 ? 1u << oct : 1u;  // 4-wide: rank order is resolved per node below
        g_oct = WIDTH == 8 ? oct : 0u;
        sp = 0;
        // a ray with a NaN or infinite component or a zero direction hits nothing (the triangle test
        // rejects it everywhere); it ends here rather than walking the whole tree, which its NaN slab
        // distances would otherwise open (a refraction past the critical angle gives such a
        // direction when the Fresnel draw is exactly 1, BxDF.h:112-141)
        if (!ray_valid(ray)) g_hits = 0u;
        best.t = ray.tmax;
        best.prim = 0xFFFFFFFFu;
        best.tri = 0xFFFFFFFFu;
        best.U = best.V = best.W = 0.f;
        best.det = 1.f;
        best.inst = 0u;
        cnt.nodes = 0u;
        cnt.tris = 0u;
    }

    // Returns false once the ray is done (stack empty, or any hit for an occlusion query).
    __device__ __forceinline__ bool step(const uint4 *__restrict__ nodes, const uint4 *__restrict__ tris,
                                         uint2 *stack, int stride) {
        if (g_hits == 0u) {
            if (TWO && level != 0u && sp == (int)blas_sp) leave();  // the instance's BLAS is done
            if (sp == 0) return false;
            --sp;
            const uint2 e = stack[sp * stride];
            if (TWO && (e.y & SPK_INST_ENTRY)) {
                // pending instances of a TLAS node: enter the next one, keep the rest
                uint32_t bits = e.y & 0xFFFFFFu;
                const uint32_t k = __builtin_ctz(bits);
                bits &= bits - 1u;
                if (bits != 0u) {
                    stack[sp * stride] = make_uint2(e.x, SPK_INST_ENTRY | bits);
                    ++sp;
                }
                enter(e.x + k);
            } else {
                g_base = e.x;
                g_imask = (e.y >> 8) & 0xffu;
                g_hits = (e.y >> 16) & 0xffu;  // pushed entries always hold a pending hit
                g_oct = e.y & 0xffu;
            }
        }
        const uint32_t r = __builtin_ctz(g_hits);
        g_hits &= g_hits - 1u;
        const uint32_t slot = r ^ g_oct;
        const uint32_t child = g_base + __builtin_popcount(g_imask & ((1u << slot) - 1u));
        if (g_hits != 0u) {
            stack[sp * stride] = make_uint2(g_base, (g_imask << 8) | (g_hits << 16) | g_oct);
            ++sp;
        }
        if (COUNT) cnt.nodes++;

        const bool nxs = (oct & 1u) != 0u, nys = (oct & 2u) != 0u, nzs = (oct & 4u) != 0u;
        const float tcull = best.t * SPK_BOX_SLACK;
        uint32_t new_hits = 0u, tri_bits = 0u;
        uint32_t child_base, tri_base, imask, node_oct;
        if (WIDTH == 8) {
            const uint4 *np = nodes + (size_t)child * 5u;
            const uint4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3], n4 = np[4];
            const float px = __uint_as_float(n0.x), py = __uint_as_float(n0.y), pz = __uint_as_float(n0.z);
            const int ex = (int)(int8_t)(n0.w & 0xffu), ey = (int)(int8_t)((n0.w >> 8) & 0xffu),
                      ez = (int)(int8_t)((n0.w >> 16) & 0xffu);
            imask = n0.w >> 24;
            child_base = n1.x;
            tri_base = n1.y;
            const float sx = exp2i(ex) * ix, sy = exp2i(ey) * iy, sz = exp2i(ez) * iz;
            const float ax = (px - ox) * ix, ay = (py - oy) * iy, az = (pz - oz) * iz;
            // near planes are qlo for a positive direction, qhi for a negative one
            const uint32_t nx0 = nxs ? n3.z : n2.x, nx1 = nxs ? n3.w : n2.y;
            const uint32_t fx0 = nxs ? n2.x : n3.z, fx1 = nxs ? n2.y : n3.w;
            const uint32_t ny0 = nys ? n4.x : n2.z, ny1 = nys ? n4.y : n2.w;
            const uint32_t fy0 = nys ? n2.z : n4.x, fy1 = nys ? n2.w : n4.y;
            const uint32_t nz0 = nzs ? n4.z : n3.x, nz1 = nzs ? n4.w : n3.y;
            const uint32_t fz0 = nzs ? n3.x : n4.z, fz1 = nzs ? n3.y : n4.w;
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const int k = s & 3;
                const uint32_t wnx = s < 4 ? nx0 : nx1, wfx = s < 4 ? fx0 : fx1;
                const uint32_t wny = s < 4 ? ny0 : ny1, wfy = s < 4 ? fy0 : fy1;
                const uint32_t wnz = s < 4 ? nz0 : nz1, wfz = s < 4 ? fz0 : fz1;
                const float tnx = __builtin_fmaf(ubyte(wnx, k), sx, ax);
                const float tny = __builtin_fmaf(ubyte(wny, k), sy, ay);
                const float tnz = __builtin_fmaf(ubyte(wnz, k), sz, az);
                const float tfx = __builtin_fmaf(ubyte(wfx, k), sx, ax);
                const float tfy = __builtin_fmaf(ubyte(wfy, k), sy, ay);
                const float tfz = __builtin_fmaf(ubyte(wfz, k), sz, az);
                const float tnear = fmaxf(fmaxf(fmaxf(tnx, tny), tnz), box_tmin);
                const float tfar = fminf(fminf(fminf(tfx, tfy), tfz) * SPK_BOX_SLACK, tcull);
                const uint32_t meta = ((s < 4 ? n1.z : n1.w) >> (8 * k)) & 0xffu;
                const bool inner = (imask >> s) & 1u;
                const bool hit = tnear <= tfar;
                // selects, not branches: every lane evaluates both outcomes of every slot
                new_hits |= (hit && inner) ? (1u << (s ^ oct)) : 0u;
                tri_bits |= (hit && !inner) ? ((1u << (meta >> 5)) - 1u) << (meta & 31u) : 0u;
            }
            node_oct = oct;
        } else {
            const uint4 *np = nodes + (size_t)child * 4u;
            const uint4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3];
            const float px = __uint_as_float(n0.x), py = __uint_as_float(n0.y), pz = __uint_as_float(n0.z);
            const int ex = (int)(int8_t)(n0.w & 0xffu), ey = (int)(int8_t)((n0.w >> 8) & 0xffu),
                      ez = (int)(int8_t)((n0.w >> 16) & 0xffu);
            imask = n0.w >> 24;
            child_base = n1.x;
            tri_base = n1.y;
            const uint32_t meta_w = n1.z;
            // n1.w: qlo_x, n2.x: qlo_y, n2.y: qlo_z, n2.z: qhi_x, n2.w: qhi_y, n3.x: qhi_z, n3.y: axis
            const float sx = exp2i(ex) * ix, sy = exp2i(ey) * iy, sz = exp2i(ez) * iz;
            const float ax = (px - ox) * ix, ay = (py - oy) * iy, az = (pz - oz) * iz;
            const uint32_t wnx = nxs ? n2.z : n1.w, wfx = nxs ? n1.w : n2.z;
            const uint32_t wny = nys ? n2.w : n2.x, wfy = nys ? n2.x : n2.w;
            const uint32_t wnz = nzs ? n3.x : n2.y, wfz = nzs ? n2.y : n3.x;
            const uint32_t axis = n3.y;
            node_oct = ((oct >> (axis < 3u ? axis : 2u)) & 1u) ? 3u : 0u;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const float tnx = __builtin_fmaf(ubyte(wnx, s), sx, ax);
                const float tny = __builtin_fmaf(ubyte(wny, s), sy, ay);
                const float tnz = __builtin_fmaf(ubyte(wnz, s), sz, az);
                const float tfx = __builtin_fmaf(ubyte(wfx, s), sx, ax);
                const float tfy = __builtin_fmaf(ubyte(wfy, s), sy, ay);
                const float tfz = __builtin_fmaf(ubyte(wfz, s), sz, az);
                const float tnear = fmaxf(fmaxf(fmaxf(tnx, tny), tnz), box_tmin);
                const float tfar = fminf(fminf(fminf(tfx, tfy), tfz) * SPK_BOX_SLACK, tcull);
                const uint32_t meta = (meta_w >> (8 * s)) & 0xffu;
                const bool inner = (imask >> s) & 1u;
                const bool hit = tnear <= tfar;
                // selects, not branches: every lane evaluates both outcomes of every slot
                new_hits |= (hit && inner) ? (1u << (s ^ node_oct)) : 0u;
                tri_bits |= (hit && !inner) ? ((1u << (meta >> 5)) - 1u) << (meta & 31u) : 0u;
            }
        }
        g_base = child_base;
        g_imask = imask;
        g_hits = new_hits;
        g_oct = node_oct;

        if (TWO && level == 0u) {
            // TLAS leaves are instances: keep this node's inner children and its other instances on
            // the stack and enter the first instance
            if (tri_bits != 0u) {
                if (g_hits != 0u) {
                    stack[sp * stride] = make_uint2(g_base, (g_imask << 8) | (g_hits << 16) | g_oct);
                    ++sp;
                    g_hits = 0u;
                }
                const uint32_t k = __builtin_ctz(tri_bits);
                const uint32_t rest = tri_bits & (tri_bits - 1u);
                if (rest != 0u) {
                    stack[sp * stride] = make_uint2(tri_base, SPK_INST_ENTRY | rest);
                    ++sp;
                }
                enter(tri_base + k);
            }
            return true;
        }

        // leaf triangles of this node
        while (tri_bits != 0u) {
            const uint32_t k = __builtin_ctz(tri_bits);
            tri_bits &= tri_bits - 1u;
            const uint32_t ti = tri_base + k;
            const uint4 *tp = tris + (size_t)ti * 3u;
            const uint4 a = tp[0], b = tp[1], c = tp[2];
            if (COUNT) cnt.tris++;
            const float4 fa = make_float4(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z), 0.f);
            const float4 fb = make_float4(__uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z), 0.f);
            const float4 fc = make_float4(__uint_as_float(c.x), __uint_as_float(c.y), __uint_as_float(c.z), 0.f);
            const TriTest tt = tri_wt(wt, ox, oy, oz, fa, fb, fc);
            if (tt.hit && tt.t > box_tmin) {
                const uint32_t gid = TWO ? id_base + a.w : a.w;
                if (tt.t < best.t || (tt.t == best.t && gid < best.prim)) {
                    best.t = tt.t;
                    best.prim = gid;
                    best.tri = ti;
                    if (TWO) best.inst = cur;
                    best.U = tt.U;
                    best.V = tt.V;
                    best.W = tt.W;
                    best.det = tt.det;
                    if (ANYHIT) {  // occlusion query: any hit in (tmin, tmax] ends it
                        g_hits = 0u;
                        sp = 0;
                        return false;
                    }
                }
            }
        }
        return true;
    }
};

// One closest-hit (or, with ANYHIT, occlusion) query. best.t must hold tmax on entry.
template <int WIDTH, int STACK, bool COUNT, bool ANYHIT = false, bool TWO = false>
__device__ __forceinline__ void traverse(const uint4 *__restrict__ nodes, const uint4 *__restrict__ tris,
                                         uint32_t root, const RayIn &ray, uint2 *stack, int stride, BestHit &best,
                                         TraverseCounters &cnt, const uint4 *insts = nullptr) {
    Traversal<WIDTH, STACK, COUNT, ANYHIT, TWO> tr;
    tr.init(ray, root, insts);
    tr.best.t = best.t;
    while (tr.step(nodes, tris, stack, stride)) {
    }
    best = tr.best;
    cnt.nodes += tr.cnt.nodes;
    cnt.tris += tr.cnt.tris;
}

// ---- analytic spheres (SphereRenderProxy.cpp:5-38), tested after the BVH: spheres[2k] = (center,
// r^2), spheres[2k+1].x = material bits. A sphere hit stores tri = SPK_SPHERE_BIT | k and prim =
// id_base + k; ties with triangles keep the smaller id (the triangle).
#define SPK_SPHERE_BIT 0x80000000u

__device__ __forceinline__ bool sphere_t(const float4 s, const RayIn &r, float &t) {
    const float cx = r.ox - s.x, cy = r.oy - s.y, cz = r.oz - s.z;
    const float half_b = (r.dx * cx + r.dy * cy) + r.dz * cz;
    const float c = ((cx * cx + cy * cy) + cz * cz) - s.w;
    const float disc = half_b * half_b - c;  // a = 1: unit directions, as the reference assumes
    if (disc < 0.f) return false;
    const float sq = sqrtf(disc);
    t = -half_b + (c >= 0.f ? -sq : sq);
    return true;
}

template <bool ANYHIT = false>
__device__ __forceinline__ void spheres_closest(const float4 *__restrict__ spheres, uint32_t n, uint32_t id_base,
                                                const RayIn &r, BestHit &best) {
    const float tmin = fmaxf(r.tmin, 0.f);
    for (uint32_t k = 0; k < n; ++k) {
        float t;
        if (!sphere_t(spheres[2 * k], r, t)) continue;
        if (t > tmin && t > 0.f && (t < best.t || (t == best.t && id_base + k < best.prim))) {
            best.t = t;
            best.prim = id_base + k;
            best.tri = SPK_SPHERE_BIT | k;
            if (ANYHIT) return;
        }
    }
}

// refined distance of the final hit (bit-identical to the oracle's); in a two-level scene on the
// object-space triangle with the instance's object-space ray (the same t parameter)
__device__ __forceinline__ float final_t(const uint4 *__restrict__ tris, const RayIn &ray, const BestHit &best,
                                         const uint4 *insts = nullptr) {
    if (best.tri & SPK_SPHERE_BIT) return best.t;  // closed form already
    const uint4 *tp = tris + (size_t)best.tri * 3u;
    const uint4 a = tp[0], b = tp[1], c = tp[2];
    const float4 fa = make_float4(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z), 0.f);
    const float4 fb = make_float4(__uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z), 0.f);
    const float4 fc = make_float4(__uint_as_float(c.x), __uint_as_float(c.y), __uint_as_float(c.z), 0.f);
    if (insts) {
        const uint4 *ip = insts + (size_t)best.inst * 4u;
        float px, py, pz, dx, dy, dz;
        object_ray(ip[0], ip[1], ip[2], ray.ox, ray.oy, ray.oz, ray.dx, ray.dy, ray.dz, px, py, pz, dx, dy, dz);
        return refine_t(px, py, pz, dx, dy, dz, fa, fb, fc, best.t);
    }
    return refine_t(ray.ox, ray.oy, ray.oz, ray.dx, ray.dy, ray.dz, fa, fb, fc, best.t);
}

// Workgroups go to the 8 XCDs round-robin (workgroup b to XCD b mod 8), each XCD with its own L2.
// Within every window of 64 consecutive workgroups, XCD x runs 8 contiguous ones, so neighbouring
// rays of the coherence order share an L2, while all XCDs still work inside the same window at the
// same time (and share one working set in the Infinity Cache). A bijection on [0, nb). Measured on
// 8M sorted incoherent rays: 1.89 ms against 1.90-1.91 for the plain order; a window of 512 gave
// nothing, and one contiguous eighth of the grid per XCD was 18% slower.
__device__ __forceinline__ uint32_t xcd_group_block(uint32_t b, uint32_t nb) {
    const uint32_t base = b & ~63u, m = min(64u, nb - base), l = b - base;
    const uint32_t x = l & 7u, s = l >> 3, q = m >> 3, r = m & 7u;
    return base + x * q + min(x, r) + s;
}

}  // namespace spk
