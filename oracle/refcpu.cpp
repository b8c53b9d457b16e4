// refcpu — CPU restatement of Sparkle's `cpu` pipeline hot path (test infrastructure only).
// Built with -ffp-contract=off and without fast-math so every float operation rounds
// exactly once, in the order written; see refcpu.h for the list of restated files.
#include "refcpu.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <thread>
#include <vector>

namespace orcl {

// ---------------------------------------------------------------------------------------------
// math: a float3 whose reductions run ((x + y) + z), matching a non-vectorised Eigen Vector3
// (libraries/include/core/math/Types.h:12-43)
// ---------------------------------------------------------------------------------------------
struct V3 {
    float x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    explicit V3(const float *p) : x(p[0]), y(p[1]), z(p[2]) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
static inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
static inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
static inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
static inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
static inline V3 cmul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
static inline float sqnorm(V3 a) { return dot(a, a); }
static inline float norm(V3 a) { return std::sqrt(sqnorm(a)); }
// Eigen normalized(): divides by the norm, returns the input when the norm is zero
static inline V3 normalized(V3 a) {
    float n = sqnorm(a);
    if (n > 0.f) {
        float s = std::sqrt(n);
        return {a.x / s, a.y / s, a.z / s};
    }
    return a;
}
static inline float maxc(V3 a) { return std::max(std::max(a.x, a.y), a.z); }

static const V3 kUp(0, 0, 1), kRight(1, 0, 0), kOnes(1, 1, 1), kZeros(0, 0, 0);
static const float kPi = 3.14159265358979323846f;  // std::numbers::pi_v<float>
static const float kInvPi = 0.318309886183790671538f;  // std::numbers::inv_pi_v<float> (Types.h:53)
static const float kEps = 1e-6f;                    // Types.h:54
static const float kTolerance = 1e-4f;              // Types.h:55
static const float kOutputLimit = 6.f;              // CameraRenderProxy.h:16

// row-major 3x4 affine transform
struct Xf {
    float m[12];
    V3 point(V3 p) const {
        return {m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
                m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
    }
    V3 dir(V3 d) const {
        return {m[0] * d.x + m[1] * d.y + m[2] * d.z, m[4] * d.x + m[5] * d.y + m[6] * d.z,
                m[8] * d.x + m[9] * d.y + m[10] * d.z};
    }
    bool identity() const {
        // value comparison: an inverted identity carries -0.0 translations
        static const float I[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
        for (int k = 0; k < 12; ++k)
            if (m[k] != I[k]) return false;
        return true;
    }
};

// ---------------------------------------------------------------------------------------------
// RNG: xoshiro128+ 1.0 (Blackman & Vigna), state seeded by SplitMix64 (Sampler.h:16-41).
// Xoshiro-cpp seeds each 32-bit state word from one SplitMix64 output; we take the low 32 bits
// of four consecutive outputs (documented choice; the exact expansion is unpinned here).
// ---------------------------------------------------------------------------------------------
struct Rng {
    uint32_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (int k = 0; k < 4; ++k) {
            x += 0x9e3779b97f4a7c15ULL;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
            z = z ^ (z >> 31);
            s[k] = static_cast<uint32_t>(z);
        }
    }
    static inline uint32_t rotl(uint32_t v, int k) { return (v << k) | (v >> (32 - k)); }
    uint32_t next() {
        const uint32_t result = s[0] + s[3];
        const uint32_t t = s[1] << 9;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 11);
        return result;
    }
    // RandomUnit(): float(u32) / float(UINT32_MAX); the denominator rounds to 2^32 (Sampler.h:40)
    float unit() { return static_cast<float>(next()) / static_cast<float>(std::numeric_limits<uint32_t>::max()); }
    // UnitDisk() (Sampler.h:43-53)
    void disk(float &x, float &y) {
        float rand = unit();
        float theta = unit() * 2.0f * kPi;
        float r = std::sqrt(rand);
        x = r * std::cos(theta);
        y = r * std::sin(theta);
    }
};

// ---------------------------------------------------------------------------------------------
// triangle tests
// ---------------------------------------------------------------------------------------------
// Moller-Trumbore as in MeshRenderProxy.cpp:128-170 (local-space t, before the world_t step)
static inline bool tri_mt(V3 o, V3 d, V3 p0, V3 e1, V3 e2, float &t, float &u, float &v) {
    const V3 h = cross(d, e2);
    float a = dot(e1, h);
    if (std::abs(a) < kEps) return false;
    const V3 s = o - p0;
    float f = 1.f / a;
    u = f * dot(s, h);
    if (u < 0.f || u > 1.f) return false;
    const V3 q = cross(s, e1);
    v = f * dot(d, q);
    if (v < 0.f || u + v > 1.f) return false;
    t = f * dot(e2, q);
    return t > kEps;
}

// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013), FP32, no FMA, no FP64 fallback:
// an edge value of exactly zero counts as inside, so a ray through a shared edge hits both
// triangles and the closest-hit tie rule (smaller t, then smaller id) decides.
struct WtRay {
    V3 o;
    int kx, ky, kz;
    float Sx, Sy, Sz;
};
static inline WtRay wt_setup(V3 o, V3 d) {
    WtRay r;
    r.o = o;
    float ax = std::abs(d.x), ay = std::abs(d.y), az = std::abs(d.z);
    int kz = 0;
    if (ay > ax) kz = 1;
    if (az > std::max(ax, ay)) kz = 2;
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if (d[kz] < 0.f) std::swap(kx, ky);
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    r.Sx = d[kx] / d[kz];
    r.Sy = d[ky] / d[kz];
    r.Sz = 1.0f / d[kz];
    return r;
}
// The FP32 edge functions are used only when a forward error bound certifies their signs; a
// triangle whose signs are not certified (the ray passes within rounding distance of an edge, or
// the triangle is seen edge-on) is re-evaluated in double-float arithmetic. This replaces the FP64
// fallback of the paper and removes the false hits that exact-zero edge values would otherwise give.
static bool tri_wt_precise(const WtRay &r, V3 v0, V3 v1, V3 v2, float &t, float &U_, float &V_, float &W_,
                           float &det_);

static const float kWtErr = 8.f * 5.9604644775390625e-08f;   // 8u, u = 2^-24
static const float kWtErrAbs = 2.f * 5.9604644775390625e-08f;  // 2u

static inline bool tri_wt(const WtRay &r, V3 v0, V3 v1, V3 v2, float &t, float &U_, float &V_, float &W_,
                          float &det_) {
    const V3 A = v0 - r.o, B = v1 - r.o, C = v2 - r.o;
    const float sAx = r.Sx * A[r.kz], sAy = r.Sy * A[r.kz];
    const float sBx = r.Sx * B[r.kz], sBy = r.Sy * B[r.kz];
    const float sCx = r.Sx * C[r.kz], sCy = r.Sy * C[r.kz];
    const float Ax = A[r.kx] - sAx;
    const float Ay = A[r.ky] - sAy;
    const float Bx = B[r.kx] - sBx;
    const float By = B[r.ky] - sBy;
    const float Cx = C[r.kx] - sCx;
    const float Cy = C[r.ky] - sCy;
    const float U = Cx * By - Cy * Bx;
    const float V = Ax * Cy - Ay * Cx;
    const float W = Bx * Ay - By * Ax;
    // magnitudes before cancellation bound the rounding of the sheared coordinates and products
    const float mAx = std::abs(A[r.kx]) + std::abs(sAx), mAy = std::abs(A[r.ky]) + std::abs(sAy);
    const float mBx = std::abs(B[r.kx]) + std::abs(sBx), mBy = std::abs(B[r.ky]) + std::abs(sBy);
    const float mCx = std::abs(C[r.kx]) + std::abs(sCx), mCy = std::abs(C[r.ky]) + std::abs(sCy);
    const float eU = kWtErr * (mCx * mBy + mCy * mBx) + kWtErrAbs * std::abs(U);
    const float eV = kWtErr * (mAx * mCy + mAy * mCx) + kWtErrAbs * std::abs(V);
    const float eW = kWtErr * (mBx * mAy + mBy * mAx) + kWtErrAbs * std::abs(W);
    if (!(std::abs(U) > eU) || !(std::abs(V) > eV) || !(std::abs(W) > eW))
        return tri_wt_precise(r, v0, v1, v2, t, U_, V_, W_, det_);
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
    const float det = U + V + W;
    if (det == 0.f) return false;
    const float Az = r.Sz * A[r.kz];
    const float Bz = r.Sz * B[r.kz];
    const float Cz = r.Sz * C[r.kz];
    const float T = U * Az + V * Bz + W * Cz;
    if ((det < 0.f && T > 0.f) || (det > 0.f && T < 0.f)) return false;
    t = T / det;
    U_ = U;
    V_ = V;
    W_ = W;
    det_ = det;
    return true;
}

// Hit distance of the selected triangle, recomputed as n.(v0 - o) / n.d in double-float
// arithmetic (pairs hi + lo built from TwoSum and FMA-based TwoProd, Dekker 1971): the edges and
// v0 - o are exact, the normal and both dot products carry ~44 bits, and only the final division
// rounds to FP32. Accurate to a few ulps even for origins within 1e-9 of the surface.
static inline void two_sum(float a, float b, float &s, float &e) {
    s = a + b;
    float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
static inline void two_prod(float a, float b, float &p, float &e) {
    p = a * b;
    e = std::fma(a, b, -p);
}
struct DD {
    float h, l;
};
static inline DD dd_diff(float a, float b) {
    DD r;
    two_sum(a, -b, r.h, r.l);
    return r;
}
static inline DD dd_mul(DD a, DD b) {
    float p, e;
    two_prod(a.h, b.h, p, e);
    e = e + (a.h * b.l + a.l * b.h);
    DD r;
    r.h = p + e;
    r.l = e - (r.h - p);
    return r;
}
static inline DD dd_add(DD a, DD b) {
    float s, e;
    two_sum(a.h, b.h, s, e);
    e = e + (a.l + b.l);
    DD r;
    two_sum(s, e, r.h, r.l);
    return r;
}
static inline DD dd_neg(DD a) { return DD{-a.h, -a.l}; }
static inline DD dd_sel(int k, DD x, DD y, DD z) { return k == 0 ? x : (k == 1 ? y : z); }

// The watertight test of tri_wt evaluated in double-float arithmetic: exact translation, shear and
// edge functions carried to ~44 bits, signs decided on the rounded high parts.
static bool tri_wt_precise(const WtRay &r, V3 v0, V3 v1, V3 v2, float &t, float &U_, float &V_, float &W_,
                           float &det_) {
    const DD A0 = dd_diff(v0.x, r.o.x), A1 = dd_diff(v0.y, r.o.y), A2 = dd_diff(v0.z, r.o.z);
    const DD B0 = dd_diff(v1.x, r.o.x), B1 = dd_diff(v1.y, r.o.y), B2 = dd_diff(v1.z, r.o.z);
    const DD C0 = dd_diff(v2.x, r.o.x), C1 = dd_diff(v2.y, r.o.y), C2 = dd_diff(v2.z, r.o.z);
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
    if ((U.h < 0.f || V.h < 0.f || W.h < 0.f) && (U.h > 0.f || V.h > 0.f || W.h > 0.f)) return false;
    const DD det = dd_add(dd_add(U, V), W);
    if (det.h == 0.f) return false;
    const DD T = dd_add(dd_add(dd_mul(U, dd_mul(Sz, Akz)), dd_mul(V, dd_mul(Sz, Bkz))), dd_mul(W, dd_mul(Sz, Ckz)));
    if ((det.h < 0.f && T.h > 0.f) || (det.h > 0.f && T.h < 0.f)) return false;
    const float detf = det.h + det.l;
    t = (T.h + T.l) / detf;
    U_ = U.h;
    V_ = V.h;
    W_ = W.h;
    det_ = detf;
    return true;
}
static inline float refine_t(V3 o, V3 d, V3 v0, V3 v1, V3 v2, float t_fallback) {
    const DD e1x = dd_diff(v1.x, v0.x), e1y = dd_diff(v1.y, v0.y), e1z = dd_diff(v1.z, v0.z);
    const DD e2x = dd_diff(v2.x, v0.x), e2y = dd_diff(v2.y, v0.y), e2z = dd_diff(v2.z, v0.z);
    const DD ax = dd_diff(v0.x, o.x), ay = dd_diff(v0.y, o.y), az = dd_diff(v0.z, o.z);
    const DD nx = dd_add(dd_mul(e1y, e2z), dd_neg(dd_mul(e1z, e2y)));
    const DD ny = dd_add(dd_mul(e1z, e2x), dd_neg(dd_mul(e1x, e2z)));
    const DD nz = dd_add(dd_mul(e1x, e2y), dd_neg(dd_mul(e1y, e2x)));
    const DD num = dd_add(dd_add(dd_mul(nx, ax), dd_mul(ny, ay)), dd_mul(nz, az));
    const DD den = dd_add(dd_add(dd_mul(nx, DD{d.x, 0.f}), dd_mul(ny, DD{d.y, 0.f})), dd_mul(nz, DD{d.z, 0.f}));
    const float numf = num.h + num.l, denf = den.h + den.l;
    if (denf == 0.f) return t_fallback;
    const float t = numf / denf;
    if (!(t > 0.f) || !std::isfinite(t)) return t_fallback;
    return t;
}

// ---------------------------------------------------------------------------------------------
// binary SAH BVH (binned), used for both the TLAS over instances and one BLAS per mesh.
// The reference uses madmann91/bvh v2 DefaultBuilder<High>; tree shape does not change closest
// hits under the traversal below (no tmax culling), only visit order for exact-t ties.
// ---------------------------------------------------------------------------------------------
struct Box {
    float lo[3], hi[3];
    Box() {
        for (int i = 0; i < 3; ++i) {
            lo[i] = std::numeric_limits<float>::max();
            hi[i] = -std::numeric_limits<float>::max();
        }
    }
    void grow(const Box &b) {
        for (int i = 0; i < 3; ++i) {
            lo[i] = std::min(lo[i], b.lo[i]);
            hi[i] = std::max(hi[i], b.hi[i]);
        }
    }
    void grow(V3 p) {
        for (int i = 0; i < 3; ++i) {
            lo[i] = std::min(lo[i], p[i]);
            hi[i] = std::max(hi[i], p[i]);
        }
    }
    double area() const {
        double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.0;
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

struct BNode {
    Box box;
    uint32_t first;  // inner: index of left child (right = first + 1); leaf: first prim slot
    uint32_t count;  // 0 = inner
};

struct Bvh {
    std::vector<BNode> nodes;
    std::vector<uint32_t> prim_ids;

    void build(const std::vector<Box> &boxes, uint32_t max_leaf) {
        const uint32_t n = static_cast<uint32_t>(boxes.size());
        prim_ids.resize(n);
        for (uint32_t i = 0; i < n; ++i) prim_ids[i] = i;
        nodes.clear();
        if (n == 0) return;
        std::vector<float> cen(3 * (size_t)n);
        for (uint32_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) cen[3 * (size_t)i + a] = 0.5f * (boxes[i].lo[a] + boxes[i].hi[a]);
        nodes.reserve(2 * (size_t)n);
        nodes.push_back(BNode());
        struct Task {
            uint32_t node, begin, end;
        };
        std::vector<Task> stack;
        stack.push_back({0, 0, n});
        const int kBins = 16;
        while (!stack.empty()) {
            Task tk = stack.back();
            stack.pop_back();
            Box nb, cb;
            for (uint32_t i = tk.begin; i < tk.end; ++i) {
                nb.grow(boxes[prim_ids[i]]);
                cb.grow(V3(&cen[3 * (size_t)prim_ids[i]]));
            }
            BNode &node = nodes[tk.node];
            node.box = nb;
            const uint32_t cnt = tk.end - tk.begin;
            auto make_leaf = [&]() {
                nodes[tk.node].first = tk.begin;
                nodes[tk.node].count = cnt;
            };
            if (cnt <= 1) {
                make_leaf();
                continue;
            }
            int best_axis = -1, best_split = -1;
            double best_cost = std::numeric_limits<double>::max();
            for (int a = 0; a < 3; ++a) {
                float ext = cb.hi[a] - cb.lo[a];
                if (!(ext > 0.f)) continue;
                Box bb[kBins];
                uint32_t bc[kBins] = {0};
                for (uint32_t i = tk.begin; i < tk.end; ++i) {
                    uint32_t p = prim_ids[i];
                    int b = (int)((cen[3 * (size_t)p + a] - cb.lo[a]) / ext * kBins);
                    b = std::min(std::max(b, 0), kBins - 1);
                    bb[b].grow(boxes[p]);
                    bc[b]++;
                }
                double rarea[kBins];
                uint32_t rcnt[kBins];
                Box acc;
                uint32_t c = 0;
                for (int b = kBins - 1; b > 0; --b) {
                    acc.grow(bb[b]);
                    c += bc[b];
                    rarea[b] = acc.area();
                    rcnt[b] = c;
                }
                Box lacc;
                uint32_t lc = 0;
                for (int b = 0; b < kBins - 1; ++b) {
                    lacc.grow(bb[b]);
                    lc += bc[b];
                    if (lc == 0 || rcnt[b + 1] == 0) continue;
                    double cost = lacc.area() * lc + rarea[b + 1] * rcnt[b + 1];
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = a;
                        best_split = b;
                    }
                }
            }
            double leaf_cost = nb.area() * cnt;
            uint32_t mid;
            if (best_axis < 0) {
                if (cnt <= max_leaf) {
                    make_leaf();
                    continue;
                }
                mid = tk.begin + cnt / 2;  // all centroids equal: split in the middle
            } else {
                if (cnt <= max_leaf && best_cost * 1.0 + nb.area() >= leaf_cost) {
                    make_leaf();
                    continue;
                }
                float ext = cb.hi[best_axis] - cb.lo[best_axis];
                auto it = std::partition(prim_ids.begin() + tk.begin, prim_ids.begin() + tk.end, [&](uint32_t p) {
                    int b = (int)((cen[3 * (size_t)p + best_axis] - cb.lo[best_axis]) / ext * kBins);
                    b = std::min(std::max(b, 0), kBins - 1);
                    return b <= best_split;
                });
                mid = static_cast<uint32_t>(it - prim_ids.begin());
                if (mid == tk.begin || mid == tk.end) mid = tk.begin + cnt / 2;
            }
            uint32_t left = static_cast<uint32_t>(nodes.size());
            nodes.push_back(BNode());
            nodes.push_back(BNode());
            nodes[tk.node].first = left;
            nodes[tk.node].count = 0;
            stack.push_back({left + 1, mid, tk.end});
            stack.push_back({left, tk.begin, mid});
        }
    }
};

// slab test of bvh v2's non-robust mode: tmin=0, tmax=+inf, never shrunk
struct SlabRay {
    float o[3], inv[3];
};
static inline SlabRay slab_setup(V3 o, V3 d) {
    SlabRay r;
    for (int a = 0; a < 3; ++a) {
        r.o[a] = o[a];
        float da = d[a];
        // bvh v2 replaces a zero direction component by a tiny value before inverting
        if (da == 0.f) da = std::numeric_limits<float>::epsilon();
        r.inv[a] = 1.f / da;
    }
    return r;
}
static inline bool slab_hit(const SlabRay &r, const Box &b, float tmax, float &tnear) {
    float t0 = 0.f, t1 = tmax;
    for (int a = 0; a < 3; ++a) {
        float ta = (b.lo[a] - r.o[a]) * r.inv[a];
        float tb = (b.hi[a] - r.o[a]) * r.inv[a];
        if (ta > tb) std::swap(ta, tb);
        t0 = std::max(t0, ta);
        t1 = std::min(t1, tb);
    }
    tnear = t0;
    return t0 <= t1;
}
// conservative box test in double with a relative widening, used by the checker modes so that
// box rounding can never drop a triangle the test below would accept
static inline bool slab_hit_robust(V3 o, V3 d, const Box &b, double tmax) {
    double t0 = 0.0, t1 = tmax;
    for (int a = 0; a < 3; ++a) {
        double ext = std::max(1e-30, ((double)b.hi[a] - (double)b.lo[a]) * 1e-6);
        double lo = (double)b.lo[a] - ext, hi = (double)b.hi[a] + ext;
        double da = d[a], oa = o[a];
        if (da == 0.0) {
            if (oa < lo || oa > hi) return false;
            continue;
        }
        double ta = (lo - oa) / da, tb = (hi - oa) / da;
        if (ta > tb) std::swap(ta, tb);
        t0 = std::max(t0, ta);
        t1 = std::min(t1, tb);
    }
    return t0 <= t1 * (1.0 + 1e-9);
}

// ---------------------------------------------------------------------------------------------
// scene: one BLAS per mesh over local-space triangles, TLAS over instance world boxes
// (MeshRenderProxy.cpp:31-65, SceneRenderProxy.cpp:31-69)
// ---------------------------------------------------------------------------------------------
struct Mesh {
    std::vector<float> pos, nrm, tan, uv;
    std::vector<uint32_t> idx;
    uint32_t ntri = 0;
    // MeshRenderProxy::BLAS::Triangle {p0, e1, e2, n} (MeshRenderProxy.cpp:78-92)
    std::vector<V3> p0, e1, e2;
    Bvh bvh;
    Box local_box;
    V3 vtx(uint32_t i) const { return V3(&pos[3 * (size_t)i]); }
    void tri(uint32_t f, V3 &a, V3 &b, V3 &c) const {
        a = vtx(idx[3 * (size_t)f]);
        b = vtx(idx[3 * (size_t)f + 1]);
        c = vtx(idx[3 * (size_t)f + 2]);
    }
};

struct Instance {
    uint32_t mesh, material;
    Xf to_world, to_local;
    uint64_t prim_offset;
};

}  // namespace orcl

namespace orcl {
struct Texture {
    uint32_t w = 0, h = 0;
    bool srgb = false;
    std::vector<float> texels;  // 4 per texel
};
}  // namespace orcl

namespace orcl {
struct Sphere {
    float c[3];
    float r2;  // scaled_radius_2_ (SphereRenderProxy.h:29-32)
    uint32_t material;
};
struct Emitter {  // world-space emissive triangle for next-event estimation
    float p[9];
    float uv[6];
    uint32_t material;
};
}  // namespace orcl

struct orcl_scene {
    std::vector<orcl::Mesh> meshes;
    std::vector<orcl::Instance> instances;  // in TLAS leaf order
    std::vector<orcl_material> materials;
    std::vector<orcl::Texture> textures;
    std::vector<orcl::Emitter> emitters;  // instance order, face order
    std::vector<orcl::Sphere> spheres;
    int32_t sky_faces[6] = {-1, -1, -1, -1, -1, -1};
    std::vector<float> emitter_cdf;       // running sum of area x mean emitted radiance
    orcl::Bvh tlas;
    uint64_t total_tris = 0;
};

namespace orcl {

struct Candidate {
    float t = std::numeric_limits<float>::max();  // IntersectionCandidate (Intersection.h:91-104)
    float u = 0, v = 0;
    uint32_t face = 0;
    int inst = -1;
    int sphere = -1;
    uint64_t gid = ~0ull;
};

// SphereRenderProxy::IntersectInternal (SphereRenderProxy.cpp:5-38): a = 1 for unit directions
static inline bool sphere_t(const Sphere &s, V3 o, V3 d, float &t) {
    const float cx = o.x - s.c[0], cy = o.y - s.c[1], cz = o.z - s.c[2];
    const float half_b = (d.x * cx + d.y * cy) + d.z * cz;
    const float c = ((cx * cx + cy * cy) + cz * cz) - s.r2;
    const float disc = half_b * half_b - c;
    if (disc < 0.f) return false;
    const float sq = std::sqrt(disc);
    t = -half_b + (c >= 0.f ? -sq : sq);
    return true;
}

// BLAS traversal + MT + world_t (MeshRenderProxy.cpp:94-194), shared candidate across meshes
static void blas_intersect_ref(const orcl_scene &sc, int ii, V3 wo, V3 wd, Candidate &cand) {
    const Instance &ins = sc.instances[ii];
    const Mesh &m = sc.meshes[ins.mesh];
    V3 lo = ins.to_local.point(wo), ld = ins.to_local.dir(wd);
    SlabRay sr = slab_setup(lo, ld);
    if (m.bvh.nodes.empty()) return;
    std::vector<uint32_t> stack;
    stack.reserve(64);
    // bvh v2 pushes the root and tests children boxes only (ordered traversal)
    uint32_t top = 0;
    stack.push_back(top);
    while (!stack.empty()) {
        uint32_t ni = stack.back();
        stack.pop_back();
        bool leaf_reached = true;
        while (m.bvh.nodes[ni].count == 0) {
            const BNode &L = m.bvh.nodes[m.bvh.nodes[ni].first];
            const BNode &R = m.bvh.nodes[m.bvh.nodes[ni].first + 1];
            float tl, tr;
            bool hl = slab_hit(sr, L.box, std::numeric_limits<float>::infinity(), tl);
            bool hr = slab_hit(sr, R.box, std::numeric_limits<float>::infinity(), tr);
            uint32_t li = m.bvh.nodes[ni].first, ri = li + 1;
            if (hl) {
                uint32_t nearI = li;
                if (hr) {
                    uint32_t farI = ri;
                    if (tl > tr) std::swap(nearI, farI);
                    stack.push_back(farI);
                }
                ni = nearI;
            } else if (hr) {
                ni = ri;
            } else {
                leaf_reached = false;
                break;
            }
        }
        if (!leaf_reached) continue;
        const BNode &leaf = m.bvh.nodes[ni];
        for (uint32_t k = leaf.first; k < leaf.first + leaf.count; ++k) {
            uint32_t f = m.bvh.prim_ids[k];
            float t, u, v;
            if (!tri_mt(lo, ld, m.p0[f], m.e1[f], m.e2[f], t, u, v)) continue;
            // world_p = transform(local.At(t)); world_t = InverseAt(world_p) (Ray.h:43-47)
            V3 lp = lo + ld * t;
            V3 wp = ins.to_world.point(lp);
            float world_t = (wp - wo).x / wd.x;
            if (world_t > 0 && world_t < cand.t) {
                cand.t = world_t;
                cand.u = u;
                cand.v = v;
                cand.face = f;
                cand.inst = ii;
            }
        }
    }
}

// TLAS traversal (SceneRenderProxy.cpp:82-126)
static void scene_intersect_ref(const orcl_scene &sc, V3 o, V3 d, Candidate &cand) {
    SlabRay sr = slab_setup(o, d);
    std::vector<uint32_t> stack;
    if (!sc.tlas.nodes.empty()) stack.push_back(0);
    while (!stack.empty()) {
        uint32_t ni = stack.back();
        stack.pop_back();
        bool leaf_reached = true;
        while (sc.tlas.nodes[ni].count == 0) {
            uint32_t li = sc.tlas.nodes[ni].first, ri = li + 1;
            float tl, tr;
            bool hl = slab_hit(sr, sc.tlas.nodes[li].box, std::numeric_limits<float>::infinity(), tl);
            bool hr = slab_hit(sr, sc.tlas.nodes[ri].box, std::numeric_limits<float>::infinity(), tr);
            if (hl) {
                uint32_t nearI = li;
                if (hr) {
                    uint32_t farI = ri;
                    if (tl > tr) std::swap(nearI, farI);
                    stack.push_back(farI);
                }
                ni = nearI;
            } else if (hr) {
                ni = ri;
            } else {
                leaf_reached = false;
                break;
            }
        }
        if (!leaf_reached) continue;
        const BNode &leaf = sc.tlas.nodes[ni];
        for (uint32_t k = leaf.first; k < leaf.first + leaf.count; ++k) blas_intersect_ref(sc, (int)k, o, d, cand);
    }
    if (cand.inst >= 0) cand.gid = sc.instances[cand.inst].prim_offset + cand.face;
    // spheres: IsCloserHit = t > 0 and t < the candidate's t
    for (size_t k = 0; k < sc.spheres.size(); ++k) {
        float t;
        if (sphere_t(sc.spheres[k], o, d, t) && t > 0.f && t < cand.t) {
            cand.t = t;
            cand.inst = -1;
            cand.sphere = (int)k;
            cand.gid = sc.total_tris + k;
        }
    }
}

// watertight closest hit, order-independent: smaller t wins, exact ties go to the smaller id
struct WtBest {
    float t;
    uint64_t gid = ~0ull;
    int inst = -1;
    int sphere = -1;
    uint32_t face = 0;
    float U = 0, V = 0, W = 0, det = 1;
};
static inline void wt_consider(WtBest &b, float t, float tmin, uint64_t gid, int inst, uint32_t face, float U,
                               float V, float W, float det) {
    if (!(t > tmin)) return;
    if (t < b.t || (t == b.t && gid < b.gid)) {
        b.t = t;
        b.gid = gid;
        b.inst = inst;
        b.face = face;
        b.U = U;
        b.V = V;
        b.W = W;
        b.det = det;
    }
}

// The watertight twin works on world-space triangles, as the device does: an instance with a
// transform has its vertices baked with the same float operations as the device upload, and its
// BVH is walked with the local-space ray only for culling (parametric t is the same in both spaces).
static inline void world_tri(const Instance &ins, const Mesh &m, uint32_t f, V3 &a, V3 &b, V3 &c) {
    m.tri(f, a, b, c);
    if (!ins.to_world.identity()) {
        a = ins.to_world.point(a);
        b = ins.to_world.point(b);
        c = ins.to_world.point(c);
    }
}

static void blas_intersect_wt(const orcl_scene &sc, int ii, V3 wo, V3 wd, float tmin, WtBest &best, bool brute) {
    const Instance &ins = sc.instances[ii];
    const Mesh &m = sc.meshes[ins.mesh];
    const bool ident = ins.to_local.identity();
    V3 lo = ident ? wo : ins.to_local.point(wo), ld = ident ? wd : ins.to_local.dir(wd);
    WtRay wr = wt_setup(wo, wd);
    auto test = [&](uint32_t f) {
        V3 a, b, c;
        world_tri(ins, m, f, a, b, c);
        float t, U, V, W, det;
        if (tri_wt(wr, a, b, c, t, U, V, W, det))
            wt_consider(best, t, tmin, ins.prim_offset + f, ii, f, U, V, W, det);
    };
    // a transformed instance is tested triangle by triangle (the rounded local ray could cull)
    if (brute || !ident) {
        for (uint32_t f = 0; f < m.ntri; ++f) test(f);
        return;
    }
    if (m.bvh.nodes.empty()) return;
    std::vector<uint32_t> stack;
    stack.push_back(0);
    while (!stack.empty()) {
        uint32_t ni = stack.back();
        stack.pop_back();
        const BNode &nd = m.bvh.nodes[ni];
        double tmax = best.gid == ~0ull ? std::numeric_limits<double>::infinity() : (double)best.t * (1.0 + 1e-5);
        if (!slab_hit_robust(lo, ld, nd.box, tmax)) continue;
        if (nd.count == 0) {
            stack.push_back(nd.first + 1);
            stack.push_back(nd.first);
        } else {
            for (uint32_t k = nd.first; k < nd.first + nd.count; ++k) test(m.bvh.prim_ids[k]);
        }
    }
}

static void scene_intersect_wt(const orcl_scene &sc, V3 o, V3 d, float tmin, float tmax, WtBest &best, bool brute) {
    best.t = tmax;
    // every instance is visited; each BLAS walk culls robustly against the best hit so far
    for (size_t ii = 0; ii < sc.instances.size(); ++ii) blas_intersect_wt(sc, (int)ii, o, d, tmin, best, brute);
}

// FP64 closest hit with flags
static double g_eps_edge = 1e-4, g_eps_tie = 1e-6, g_eps_graze = 1e-3;

struct F64Best {
    double t = std::numeric_limits<double>::infinity();
    uint64_t gid = ~0ull;
    double minb = 0, cosn = 1;
    double t2 = std::numeric_limits<double>::infinity();     // second-best t among accepted hits
    double t_out = std::numeric_limits<double>::infinity();  // closest near-edge rejected candidate
};

static void f64_tri(const double o[3], const double d[3], const double a[3], const double b[3], const double c[3],
                    double tmin, uint64_t gid, F64Best &best) {
    double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double h[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    double det = e1[0] * h[0] + e1[1] * h[1] + e1[2] * h[2];
    if (det == 0.0) return;
    double s[3] = {o[0] - a[0], o[1] - a[1], o[2] - a[2]};
    double u = (s[0] * h[0] + s[1] * h[1] + s[2] * h[2]) / det;
    double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
    double v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) / det;
    double w = 1.0 - u - v;
    double t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) / det;
    if (!(t > tmin)) return;
    double minb = std::min(u, std::min(v, w));
    if (minb < -g_eps_edge) return;
    if (minb < 0.0) {
        // just outside: the FP32 test may legitimately accept it
        best.t_out = std::min(best.t_out, t);
        return;
    }
    if (t < best.t || (t == best.t && gid < best.gid)) {
        best.t2 = std::min(best.t2, best.t);
        best.t = t;
        best.gid = gid;
        best.minb = minb;
        double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        double dd = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        best.cosn = nn > 0 ? std::abs(n[0] * d[0] + n[1] * d[1] + n[2] * d[2]) / (nn * dd) : 0.0;
    } else {
        best.t2 = std::min(best.t2, t);
    }
}

static void scene_intersect_f64(const orcl_scene &sc, V3 o, V3 d, float tmin, float tmax, F64Best &best) {
    best.t = tmax;
    for (size_t ii = 0; ii < sc.instances.size(); ++ii) {
        const Instance &ins = sc.instances[ii];
        const Mesh &m = sc.meshes[ins.mesh];
        if (!ins.to_world.identity()) {
            // a transformed instance is checked on its baked world-space triangles (the geometry the
            // device intersects), triangle by triangle
            const double ow[3] = {o.x, o.y, o.z}, dw[3] = {d.x, d.y, d.z};
            for (uint32_t f = 0; f < m.ntri; ++f) {
                V3 a, b, c;
                world_tri(ins, m, f, a, b, c);
                double A[3] = {a.x, a.y, a.z}, B[3] = {b.x, b.y, b.z}, C[3] = {c.x, c.y, c.z};
                f64_tri(ow, dw, A, B, C, (double)tmin, ins.prim_offset + f, best);
            }
            continue;
        }
        // instance transforms are applied in double
        double od[3], dd[3];
        const float *M = ins.to_local.m;
        for (int r = 0; r < 3; ++r) {
            od[r] = (double)M[4 * r] * o.x + (double)M[4 * r + 1] * o.y + (double)M[4 * r + 2] * o.z + M[4 * r + 3];
            dd[r] = (double)M[4 * r] * d.x + (double)M[4 * r + 1] * d.y + (double)M[4 * r + 2] * d.z;
        }
        V3 lo((float)od[0], (float)od[1], (float)od[2]), ld((float)dd[0], (float)dd[1], (float)dd[2]);
        if (m.bvh.nodes.empty()) continue;
        std::vector<uint32_t> stack;
        stack.push_back(0);
        while (!stack.empty()) {
            uint32_t ni = stack.back();
            stack.pop_back();
            const BNode &nd = m.bvh.nodes[ni];
            const double cull = best.t * (1.0 + 1e-4);
            // widen by the float rounding of the local ray, then test
            Box wb = nd.box;
            for (int a = 0; a < 3; ++a) {
                float pad = 1e-5f * std::max(1.f, std::max(std::abs(wb.lo[a]), std::abs(wb.hi[a])));
                wb.lo[a] -= pad;
                wb.hi[a] += pad;
            }
            if (!slab_hit_robust(lo, ld, wb, cull)) continue;
            if (nd.count == 0) {
                stack.push_back(nd.first + 1);
                stack.push_back(nd.first);
                continue;
            }
            for (uint32_t k = nd.first; k < nd.first + nd.count; ++k) {
                uint32_t f = m.bvh.prim_ids[k];
                V3 a, b, c;
                m.tri(f, a, b, c);
                double A[3] = {a.x, a.y, a.z}, B[3] = {b.x, b.y, b.z}, C[3] = {c.x, c.y, c.z};
                f64_tri(od, dd, A, B, C, (double)tmin, ins.prim_offset + f, best);
            }
        }
    }
}

}  // namespace orcl

using namespace orcl;

extern "C" {

orcl_scene *orcl_scene_create(const orcl_mesh *meshes, uint32_t num_meshes, const orcl_instance *instances,
                              uint32_t num_instances, const orcl_material *materials, uint32_t num_materials) {
    auto *sc = new orcl_scene();
    sc->meshes.resize(num_meshes);
    for (uint32_t mi = 0; mi < num_meshes; ++mi) {
        const orcl_mesh &src = meshes[mi];
        Mesh &m = sc->meshes[mi];
        m.pos.assign(src.positions, src.positions + 3 * (size_t)src.num_vertices);
        m.idx.assign(src.indices, src.indices + 3 * (size_t)src.num_triangles);
        if (src.normals) m.nrm.assign(src.normals, src.normals + 3 * (size_t)src.num_vertices);
        if (src.tangents) m.tan.assign(src.tangents, src.tangents + 4 * (size_t)src.num_vertices);
        if (src.uvs) m.uv.assign(src.uvs, src.uvs + 2 * (size_t)src.num_vertices);
        m.ntri = src.num_triangles;
        m.p0.resize(m.ntri);
        m.e1.resize(m.ntri);
        m.e2.resize(m.ntri);
        std::vector<Box> boxes(m.ntri);
        for (uint32_t f = 0; f < m.ntri; ++f) {
            V3 a, b, c;
            m.tri(f, a, b, c);
            m.p0[f] = a;
            m.e1[f] = b - a;
            m.e2[f] = c - a;
            boxes[f].grow(a);
            boxes[f].grow(b);
            boxes[f].grow(c);
            m.local_box.grow(boxes[f]);
        }
        m.bvh.build(boxes, 8);
    }
    // TLAS over instance world boxes (AABB::TransformTo, AABB.h:57-79)
    std::vector<Instance> inst(num_instances);
    std::vector<Box> wboxes(num_instances);
    for (uint32_t i = 0; i < num_instances; ++i) {
        inst[i].mesh = instances[i].mesh;
        inst[i].material = instances[i].material;
        std::memcpy(inst[i].to_world.m, instances[i].world_from_local, sizeof(float) * 12);
        std::memcpy(inst[i].to_local.m, instances[i].local_from_world, sizeof(float) * 12);
        const Box &lb = sc->meshes[inst[i].mesh].local_box;
        for (int c = 0; c < 8; ++c) {
            V3 p((c & 1) ? lb.hi[0] : lb.lo[0], (c & 2) ? lb.hi[1] : lb.lo[1], (c & 4) ? lb.hi[2] : lb.lo[2]);
            wboxes[i].grow(inst[i].to_world.point(p));
        }
    }
    // global triangle ids follow the caller's instance order
    uint64_t off = 0;
    for (uint32_t i = 0; i < num_instances; ++i) {
        inst[i].prim_offset = off;
        off += sc->meshes[inst[i].mesh].ntri;
    }
    sc->total_tris = off;
    sc->tlas.build(wboxes, 1);
    // reorder primitives by TLAS leaf order (SceneRenderProxy.cpp:59-68)
    sc->instances.resize(num_instances);
    for (uint32_t i = 0; i < num_instances; ++i) sc->instances[i] = inst[sc->tlas.prim_ids[i]];
    sc->materials.assign(materials, materials + num_materials);
    // emitters for next-event estimation: world-space emissive triangles in the caller's instance
    // order, weighted by area x mean emitted radiance (factor)
    double acc = 0.0;
    for (uint32_t i = 0; i < num_instances; ++i) {
        const orcl_material &mt = materials[inst[i].material];
        const float en = std::sqrt(mt.emissive[0] * mt.emissive[0] + mt.emissive[1] * mt.emissive[1] +
                                   mt.emissive[2] * mt.emissive[2]);
        if (!(en > kEps)) continue;
        const Mesh &m = sc->meshes[inst[i].mesh];
        for (uint32_t f = 0; f < m.ntri; ++f) {
            Emitter e;
            V3 v[3];
            m.tri(f, v[0], v[1], v[2]);
            for (int k = 0; k < 3; ++k) {
                const V3 w = inst[i].to_world.point(v[k]);
                e.p[3 * k] = w.x, e.p[3 * k + 1] = w.y, e.p[3 * k + 2] = w.z;
                const uint32_t vi = m.idx[3 * (size_t)f + k];
                e.uv[2 * k] = m.uv.empty() ? 0.f : m.uv[2 * vi];
                e.uv[2 * k + 1] = m.uv.empty() ? 0.f : m.uv[2 * vi + 1];
            }
            e.material = inst[i].material;
            const double e1[3] = {(double)e.p[3] - e.p[0], (double)e.p[4] - e.p[1], (double)e.p[5] - e.p[2]};
            const double e2[3] = {(double)e.p[6] - e.p[0], (double)e.p[7] - e.p[1], (double)e.p[8] - e.p[2]};
            const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                         cz = e1[0] * e2[1] - e1[1] * e2[0];
            acc += 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz) * (mt.emissive[0] + mt.emissive[1] + mt.emissive[2]) / 3.0;
            sc->emitters.push_back(e);
            sc->emitter_cdf.push_back((float)acc);
        }
    }
    return sc;
}

void orcl_scene_destroy(orcl_scene *scene) { delete scene; }

void orcl_scene_set_sky_faces(orcl_scene *scene, const int32_t *faces) {
    for (int f = 0; f < 6; ++f) scene->sky_faces[f] = faces ? faces[f] : -1;
}

void orcl_scene_set_spheres(orcl_scene *scene, const orcl_sphere *spheres, uint32_t num_spheres) {
    scene->spheres.resize(num_spheres);
    for (uint32_t k = 0; k < num_spheres; ++k) {
        orcl::Sphere &d = scene->spheres[k];
        for (int a = 0; a < 3; ++a) d.c[a] = spheres[k].center[a];
        d.r2 = spheres[k].radius * spheres[k].radius;
        d.material = spheres[k].material;
    }
}

void orcl_scene_set_textures(orcl_scene *scene, const orcl_texture *textures, uint32_t num_textures) {
    scene->textures.assign(num_textures, orcl::Texture());
    for (uint32_t k = 0; k < num_textures; ++k) {
        const orcl_texture &t = textures[k];
        orcl::Texture &d = scene->textures[k];
        d.w = t.width;
        d.h = t.height;
        d.srgb = t.format == 1u;
        const size_t n = 4 * (size_t)t.width * t.height;
        d.texels.resize(n);
        if (t.format == 2u) {
            std::memcpy(d.texels.data(), t.pixels, n * sizeof(float));
        } else {
            // ColorToVec (Utilities.h:377-380)
            const uint8_t *src = static_cast<const uint8_t *>(t.pixels);
            for (size_t i = 0; i < n; ++i) d.texels[i] = static_cast<float>(src[i]) / 255.f;
        }
    }
}
// Twin of the device's two-level traversal (no counterpart in the reference): the library uploads,
// per instance, object_from_world = the inverse of the 3x3 part computed in double and rounded once,
// with translation -inverse * t in double, rounded once; the device transforms the world ray with it
// in the order ((a0 x + a1 y) + a2 z) + a3 and tests the instance's object-space triangles. The
// closest hit (t, then global id) over all instances does not depend on the TLAS or BLAS shapes.
static void object_from_world(const Xf &w, float r[12]) {
    const float *M = w.m;
    const double a[9] = {M[0], M[1], M[2], M[4], M[5], M[6], M[8], M[9], M[10]};
    const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                       a[2] * (a[3] * a[7] - a[4] * a[6]);
    const double inv[9] = {(a[4] * a[8] - a[5] * a[7]) / det, (a[2] * a[7] - a[1] * a[8]) / det,
                           (a[1] * a[5] - a[2] * a[4]) / det, (a[5] * a[6] - a[3] * a[8]) / det,
                           (a[0] * a[8] - a[2] * a[6]) / det, (a[2] * a[3] - a[0] * a[5]) / det,
                           (a[3] * a[7] - a[4] * a[6]) / det, (a[1] * a[6] - a[0] * a[7]) / det,
                           (a[0] * a[4] - a[1] * a[3]) / det};
    const double t[3] = {M[3], M[7], M[11]};
    for (int row = 0; row < 3; ++row) {
        r[4 * row] = (float)inv[3 * row];
        r[4 * row + 1] = (float)inv[3 * row + 1];
        r[4 * row + 2] = (float)inv[3 * row + 2];
        r[4 * row + 3] = (float)-(inv[3 * row] * t[0] + inv[3 * row + 1] * t[1] + inv[3 * row + 2] * t[2]);
    }
}

static inline void object_ray_twin(const float *r, V3 wo, V3 wd, V3 &o, V3 &d) {
    o = V3(((r[0] * wo.x + r[1] * wo.y) + r[2] * wo.z) + r[3], ((r[4] * wo.x + r[5] * wo.y) + r[6] * wo.z) + r[7],
           ((r[8] * wo.x + r[9] * wo.y) + r[10] * wo.z) + r[11]);
    d = V3((r[0] * wd.x + r[1] * wd.y) + r[2] * wd.z, (r[4] * wd.x + r[5] * wd.y) + r[6] * wd.z,
           (r[8] * wd.x + r[9] * wd.y) + r[10] * wd.z);
}

uint64_t orcl_scene_num_triangles(const orcl_scene *scene) { return scene->total_tris; }

void orcl_set_fp64_eps(double eps_edge, double eps_tie, double eps_graze) {
    g_eps_edge = eps_edge;
    g_eps_tie = eps_tie;
    g_eps_graze = eps_graze;
}

static void parallel_for(uint64_t n, int num_threads, const std::function<void(uint64_t, uint64_t)> &fn);

int orcl_trace(const orcl_scene *scene, const orcl_ray *rays, uint64_t num_rays, int mode, int num_threads,
               orcl_hit *hits, uint32_t *flags, double *t64) {
    if (!scene || (!rays && num_rays) || (!hits && num_rays)) return -1;
    const orcl_scene &sc = *scene;
    std::vector<float> inv(mode == ORCL_TRACE_WT_TWO ? 12 * sc.instances.size() : 0);
    for (size_t k = 0; 12 * k < inv.size(); ++k) object_from_world(sc.instances[k].to_world, &inv[12 * k]);
    parallel_for(num_rays, num_threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) {
            const orcl_ray &r = rays[i];
            V3 o(r.org), d(r.dir);
            hits[i].prim = 0xFFFFFFFFu;
            hits[i].t = std::numeric_limits<float>::infinity();
            if (mode == ORCL_TRACE_WT_TWO) {
                bool ok = std::isfinite(o.x) && std::isfinite(o.y) && std::isfinite(o.z) && std::isfinite(d.x) &&
                          std::isfinite(d.y) && std::isfinite(d.z) && (d.x != 0.f || d.y != 0.f || d.z != 0.f);
                if (!ok || std::isnan(r.tmin) || std::isnan(r.tmax)) continue;
                WtBest best;
                best.t = r.tmax;
                const float tmin = std::max(r.tmin, 0.f);
                for (size_t ii = 0; ii < sc.instances.size(); ++ii) {
                    const Instance &ins = sc.instances[ii];
                    const Mesh &m = sc.meshes[ins.mesh];
                    V3 lo, ld;
                    object_ray_twin(&inv[12 * ii], o, d, lo, ld);
                    WtRay wr = wt_setup(lo, ld);
                    for (uint32_t f = 0; f < m.ntri; ++f) {
                        V3 a, bb, c;
                        m.tri(f, a, bb, c);
                        float t, U, V, W, det;
                        if (tri_wt(wr, a, bb, c, t, U, V, W, det))
                            wt_consider(best, t, tmin, ins.prim_offset + f, (int)ii, f, U, V, W, det);
                    }
                }
                for (size_t k = 0; k < sc.spheres.size(); ++k) {
                    float t;
                    const uint64_t gid = sc.total_tris + k;
                    if (sphere_t(sc.spheres[k], o, d, t) && t > tmin && t > 0.f &&
                        (t < best.t || (t == best.t && gid < best.gid))) {
                        best.t = t;
                        best.gid = gid;
                        best.sphere = (int)k;
                    }
                }
                if (best.sphere >= 0) {
                    hits[i].prim = (uint32_t)best.gid;
                    hits[i].t = best.t;
                } else if (best.gid != ~0ull) {
                    const Instance &ins = sc.instances[best.inst];
                    V3 lo, ld, a, bb, c;
                    object_ray_twin(&inv[12 * (size_t)best.inst], o, d, lo, ld);
                    sc.meshes[ins.mesh].tri(best.face, a, bb, c);
                    hits[i].prim = (uint32_t)best.gid;
                    hits[i].t = refine_t(lo, ld, a, bb, c, best.t);
                }
            } else if (mode == ORCL_TRACE_REF_MT) {
                Candidate c;
                scene_intersect_ref(sc, o, d, c);
                if (c.inst >= 0 || c.sphere >= 0) {
                    hits[i].prim = (uint32_t)c.gid;
                    hits[i].t = c.t;
                }
            } else if (mode == ORCL_TRACE_WT || mode == ORCL_TRACE_WT_BRUTE) {
                WtBest best;
                bool ok = std::isfinite(o.x) && std::isfinite(o.y) && std::isfinite(o.z) && std::isfinite(d.x) &&
                          std::isfinite(d.y) && std::isfinite(d.z) && (d.x != 0.f || d.y != 0.f || d.z != 0.f);
                if (!ok) continue;
                if (std::isnan(r.tmin) || std::isnan(r.tmax)) continue;
                // a negative tmin is treated as 0, as on the device
                scene_intersect_wt(sc, o, d, std::max(r.tmin, 0.f), r.tmax, best, mode == ORCL_TRACE_WT_BRUTE);
                // spheres after the triangles, as on the device: smaller t, then smaller id
                for (size_t k = 0; k < sc.spheres.size(); ++k) {
                    float t;
                    const uint64_t gid = sc.total_tris + k;
                    if (sphere_t(sc.spheres[k], o, d, t) && t > std::max(r.tmin, 0.f) && t > 0.f &&
                        (t < best.t || (t == best.t && gid < best.gid))) {
                        best.t = t;
                        best.gid = gid;
                        best.sphere = (int)k;
                    }
                }
                if (best.sphere >= 0) {
                    hits[i].prim = (uint32_t)best.gid;
                    hits[i].t = best.t;
                } else if (best.gid != ~0ull) {
                    const Instance &ins = sc.instances[best.inst];
                    const Mesh &m = sc.meshes[ins.mesh];
                    V3 a, bb, c;
                    world_tri(ins, m, best.face, a, bb, c);
                    hits[i].prim = (uint32_t)best.gid;
                    hits[i].t = refine_t(o, d, a, bb, c, best.t);
                }
            } else if (mode == ORCL_TRACE_FP64) {
                F64Best best;
                scene_intersect_f64(sc, o, d, r.tmin, r.tmax, best);
                uint32_t fl = 0;
                if (best.gid != ~0ull) {
                    hits[i].prim = (uint32_t)best.gid;
                    hits[i].t = (float)best.t;
                    if (best.minb < g_eps_edge) fl |= ORCL_FLAG_EDGE;
                    if (best.t2 <= best.t * (1.0 + g_eps_tie)) fl |= ORCL_FLAG_TIE;
                    if (best.cosn < g_eps_graze) fl |= ORCL_FLAG_GRAZE;
                }
                if (std::isfinite(best.t_out) && (best.gid == ~0ull || best.t_out <= best.t * (1.0 + g_eps_tie)))
                    fl |= ORCL_FLAG_EDGE;
                if (flags) flags[i] = fl;
                if (t64) t64[i] = best.gid != ~0ull ? best.t : std::numeric_limits<double>::infinity();
            } else {
                continue;
            }
        }
    });
    return 0;
}

int orcl_tri_mt(const float *org, const float *dir, const float *v0, const float *v1, const float *v2, float *t,
                float *u, float *v) {
    V3 a(v0), b(v1), c(v2);
    return tri_mt(V3(org), V3(dir), a, b - a, c - a, *t, *u, *v) ? 1 : 0;
}

int orcl_tri_wt(const float *org, const float *dir, float tmin, const float *v0, const float *v1, const float *v2,
                float *t, float *u, float *v) {
    WtRay r = wt_setup(V3(org), V3(dir));
    float tt, U, V, W, det;
    if (!tri_wt(r, V3(v0), V3(v1), V3(v2), tt, U, V, W, det)) return 0;
    if (!(tt > tmin)) return 0;
    *t = tt;
    *u = V / det;  // weight of v1, as Moller-Trumbore's u
    *v = W / det;  // weight of v2
    return 1;
}

float orcl_refine_t(const float *org, const float *dir, const float *v0, const float *v1, const float *v2) {
    return refine_t(V3(org), V3(dir), V3(v0), V3(v1), V3(v2), std::numeric_limits<float>::quiet_NaN());
}

void orcl_xoshiro_stream(uint64_t seed, uint32_t *out, uint32_t n) {
    Rng r(seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = r.next();
}

void orcl_xoshiro_advance(uint32_t state[4], uint64_t n, uint32_t *out) {
    orcl::Rng r(0);
    for (int k = 0; k < 4; ++k) r.s[k] = state[k];
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t v = r.next();
        if (out) out[i] = v;
    }
    for (int k = 0; k < 4; ++k) state[k] = r.s[k];
}

void orcl_random_unit_stream(uint64_t seed, float *out, uint32_t n) {
    Rng r(seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = r.unit();
}

}  // extern "C"

static void parallel_for(uint64_t n, int num_threads, const std::function<void(uint64_t, uint64_t)> &fn) {
    if (num_threads <= 1 || n < 256) {
        fn(0, n);
        return;
    }
    // dynamic chunks so that uneven rays do not serialise on one thread
    const uint64_t chunk = std::max<uint64_t>(64, n / ((uint64_t)num_threads * 64));
    std::atomic<uint64_t> next(0);
    std::vector<std::thread> pool;
    for (int k = 0; k < num_threads; ++k)
        pool.emplace_back([&]() {
            for (;;) {
                uint64_t b = next.fetch_add(chunk);
                if (b >= n) break;
                fn(b, std::min(n, b + chunk));
            }
        });
    for (auto &th : pool) th.join();
}

// =============================================================================================
// integrator: BasePass / RenderPixel / SamplePixel (CPURenderer.cpp:220-450)
// =============================================================================================
namespace orcl {

struct Hit {
    bool hit = false;
    float t = 0;
    V3 location, normal, tangent;
    V3 gnormal;  // unnormalised geometric normal of the world-space triangle (e1 x e2)
    uint32_t material = 0;
    bool sphere = false;
    float uvx = 0, uvy = 0;
    const Instance *inst = nullptr;
};

template <class T>
static inline T lerp3(const T &v0, const T &v1, const T &v2, float u, float v) {
    return v0 + (v1 - v0) * u + (v2 - v0) * v;  // Utilities.h:114-117
}

// WrapMod (Utilities.h:526-529)
static inline unsigned wrap_mod(int a, unsigned b) { return (static_cast<unsigned>(a % static_cast<int>(b)) + b) % b; }

// SRGBtoLinear (Utilities.h:416-425), one channel
static inline float srgb_to_linear(float c) {
    float cutoff = c < 0.04045f ? 1.f : 0.f;
    float higher = std::pow((c + 0.055f) / 1.055f, 2.4f);
    float lower = c / 12.92f;
    return higher * (1.f - cutoff) + lower * cutoff;
}

// Image::Sample (io/Image.h:112-152): uv scaled by (width - 1, height - 1), Decompose truncates
// (Utilities.h:124-128), bilinear with wrapping, pseudo-sRGB decode after filtering
static V3 texture_sample(const Texture &t, float u, float v) {
    const float px = u * static_cast<float>(t.w - 1), py = v * static_cast<float>(t.h - 1);
    const int iu = static_cast<int>(px), iv = static_cast<int>(py);
    const float fu = px - static_cast<float>(iu), fv = py - static_cast<float>(iv);
    auto at = [&](unsigned x, unsigned y) { return &t.texels[4 * ((size_t)y * t.w + x)]; };
    const float *s00 = at(wrap_mod(iu, t.w), wrap_mod(iv, t.h));
    const float *s10 = at(wrap_mod(iu + 1, t.w), wrap_mod(iv, t.h));
    const float *s01 = at(wrap_mod(iu, t.w), wrap_mod(iv + 1, t.h));
    const float *s11 = at(wrap_mod(iu + 1, t.w), wrap_mod(iv + 1, t.h));
    float r[3];
    for (int c = 0; c < 3; ++c) {
        const float l0 = s00[c] + (s10[c] - s00[c]) * fu;
        const float l1 = s01[c] + (s11[c] - s01[c]) * fu;
        r[c] = l0 + (l1 - l0) * fv;
        if (t.srgb) r[c] = srgb_to_linear(r[c]);
    }
    return V3(r[0], r[1], r[2]);
}

// MaterialRenderProxy getters at a texture coordinate (MaterialRenderProxy.h:117-162)
struct MatEval {
    V3 base, emissive;
    float metallic, roughness;
};
static MatEval eval_material(const orcl_scene &sc, const orcl_material &m, float tu, float tv) {
    MatEval e;
    e.base = V3(m.base_color);
    e.emissive = V3(m.emissive);
    e.metallic = m.metallic;
    e.roughness = m.roughness;
    if (m.base_color_texture >= 0) e.base = cmul(texture_sample(sc.textures[m.base_color_texture], tu, tv), e.base);
    if (m.metallic_roughness_texture >= 0) {
        const V3 t = texture_sample(sc.textures[m.metallic_roughness_texture], tu, tv);
        e.metallic = t.z * m.metallic;
        e.roughness = t.y * m.roughness;
    }
    if (m.emissive_texture >= 0)
        e.emissive = cmul(texture_sample(sc.textures[m.emissive_texture], tu, tv), e.emissive);
    return e;
}

// GetPossibleMajorAxis (Utilities.h:222-234)
static inline V3 major_axis(V3 n) {
    const float inv_sqrt3 = 0.577350269189625764509f;
    if (std::fabs(n.x) < inv_sqrt3) return V3(1, 0, 0);
    if (std::fabs(n.y) < inv_sqrt3) return V3(0, 1, 0);
    return V3(0, 0, 1);
}

// SceneRenderProxy::Intersect<false> + MeshRenderProxy::GetIntersection (MeshRenderProxy.cpp:340-361)
static Hit scene_hit(const orcl_scene &sc, V3 o, V3 d) {
    Candidate c;
    scene_intersect_ref(sc, o, d, c);
    Hit h;
    if (c.sphere >= 0) {
        // SphereRenderProxy::GetIntersection (SphereRenderProxy.h:22-31): uv 0, major-axis tangent
        const Sphere &sp = sc.spheres[c.sphere];
        h.t = c.t;
        h.location = o + d * c.t;
        h.normal = normalized(h.location - V3(sp.c[0], sp.c[1], sp.c[2]));
        h.tangent = major_axis(h.normal);
        h.material = sp.material;
        h.sphere = true;
        h.hit = true;
        return h;
    }
    if (c.inst < 0) return h;
    const Instance &ins = sc.instances[c.inst];
    const Mesh &m = sc.meshes[ins.mesh];
    const uint32_t i0 = m.idx[3 * (size_t)c.face], i1 = m.idx[3 * (size_t)c.face + 1],
                   i2 = m.idx[3 * (size_t)c.face + 2];
    if (!m.uv.empty()) {
        h.uvx = m.uv[2 * i0] + (m.uv[2 * i1] - m.uv[2 * i0]) * c.u + (m.uv[2 * i2] - m.uv[2 * i0]) * c.v;
        h.uvy = m.uv[2 * i0 + 1] + (m.uv[2 * i1 + 1] - m.uv[2 * i0 + 1]) * c.u +
                (m.uv[2 * i2 + 1] - m.uv[2 * i0 + 1]) * c.v;
    }
    V3 n(0, 0, 0);
    if (!m.nrm.empty()) n = lerp3(V3(&m.nrm[3 * i0]), V3(&m.nrm[3 * i1]), V3(&m.nrm[3 * i2]), c.u, c.v);
    V3 tg = kRight;
    float handedness = 1.f;
    if (!m.tan.empty()) {
        tg = lerp3(V3(&m.tan[4 * i0]), V3(&m.tan[4 * i1]), V3(&m.tan[4 * i2]), c.u, c.v);
        handedness = m.tan[4 * i0 + 3] + (m.tan[4 * i1 + 3] - m.tan[4 * i0 + 3]) * c.u +
                     (m.tan[4 * i2 + 3] - m.tan[4 * i0 + 3]) * c.v;
    }
    const orcl_material &mat = sc.materials[ins.material];
    if (mat.normal_texture >= 0) {
        // TangentSpaceToWorldSpace (Utilities.h:398-404) with GetNormal = sample * 2 - 1
        const V3 tn = texture_sample(sc.textures[mat.normal_texture], h.uvx, h.uvy) * 2.f - kOnes;
        const V3 bit = cross(n, tg) * handedness;
        n = normalized(tg * tn.x + bit * tn.y + n * tn.z);
    }
    // TransformDirectionTangentSpace on the inverse = the forward 3x3 (Transform.h:105-108)
    h.normal = normalized(ins.to_world.dir(n));
    h.tangent = ins.to_world.dir(tg);
    {
        const V3 w0 = ins.to_world.point(m.vtx(i0)), w1 = ins.to_world.point(m.vtx(i1)),
                 w2 = ins.to_world.point(m.vtx(i2));
        h.gnormal = cross(w1 - w0, w2 - w0);
    }
    h.t = c.t;
    h.location = o + d * c.t;  // Intersection::Update -> ray.At(t)
    h.inst = &ins;
    h.material = ins.material;
    h.hit = true;
    return h;
}

// GetLocalAxisFromNormal(normal, major_axis) (Utilities.h:236-244)
static inline void local_axes(V3 n, V3 major, V3 &u, V3 &v, V3 &w) {
    u = normalized(cross(n, major));
    v = cross(n, u);
    w = n;
}

static inline float schlick(float c, float r0) { return r0 + (1.f - r0) * std::pow((1.f - c), 5.f); }
static inline V3 schlick3(float c, V3 r0) {
    float p = std::pow((1.f - c), 5.f);
    return r0 + (kOnes - r0) * p;
}

static inline float smith_ggx_correlated(float cos_o, float cos_i, float roughness) {
    float a = roughness * roughness;
    float a2 = a * a;
    float ggx_i = cos_o * std::sqrt((-cos_i * a2 + cos_i) * cos_i + a2);
    float ggx_o = cos_i * std::sqrt((-cos_o * a2 + cos_o) * cos_o + a2);
    return 2.f * cos_o * cos_i / (ggx_o + ggx_i);
}

static inline float geometry_schlick_ggx(float cos_theta, float roughness) {
    float a = roughness * roughness;
    float a2 = a * a;
    float tan2_v = (1.f - cos_theta * cos_theta) / (cos_theta * cos_theta);
    return 2.f / (1.f + std::sqrt(1.f + a2 * tan2_v));
}

// DistributionVn::Sample (Sampler.h:133-165)
static V3 sample_vndf(Rng &rng, V3 w_o, float roughness) {
    float a = roughness * roughness;
    float u1 = rng.unit();
    float u2 = rng.unit();
    V3 v = normalized(V3(w_o.x * a, w_o.y * a, w_o.z));
    V3 t1 = (v.z < 0.999f) ? normalized(cross(v, kUp)) : kRight;
    V3 t2 = cross(v, t1);
    float r = std::sqrt(u1);
    float phi = 2.f * kPi * u2;
    float p1 = r * std::cos(phi);
    float p2 = r * std::sin(phi);
    float s = 0.5f * (1.f + v.z);
    p2 = (1.f - s) * std::sqrt(1.f - p1 * p1) + s * p2;
    V3 n = p1 * t1 + p2 * t2 + std::sqrt(std::max(0.f, 1.f - p1 * p1 - p2 * p2)) * v;
    return normalized(V3(a * n.x, a * n.y, std::max(0.f, n.z)));
}

struct Sample {
    bool valid = false;
    V3 throughput, wi;
};

// SpecularBxDF::Sample (BxDF.h:48-109)
static Sample specular_sample(Rng &rng, V3 wo, V3 base, float roughness, float metallic) {
    Sample s;
    V3 wm = sample_vndf(rng, wo, roughness);
    V3 wi = wm * (dot(wo, wm) * 2.f) - wo;  // Reflect (Utilities.h:206-209)
    float cos_o = std::max(std::abs(wo.z), 0.f);
    float cos_i = std::max(std::abs(wi.z), 0.f);
    float cos_m = std::max(std::abs(wm.z), 0.f);
    float local_cos_i = dot(wi, wm);
    V3 fres;
    if (cos_o > kEps && cos_i > kEps && cos_m > kEps) {
        if (rng.unit() < metallic) {
            fres = schlick3(local_cos_i, base);
            s.valid = true;
        } else {
            float f = schlick(local_cos_i, 0.04f);
            if (rng.unit() < f) {
                fres = kOnes;
                s.valid = true;
            }
        }
    }
    if (!s.valid) return s;
    float occlusion = smith_ggx_correlated(cos_o, cos_i, roughness);
    float normalizer = geometry_schlick_ggx(cos_o, roughness) + kEps;
    s.throughput = fres * occlusion / normalizer;
    s.wi = wi;
    return s;
}

// LambertianBxDF::Sample (BxDF.h:32-46) with CosineWeightedHemiSphere (Sampler.h:79-95)
static Sample lambert_sample(Rng &rng, V3 base) {
    Sample s;
    s.valid = true;
    float x, y;
    rng.disk(x, y);
    float z = std::sqrt(std::max(0.f, 1.0f - (x * x + y * y)));
    s.wi = V3(x, y, z);
    s.throughput = base;
    return s;
}

// FrDielectric (Utilities.h:301-324)
static float fr_dielectric(float cos_i, float eta_i, float eta_t) {
    if (!(cos_i > 0.f)) {
        std::swap(eta_i, eta_t);
        cos_i = -cos_i;
    }
    float sin_i = std::sqrt(1 - cos_i * cos_i);
    float ref_idx = eta_i / eta_t;
    float sin_t = ref_idx * sin_i;
    if (sin_t >= 1.f) return 1.f;
    float r0 = (1.f - ref_idx) / (1.f + ref_idx);
    r0 = r0 * r0;
    return schlick(cos_i, r0);
}

// DieletricBxDF::Sample (BxDF.h:112-141)
static Sample dielectric_sample(Rng &rng, V3 wo, V3 base, float eta) {
    Sample s;
    float fres = fr_dielectric(wo.z, 1.0f, eta);
    s.valid = true;
    if (rng.unit() < fres) {
        s.wi = V3(-wo.x, -wo.y, wo.z);
    } else {
        bool entering = wo.z > 0;
        float e = entering ? 1.0f / eta : eta / 1.0f;
        float cos_i = wo.z;
        float sin_i2 = 1.f - cos_i * cos_i;
        float sin_t2 = e * e * sin_i2;
        float cos_t = std::sqrt(1 - sin_t2) * (cos_i > 0 ? -1.f : 1.f);
        s.wi = normalized(V3(-e * wo.x, -e * wo.y, cos_t));
    }
    s.throughput = base;
    return s;
}

// MaterialRenderProxy::SampleSurface for Pbr / Dielectric proxies
static V3 sample_surface(Rng &rng, const orcl_material &mat, const MatEval &me, V3 ray_dir, V3 normal, V3 tangent,
                         V3 &w_i, bool *diffuse = nullptr) {
    V3 base = me.base;
    V3 wo = -ray_dir;
    V3 u, v, w;
    local_axes(normal, tangent, u, v, w);
    V3 lwo(dot(wo, u), dot(wo, v), dot(wo, w));
    Sample s;
    if (mat.type == ORCL_MAT_DIELECTRIC) {
        s = dielectric_sample(rng, lwo, base, mat.eta);
    } else {
        s = specular_sample(rng, lwo, base, me.roughness, me.metallic);
        if (!s.valid) {
            s = lambert_sample(rng, base);
            if (diffuse) *diffuse = true;
        }
    }
    V3 u2, v2, w2;
    local_axes(normal, tangent, u2, v2, w2);
    w_i = normalized(u2 * s.wi.x + v2 * s.wi.y + w2 * s.wi.z);
    return s.throughput;
}

struct PixelResult {
    V3 color;
    V3 normal;
    float valid = 1.f;
};

static std::atomic<uint64_t> g_rays{0};

// Image2DCube::DirectionToTextureCoordinate + Sample (io/Image.cpp:298-357)
static V3 sky_cube_sample(const orcl_scene &sc, V3 dir) {
    const float x = dir.x, y = dir.y, z = dir.z;
    const float ax = std::fabs(x), ay = std::fabs(y), az = std::fabs(z);
    int face;
    float u, v;
    if (ax >= ay && ax >= az) {
        if (x > 0) { face = 0; u = -z / x; v = -y / x; } else { face = 1; u = z / -x; v = -y / -x; }
    } else if (ay >= ax && ay >= az) {
        if (y > 0) { face = 2; u = x / y; v = z / y; } else { face = 3; u = x / -y; v = -z / -y; }
    } else {
        if (z > 0) { face = 4; u = x / z; v = -y / z; } else { face = 5; u = -x / -z; v = -y / -z; }
    }
    u = (u + 1.f) * 0.5f;
    v = (v + 1.f) * 0.5f;
    return texture_sample(sc.textures[sc.sky_faces[face]], u, v);
}

// area pdf of a point on an emitter of this material under the emitter sampling below
static float light_pdf_area(const orcl_scene &sc, const orcl_material &m) {
    return (m.emissive[0] + m.emissive[1] + m.emissive[2]) / 3.f / sc.emitter_cdf.back();
}

// Next-event estimation as the device does it (no reference counterpart): one emissive triangle
// picked by area x mean radiance, one uniform point on it, Lambert BRDF base/pi, two-sided
// emitters, occlusion over (1e-4, dist - 1e-4) of the segment. Returns the contribution before
// visibility and sets `visible`.
static V3 nee_sample(const orcl_scene &sc, Rng &nr, V3 loc, V3 normal, V3 base, V3 T, bool &visible) {
    visible = false;
    const float total = sc.emitter_cdf.back();
    const float r0 = nr.unit() * total;
    size_t lo = 0, hi = sc.emitters.size() - 1;
    while (lo < hi) {
        const size_t mid = (lo + hi) / 2;
        if (sc.emitter_cdf[mid] < r0) lo = mid + 1; else hi = mid;
    }
    const Emitter &e = sc.emitters[lo];
    const float w_tri = sc.emitter_cdf[lo] - (lo ? sc.emitter_cdf[lo - 1] : 0.f);
    const float r1 = nr.unit(), r2 = nr.unit();
    const V3 p0(&e.p[0]), p1(&e.p[3]), p2(&e.p[6]);
    const float su = std::sqrt(r1);
    const float b0 = 1.f - su, b1 = r2 * su;
    const V3 y = p0 * b0 + p1 * b1 + p2 * (1.f - b0 - b1);
    const V3 ng = cross(p1 - p0, p2 - p0);
    const float area2 = std::sqrt(sqnorm(ng));
    const V3 dl = y - loc;
    const float dist2 = sqnorm(dl);
    const float dist = std::sqrt(dist2);
    const V3 wl = dl / dist;
    const float cos_x = dot(normal, wl);
    const float cos_y = std::fabs(dot(ng, wl)) / area2;
    const float bu = b1, bv = 1.f - b0 - b1;
    const MatEval lme = eval_material(sc, sc.materials[e.material], e.uv[0] + (e.uv[2] - e.uv[0]) * bu + (e.uv[4] - e.uv[0]) * bv,
                                      e.uv[1] + (e.uv[3] - e.uv[1]) * bu + (e.uv[5] - e.uv[1]) * bv);
    if (!(norm(lme.emissive) > kEps) || !(cos_x > 0.f) || !(cos_y > 0.f) || !(dist > 2e-4f) || !(w_tri > 0.f))
        return V3(0, 0, 0);
    const float area = 0.5f * area2;
    // balance heuristic against the cosine-sampled direction
    const float p_light = light_pdf_area(sc, sc.materials[e.material]) * dist2 / cos_y;
    const float p_bsdf = cos_x / kPi;
    const float g = cos_x * cos_y / dist2 * area * total / w_tri / kPi * (p_light / (p_light + p_bsdf));
    Candidate c;
    scene_intersect_ref(sc, loc + wl * kTolerance, wl, c);
    visible = !(c.inst >= 0 && c.t < dist - 2.f * kTolerance);
    return cmul(cmul(T, base), lme.emissive) * g;
}

// VisualizeInteger (Utilities.h:93-112): a 7-colour high-contrast palette / 255, black from 7 on
static V3 visualize_integer(unsigned number) {
    static const float palette[7][3] = {{137, 49, 239}, {242, 202, 25}, {255, 0, 189}, {0, 87, 233},
                                        {135, 233, 17}, {225, 24, 69},  {255, 255, 255}};
    if (number >= 7) return kZeros;
    return V3(palette[number][0], palette[number][1], palette[number][2]) / 255.0f;
}

static PixelResult sample_pixel(const orcl_scene &sc, const orcl_camera &cam, const orcl_frame &fr, Rng &rng,
                                float u, float v, uint64_t &nrays, Rng *nee_rng = nullptr) {
    // SetupViewRay (CPURenderer.cpp:220-238)
    float nx, ny;
    rng.disk(nx, ny);
    nx = nx * cam.lens_radius;
    ny = ny * cam.lens_radius;
    V3 lens = V3(cam.right) * nx + V3(cam.up) * ny;
    V3 o = V3(cam.position) + lens;
    V3 focus = V3(cam.lower_left) + V3(cam.max_u) * u + V3(cam.max_v) * v;
    V3 d = normalized(focus - o);

    V3 T = kOnes;
    PixelResult res;
    unsigned bounce = 0;
    bool skip_emission = false;  // the previous vertex did next-event estimation
    float prev_pdf = 0.f;
    const bool nee = nee_rng && !sc.emitters.empty();
    for (; bounce < fr.max_bounce; bounce++) {
        Hit h = scene_hit(sc, o, d);
        nrays++;
        if (!h.hit) {
            if (fr.sky_enabled) {
                V3 sky;
                if (sc.sky_faces[0] >= 0) {
                    sky = sky_cube_sample(sc, d);
                } else {
                    float t = 0.5f * d.z + 0.5f;
                    sky = kOnes + (V3(fr.sky_color) - kOnes) * t;
                }
                res.color = res.color + cmul(sky, T);
            }
            break;
        }
        const orcl_material &mat = sc.materials[h.material];
        const MatEval me = eval_material(sc, mat, h.uvx, h.uvy);
        if (bounce == 0) res.normal = h.normal;
        if (norm(me.emissive) > kEps) {
            if (!skip_emission || h.sphere) {  // spheres are not among the sampled emitters
                res.color = res.color + cmul(me.emissive, T);
            } else {
                const float cos_y = std::fabs(dot(h.gnormal, d)) / std::sqrt(sqnorm(h.gnormal));
                const float p_light = light_pdf_area(sc, mat) * h.t * h.t / cos_y;
                if (cos_y > 0.f) res.color = res.color + cmul(me.emissive, T) * (prev_pdf / (prev_pdf + p_light));
            }
            break;
        }
        V3 wi;
        bool diffuse = false;
        V3 f = sample_surface(rng, mat, me, d, h.normal, h.tangent, wi, &diffuse);
        switch (fr.debug_mode) {
            case 1:  // Debug: the reference's debug colour, zero unless a debug pixel sets it
                res.color = kZeros;
                return res;
            case 3:  // Normal
                res.color = (h.normal + kOnes) * 0.5f;
                return res;
            case 4:  // RayDirection
                res.color = (wi + kOnes) * 0.5f;
                return res;
            case 7:
                res.color = kOnes * me.metallic;
                return res;
            case 8:
                res.color = kOnes * me.roughness;
                return res;
            case 9:
                res.color = me.base;
                return res;
            case 10:
                res.color = me.emissive;
                return res;
            case 11:
                res.color = kOnes * (dot(h.location - V3(cam.position), V3(cam.front)) / cam.far_plane);
                return res;
            default:
                break;
        }
        const bool last = bounce + 1 >= fr.max_bounce;
        if (nee && diffuse && !last) {
            bool visible = false;
            const V3 contrib = nee_sample(sc, *nee_rng, h.location, h.normal, me.base, T, visible);
            if (visible) res.color = res.color + contrib;
        }
        skip_emission = nee && diffuse;
        if (skip_emission) prev_pdf = std::max(dot(wi, h.normal), 0.f) / kPi;
        T = cmul(T, f);
        if (sqnorm(T) < kEps) {
            res.valid = -1.f;
            break;
        }
        // Russian roulette (CPURenderer.cpp:350-364)
        if (bounce >= 3) {
            float p = maxc(T);
            p = std::min(std::max(p, 0.05f), 1.0f);
            const float r = rng.unit();
            if (!fr.rr_off) {
                if (r > p) break;
                T = T / p;
            }
        }
        o = h.location + wi * kTolerance;
        d = wi;
    }
    // debug modes that rewrite the colour after the loop (CPURenderer.cpp:370-391); `bounce` is the
    // segment the path ended in, max_bounce when the loop ran out
    switch (fr.debug_mode) {
        case 1:  // Debug
            res.color = kZeros;
            break;
        case 2:  // RayDepth: VisualizeInteger (Utilities.h:93-112)
            res.color = visualize_integer(bounce);
            break;
        case 5:  // IndirectLighting
            if (bounce <= 1) res.color = kZeros;
            break;
        case 6:  // DirectionalLighting
            if (bounce > 1) res.color = kZeros;
            break;
        default:
            break;
    }
    return res;
}

}  // namespace orcl

extern "C" {

int orcl_render_frame(const orcl_scene *scene, const orcl_camera *camera, const orcl_frame *frame, int num_threads,
                      float *rgba, float *normal) {
    if (!scene || !camera || !frame || !rgba) return -1;
    const orcl_frame fr = *frame;
    const orcl_camera cam = *camera;
    const float pw = 1.f / static_cast<float>(fr.width - 1);
    const float ph = 1.f / static_cast<float>(fr.height - 1);
    g_rays = 0;
    // one task per row (CPURenderer.cpp:435-449)
    parallel_for(fr.height, num_threads, [&](uint64_t jb, uint64_t je) {
        uint64_t nrays = 0;
        for (uint64_t j = jb; j < je; ++j) {
            for (uint32_t i = 0; i < fr.width; ++i) {
                const uint32_t seed32 = (uint32_t)j * fr.width + i + fr.frame_seed * fr.width * fr.height;
                uint64_t seed = seed32;
                if (fr.seed_hi) {
                    seed = (uint64_t)j * fr.width + i + (uint64_t)fr.frame_seed * fr.width * fr.height;
                }
                Rng rng(seed);
                Rng nee_rng(seed ^ 0x5bd1e9955bd1e995ull);
                float u = (static_cast<float>(i) + rng.unit()) * pw;
                float v = (static_cast<float>(j) + rng.unit()) * ph;
                PixelResult r = sample_pixel(*scene, cam, fr, rng, u, v, nrays, fr.nee ? &nee_rng : nullptr);
                const size_t p = (size_t)j * fr.width + i;
                const float limit = fr.output_limit > 0.f ? fr.output_limit : kOutputLimit;
                rgba[4 * p + 0] = std::min(r.color.x, limit);
                rgba[4 * p + 1] = std::min(r.color.y, limit);
                rgba[4 * p + 2] = std::min(r.color.z, limit);
                rgba[4 * p + 3] = r.valid;
                if (normal) {
                    normal[3 * p + 0] = r.normal.x;
                    normal[3 * p + 1] = r.normal.y;
                    normal[3 * p + 2] = r.normal.z;
                }
            }
        }
        g_rays += nrays;
    });
    return 0;
}

uint64_t orcl_last_render_rays(void) { return g_rays.load(); }

// CPURenderer::DenoisePass temporal part (CPURenderer.cpp:527-542): acc = new + (acc - new) * n/(n+spp)
void orcl_accumulate(const float *rgba, float *accum, uint64_t num_pixels, uint32_t cumulated, uint32_t spp) {
    const float ma = static_cast<float>(cumulated) / static_cast<float>(cumulated + spp);
    for (uint64_t i = 0; i < 4 * num_pixels; ++i) accum[i] = rgba[i] + (accum[i] - rgba[i]) * ma;
}

// SpatialDenoisePixel over the whole frame (CPURenderer.cpp:452-504, 511-525), in the order a single
// thread runs it: rows j = 0..H-1, columns i = 0..W-1. Every valid, non-sky pixel copies its colour
// into the first of its 8 neighbours (fixed direction order) that is in the image, invalid
// (w <= 0) and has a normal within dot >= 0.99 of its own; later writers overwrite earlier ones.
// (The reference runs rows in parallel, so where two rows target the same pixel its winner is not
// fixed; this order is the single-threaded one.)
void orcl_spatial_fill(const float *rgba, const float *normal, float *out, uint32_t w, uint32_t h) {
    static const int dirs[8][2] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}, {-1, -1}, {1, -1}, {-1, 1}, {1, 1}};
    std::memcpy(out, rgba, sizeof(float) * 4 * (size_t)w * h);
    for (uint32_t j = 0; j < h; ++j) {
        for (uint32_t i = 0; i < w; ++i) {
            const size_t p = (size_t)j * w + i;
            const float *n = normal + 3 * p;
            const bool valid = rgba[4 * p + 3] > 0.f;
            const bool sky = valid && n[0] == 0.f && n[1] == 0.f && n[2] == 0.f;
            if (!valid || sky) continue;
            for (int k = 0; k < 8; ++k) {
                const int si = (int)i + dirs[k][0], sj = (int)j + dirs[k][1];
                if (si < 0 || si >= (int)w || sj < 0 || sj >= (int)h) continue;
                const size_t q = (size_t)sj * w + si;
                if (rgba[4 * q + 3] > 0.f) continue;
                const float *m = normal + 3 * q;
                const float sim = (m[0] * n[0] + m[1] * n[1]) + m[2] * n[2];
                if (sim < 0.99f) continue;
                out[4 * q] = rgba[4 * p];
                out[4 * q + 1] = rgba[4 * p + 1];
                out[4 * q + 2] = rgba[4 * p + 2];
                out[4 * q + 3] = 1.f;
                break;
            }
        }
    }
}

void orcl_camera_setup(const float position[3], const float right[3], const float front[3], const float up[3],
                       float vertical_fov, float focus_distance, float aperture_radius, float aspect,
                       orcl_camera *c) {
    float h = std::tan(vertical_fov / 2.f);
    float height = 2.f * h * focus_distance;
    float width = aspect * height;
    V3 R(right), F(front), U(up), P(position);
    V3 mu = R * width, mv = U * height;
    V3 ll = P + F * focus_distance - mu * .5f - mv * .5f;
    for (int a = 0; a < 3; ++a) {
        c->position[a] = P[a];
        c->right[a] = R[a];
        c->up[a] = U[a];
        c->front[a] = F[a];
        c->max_u[a] = mu[a];
        c->max_v[a] = mv[a];
        c->lower_left[a] = ll[a];
    }
    c->lens_radius = aperture_radius;
    c->far_plane = 1000.f;
}

}  // extern "C"

// =============================================================================================
// learned radiance cache, FP64 (no reference counterpart): multiresolution hash grid
// (Mueller et al. 2022, primes 1, 2654435761, 805459861) and a bias-free ReLU MLP trained with
// a relative L2 loss (Mueller et al. 2021) and Adam (Kingma & Ba 2015).
// =============================================================================================
namespace orcl {
static inline uint32_t level_resolution(const orcl_cache_config *cfg, uint32_t level) {
    return (uint32_t)std::floor((double)cfg->base_resolution * std::pow((double)cfg->per_level_scale, (double)level));
}
}  // namespace orcl

extern "C" {

uint32_t orcl_hashgrid_index(const orcl_cache_config *cfg, uint32_t level, uint32_t x, uint32_t y, uint32_t z) {
    const uint32_t res = orcl::level_resolution(cfg, level);
    const uint64_t T = 1ull << cfg->log2_table;
    const uint64_t side = (uint64_t)res + 1;
    if (side * side * side <= T) return (uint32_t)(x + side * (y + side * z));
    const uint32_t h = (x * 1u) ^ (y * 2654435761u) ^ (z * 805459861u);
    return h & (uint32_t)(T - 1);
}

void orcl_hashgrid_encode(const orcl_cache_config *cfg, const float *table, const float *points, uint32_t n,
                          double *features) {
    const uint32_t L = cfg->levels, F = cfg->features;
    const uint64_t T = 1ull << cfg->log2_table;
    for (uint32_t i = 0; i < n; ++i) {
        for (uint32_t l = 0; l < L; ++l) {
            const uint32_t res = orcl::level_resolution(cfg, l);
            uint32_t cell[3];
            double frac[3];
            for (int a = 0; a < 3; ++a) {
                // cell selection in float, as the device computes it
                float p = std::min(std::max(points[3 * (size_t)i + a], 0.f), 1.f);
                float pos = p * (float)res;
                float c = std::floor(pos);
                c = std::min(std::max(c, 0.f), (float)(res - 1));
                cell[a] = (uint32_t)c;
                frac[a] = (double)(pos - c);
            }
            for (uint32_t f = 0; f < F; ++f) features[(size_t)i * L * F + l * F + f] = 0.0;
            for (int k = 0; k < 8; ++k) {
                uint32_t cx = cell[0] + (k & 1), cy = cell[1] + ((k >> 1) & 1), cz = cell[2] + ((k >> 2) & 1);
                double w = ((k & 1) ? frac[0] : 1.0 - frac[0]) * (((k >> 1) & 1) ? frac[1] : 1.0 - frac[1]) *
                           (((k >> 2) & 1) ? frac[2] : 1.0 - frac[2]);
                uint32_t idx = orcl_hashgrid_index(cfg, l, cx, cy, cz);
                for (uint32_t f = 0; f < F; ++f)
                    features[(size_t)i * L * F + l * F + f] += w * (double)table[((size_t)l * T + idx) * F + f];
            }
        }
    }
}

double orcl_mlp_loss_grad(const orcl_cache_config *cfg, const double *weights, const double *features,
                          const double *targets, uint32_t n, double *outputs, double *dweights, double *dfeatures) {
    const uint32_t in = cfg->inputs ? cfg->inputs : cfg->levels * cfg->features, H = cfg->hidden,
                   NL = cfg->hidden_layers, O = cfg->outputs;
    // layer shapes: [H x in], (NL-1) x [H x H], [O x H]
    std::vector<uint32_t> rows, cols;
    rows.push_back(H);
    cols.push_back(in);
    for (uint32_t k = 1; k < NL; ++k) {
        rows.push_back(H);
        cols.push_back(H);
    }
    rows.push_back(O);
    cols.push_back(H);
    const size_t nlayers = rows.size();
    std::vector<size_t> off(nlayers + 1, 0);
    for (size_t k = 0; k < nlayers; ++k) off[k + 1] = off[k] + (size_t)rows[k] * cols[k];
    if (dweights) std::fill(dweights, dweights + off[nlayers], 0.0);
    double loss = 0.0;
    const double inv = 1.0 / ((double)n * O);
    std::vector<std::vector<double>> act(nlayers + 1);
    for (uint32_t s = 0; s < n; ++s) {
        act[0].assign(features + (size_t)s * in, features + (size_t)s * in + in);
        for (size_t k = 0; k < nlayers; ++k) {
            act[k + 1].assign(rows[k], 0.0);
            for (uint32_t r = 0; r < rows[k]; ++r) {
                double acc = 0.0;
                for (uint32_t c = 0; c < cols[k]; ++c) acc += weights[off[k] + (size_t)r * cols[k] + c] * act[k][c];
                act[k + 1][r] = (k + 1 < nlayers) ? std::max(acc, 0.0) : acc;
            }
        }
        std::vector<double> g(O);
        for (uint32_t o = 0; o < O; ++o) {
            double y = act[nlayers][o], tg = targets[(size_t)s * O + o];
            if (outputs) outputs[(size_t)s * O + o] = y;
            double den = y * y + 0.01;  // the denominator is treated as a constant (stop-gradient)
            loss += (y - tg) * (y - tg) / den * inv;
            g[o] = 2.0 * (y - tg) / den * inv;
        }
        for (size_t k = nlayers; k-- > 0;) {
            std::vector<double> gin(cols[k], 0.0);
            for (uint32_t r = 0; r < rows[k]; ++r) {
                for (uint32_t c = 0; c < cols[k]; ++c) {
                    if (dweights) dweights[off[k] + (size_t)r * cols[k] + c] += g[r] * act[k][c];
                    gin[c] += weights[off[k] + (size_t)r * cols[k] + c] * g[r];
                }
            }
            if (k > 0)
                for (uint32_t c = 0; c < cols[k]; ++c) gin[c] = act[k][c] > 0.0 ? gin[c] : 0.0;
            g.swap(gin);
        }
        if (dfeatures)
            for (uint32_t c = 0; c < in; ++c) dfeatures[(size_t)s * in + c] = g[c];
    }
    return loss;
}

void orcl_adam(double *param, const double *grad, double *m, double *v, uint64_t count, double lr, double beta1,
               double beta2, double eps, uint32_t step) {
    const double c1 = 1.0 - std::pow(beta1, (double)step), c2 = 1.0 - std::pow(beta2, (double)step);
    for (uint64_t i = 0; i < count; ++i) {
        m[i] = beta1 * m[i] + (1.0 - beta1) * grad[i];
        v[i] = beta2 * v[i] + (1.0 - beta2) * grad[i] * grad[i];
        const double mh = m[i] / c1, vh = v[i] / c2;
        param[i] -= lr * mh / (std::sqrt(vh) + eps);
    }
}

}  // extern "C"

// n samples of one lobe in the local frame (normal +z), all from one xoshiro stream seeded with `seed`:
// kind 0 Lambertian, 1 specular (GGX VNDF, metallic / Fresnel choice), 2 dielectric, 3 the VNDF
// microfacet normal alone. wi: n x 3; thr: n x 3 (may be null); valid: n (may be null).
extern "C" void orcl_bxdf_sample(int kind, const float *wo, const float *base, float roughness, float metallic,
                                 float eta, uint64_t seed, uint32_t n, float *wi, float *thr, int32_t *valid) {
    orcl::Rng rng(seed);
    const orcl::V3 o(wo[0], wo[1], wo[2]), b(base[0], base[1], base[2]);
    for (uint32_t k = 0; k < n; ++k) {
        orcl::Sample s;
        if (kind == 0) {
            s = orcl::lambert_sample(rng, b);
        } else if (kind == 1) {
            s = orcl::specular_sample(rng, o, b, roughness, metallic);
        } else if (kind == 2) {
            s = orcl::dielectric_sample(rng, o, b, eta);
        } else {
            s.valid = true;
            s.wi = orcl::sample_vndf(rng, o, roughness);
            s.throughput = b;
        }
        wi[3 * (size_t)k] = s.wi.x, wi[3 * (size_t)k + 1] = s.wi.y, wi[3 * (size_t)k + 2] = s.wi.z;
        if (thr) thr[3 * (size_t)k] = s.throughput.x, thr[3 * (size_t)k + 1] = s.throughput.y,
                 thr[3 * (size_t)k + 2] = s.throughput.z;
        if (valid) valid[k] = s.valid ? 1 : 0;
    }
}

// The shading frame SampleSurface builds (u, v, w) from a normal and a tangent
extern "C" void orcl_local_axes(const float *normal, const float *tangent, float *u, float *v, float *w) {
    orcl::V3 a, b, c;
    orcl::local_axes(orcl::V3(normal[0], normal[1], normal[2]), orcl::V3(tangent[0], tangent[1], tangent[2]), a, b,
                     c);
    u[0] = a.x, u[1] = a.y, u[2] = a.z;
    v[0] = b.x, v[1] = b.y, v[2] = b.z;
    w[0] = c.x, w[1] = c.y, w[2] = c.z;
}

// ---- textured sky light cook (SkyLightCookJob::Execute, SkyLight.cpp:83-135) ----
// float -> binary16 -> float with round to nearest even: 11 significant bits for normal halves
// (|x| >= 2^-14), a quantum of 2^-24 below, infinity past the largest half (65504)
static float to_half_and_back(float x) {
    if (!std::isfinite(x) || x == 0.f) return x;
    const double a = std::fabs((double)x);
    double r;
    if (a < std::ldexp(1.0, -14)) {
        r = std::nearbyint(std::ldexp(a, 24)) * std::ldexp(1.0, -24);
    } else {
        int e;
        const double m = std::frexp(a, &e);  // a = m 2^e, m in [0.5, 1)
        r = std::ldexp(std::nearbyint(std::ldexp(m, 11)), e - 11);
    }
    if (r > 65504.0) r = std::numeric_limits<double>::infinity();
    return static_cast<float>(std::copysign(r, (double)x));
}

extern "C" void orcl_sky_cook(const float *equirect, uint32_t w, uint32_t h, uint32_t face_size, float *faces) {
    Texture src;
    src.w = w;
    src.h = h;
    src.texels.assign(equirect, equirect + 4 * (size_t)w * h);
    for (uint32_t face = 0; face < 6; ++face) {
        for (uint32_t i = 0; i < face_size; ++i) {
            for (uint32_t j = 0; j < face_size; ++j) {
                const float u = (static_cast<float>(i) + 0.5f) / static_cast<float>(face_size) * 2.f - 1.f;
                const float v = (static_cast<float>(j) + 0.5f) / static_cast<float>(face_size) * 2.f - 1.f;
                // Image2DCube::TextureCoordinateToDirection (io/Image.h:287-315): components by
                // negation, so a zero keeps its sign (atan2(-0, x < 0) = -pi picks the map's left edge)
                V3 d;
                switch (face) {
                    case 0: d = V3(1.f, -v, -u); break;
                    case 1: d = V3(-1.f, -v, u); break;
                    case 2: d = V3(u, 1.f, v); break;
                    case 3: d = V3(u, -1.f, -v); break;
                    case 4: d = V3(u, -v, 1.f); break;
                    default: d = V3(-u, -v, -1.f); break;
                }
                d = normalized(d);
                // CartesianToEquirectangular (Utilities.h:510-518)
                const float eu = std::atan2(d.y, d.x) * (kInvPi * 0.5f) + 0.5f;
                const float ev = std::asin(-d.z) * kInvPi + 0.5f;
                const V3 c = texture_sample(src, eu, ev);
                float *o = faces + 4 * (((size_t)face * face_size + j) * face_size + i);
                o[0] = to_half_and_back(c.x);
                o[1] = to_half_and_back(c.y);
                o[2] = to_half_and_back(c.z);
                o[3] = 1.f;
            }
        }
    }
}

extern "C" float orcl_half_round(float x) { return to_half_and_back(x); }

extern "C" void orcl_texture_sample(const orcl_scene *scene, int index, float u, float v, float *out) {
    const orcl::V3 r = orcl::texture_sample(scene->textures[index], u, v);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}
