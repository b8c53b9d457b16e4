// Host BVH builder for the device scene (see bvh_build.h).
#include "bvh_build.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <thread>

namespace spk {
namespace {

struct Aabb {
    float lo[3], hi[3];
    Aabb() {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::numeric_limits<float>::infinity();
            hi[a] = -std::numeric_limits<float>::infinity();
        }
    }
    void grow(const Aabb &b) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    void grow_point(const float *p) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    bool empty() const { return !(lo[0] <= hi[0]); }
    double area() const {
        if (empty()) return 0.0;
        double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

struct BinNode {
    Aabb box;
    uint32_t left = 0;   // inner: left child (right = left + 1); leaf: slot in the index array
    uint32_t count = 0;  // triangles in the subtree
    bool leaf = false;
};

class BinaryBuilder {
public:
    BinaryBuilder(const std::vector<Aabb> &boxes, const std::vector<float> &cent, unsigned threads)
        : boxes_(boxes), cent_(cent), threads_(std::max(1u, threads)) {}

    void build(std::vector<BinNode> &nodes, std::vector<uint32_t> &ids) {
        const uint32_t n = static_cast<uint32_t>(boxes_.size());
        ids.resize(n);
        for (uint32_t i = 0; i < n; ++i) ids[i] = i;
        nodes.assign(std::max<size_t>(1, 2 * (size_t)n - 1), BinNode());
        next_.store(1);
        ids_ = ids.data();
        nodes_ = nodes.data();
        active_.store(1);
        if (n > 0) recurse(0, 0, n, 0);
        nodes.resize(next_.load());
    }

private:
    static constexpr int kBins = 32;

    void recurse(uint32_t node, uint32_t begin, uint32_t end, int depth) {
        for (;;) {
            BinNode &nd = nodes_[node];
            Aabb nb, cb;
            for (uint32_t i = begin; i < end; ++i) {
                nb.grow(boxes_[ids_[i]]);
                cb.grow_point(&cent_[3 * (size_t)ids_[i]]);
            }
            nd.box = nb;
            nd.count = end - begin;
            if (end - begin == 1) {
                nd.leaf = true;
                nd.left = begin;
                return;
            }
            uint32_t mid = split(begin, end, cb);
            uint32_t l = next_.fetch_add(2);
            nd.leaf = false;
            nd.left = l;
            // spawn the larger half on another thread near the top of the tree
            const bool spawn = (end - begin) > 8192 && active_.load() < (int)threads_;
            if (spawn) {
                active_.fetch_add(1);
                auto fut = std::async(std::launch::async, [this, l, begin, mid, depth]() {
                    recurse(l, begin, mid, depth + 1);
                    active_.fetch_sub(1);
                });
                recurse(l + 1, mid, end, depth + 1);
                fut.get();
                return;
            }
            recurse(l, begin, mid, depth + 1);
            node = l + 1;
            begin = mid;
            ++depth;
        }
    }

    uint32_t split(uint32_t begin, uint32_t end, const Aabb &cb) {
        const uint32_t n = end - begin;
        int best_axis = -1, best_bin = -1;
        double best = std::numeric_limits<double>::infinity();
        for (int a = 0; a < 3; ++a) {
            const float ext = cb.hi[a] - cb.lo[a];
            if (!(ext > 0.f)) continue;
            const float k = kBins / ext;
            Aabb bb[kBins];
            uint32_t bc[kBins] = {0};
            for (uint32_t i = begin; i < end; ++i) {
                const uint32_t p = ids_[i];
                int b = (int)((cent_[3 * (size_t)p + a] - cb.lo[a]) * k);
                b = std::min(std::max(b, 0), kBins - 1);
                bb[b].grow(boxes_[p]);
                bc[b]++;
            }
            double ra[kBins];
            uint32_t rc[kBins];
            Aabb acc;
            uint32_t c = 0;
            for (int b = kBins - 1; b > 0; --b) {
                acc.grow(bb[b]);
                c += bc[b];
                ra[b] = acc.area();
                rc[b] = c;
            }
            Aabb lacc;
            uint32_t lc = 0;
            for (int b = 0; b < kBins - 1; ++b) {
                lacc.grow(bb[b]);
                lc += bc[b];
                if (lc == 0 || rc[b + 1] == 0) continue;
                const double cost = lacc.area() * lc + ra[b + 1] * rc[b + 1];
                if (cost < best) {
                    best = cost;
                    best_axis = a;
                    best_bin = b;
                }
            }
        }
        if (best_axis < 0) return begin + n / 2;
        const float ext = cb.hi[best_axis] - cb.lo[best_axis];
        const float k = kBins / ext;
        uint32_t *mid = std::partition(ids_ + begin, ids_ + end, [&](uint32_t p) {
            int b = (int)((cent_[3 * (size_t)p + best_axis] - cb.lo[best_axis]) * k);
            b = std::min(std::max(b, 0), kBins - 1);
            return b <= best_bin;
        });
        uint32_t m = static_cast<uint32_t>(mid - ids_);
        if (m == begin || m == end) m = begin + n / 2;
        return m;
    }

    const std::vector<Aabb> &boxes_;
    const std::vector<float> &cent_;
    unsigned threads_;
    std::atomic<uint32_t> next_{0};
    std::atomic<int> active_{0};
    uint32_t *ids_ = nullptr;
    BinNode *nodes_ = nullptr;
};

// power-of-two quantisation of one axis for a set of child intervals, rounded outward with one
// extra step of slack on each side so that FP32 slab tests cannot reject a contained triangle
static bool quantise_axis(const float *clo, const float *chi, int k, float ulo, float uhi, float &p, int &e,
                          uint8_t *qlo, uint8_t *qhi) {
    const double ext = (double)uhi - (double)ulo;
    int ee = ext > 0.0 ? (int)std::ceil(std::log2(ext / 252.0)) : -126;
    ee = std::max(ee, -126);
    for (; ee <= 127; ++ee) {
        const double step = std::ldexp(1.0, ee);
        float pp = (float)((double)ulo - step);
        if ((double)pp > (double)ulo - step) pp = std::nextafter(pp, -std::numeric_limits<float>::infinity());
        bool ok = std::isfinite(pp);
        for (int c = 0; c < k && ok; ++c) {
            const double lo = std::floor(((double)clo[c] - (double)pp) / step) - 1.0;
            const double hi = std::ceil(((double)chi[c] - (double)pp) / step) + 1.0;
            if (lo < 0.0 || hi > 255.0) {
                ok = false;
                break;
            }
            qlo[c] = (uint8_t)lo;
            qhi[c] = (uint8_t)hi;
        }
        if (ok) {
            p = pp;
            e = ee;
            return true;
        }
    }
    return false;
}

struct WideChild {
    uint32_t bin;  // binary node
    bool leaf;
};

// Optimal SAH collapse of the binary tree into W-wide nodes (the dynamic programme of Ylitie, Karras
// and Laine, "Efficient incoherent ray traversal on GPUs through compressed wide BVHs", HPG 2017).
// cover(x, k) is the least SAH cost of covering x's subtree with at most k child slots of one wide
// node; a slot holds either a leaf of at most L triangles (cost c_tri x area x triangles) or a wide
// node of its own (c_node x area + its best cover with W slots). Areas are relative to the root's.
class Collapse {
public:
    Collapse(const std::vector<BinNode> &bn, uint32_t W, uint32_t L, float c_node, float c_tri)
        : bn_(bn), W_(W), cover_((size_t)bn.size() * W), split_((size_t)bn.size() * W), leaf_(bn.size()),
          wide_split_(bn.size()) {
        const double root = std::max(bn[0].box.area(), 1e-300);
        const float inf = std::numeric_limits<float>::infinity();
        // children have larger indices than their parent: one backward sweep is bottom-up
        for (size_t x = bn.size(); x-- > 0;) {
            const BinNode &nd = bn[x];
            const float a = (float)(nd.box.area() / root);
            const float leaf_cost = nd.count <= L ? c_tri * a * (float)nd.count : inf;
            float *cv = &cover_[x * W];
            uint8_t *sp = &split_[x * W];
            if (nd.leaf) {
                for (uint32_t k = 0; k < W; ++k) cv[k] = leaf_cost, sp[k] = 0;
                leaf_[x] = 1;
                continue;
            }
            const float *cl = &cover_[(size_t)nd.left * W], *cr = &cover_[(size_t)(nd.left + 1) * W];
            float dist[9];
            uint8_t dsp[9];
            for (uint32_t k = 2; k <= W; ++k) {
                dist[k] = inf;
                dsp[k] = 1;
                for (uint32_t i = 1; i < k; ++i) {
                    const float c = cl[i - 1] + cr[k - i - 1];
                    if (c < dist[k]) dist[k] = c, dsp[k] = (uint8_t)i;
                }
            }
            const float inner_cost = c_node * a + dist[W];
            const float slot = std::min(leaf_cost, inner_cost);
            leaf_[x] = leaf_cost <= inner_cost;
            cv[0] = slot;
            sp[0] = 0;
            for (uint32_t k = 2; k <= W; ++k) {
                const bool spread = dist[k] < slot;
                cv[k - 1] = spread ? dist[k] : slot;
                sp[k - 1] = spread ? dsp[k] : 0;
            }
            wide_split_[x] = dsp[W];
        }
    }
    // the child slots of x as a wide node: its best cover with W slots, split between its two children
    void children(uint32_t x, std::vector<WideChild> &out) const {
        const uint32_t i = wide_split_[x];
        gather(bn_[x].left, i, out);
        gather(bn_[x].left + 1, W_ - i, out);
    }

private:
    void gather(uint32_t y, uint32_t k, std::vector<WideChild> &out) const {
        const uint32_t i = split_[(size_t)y * W_ + k - 1];
        if (i == 0) {
            out.push_back({y, leaf_[y] != 0});
            return;
        }
        gather(bn_[y].left, i, out);
        gather(bn_[y].left + 1, k - i, out);
    }
    const std::vector<BinNode> &bn_;
    uint32_t W_;
    std::vector<float> cover_;
    std::vector<uint8_t> split_, leaf_, wide_split_;
};

}  // namespace

std::string build_wide_bvh(const BuildInput &in, BuildOutput &out) {
    auto t0 = std::chrono::steady_clock::now();
    const uint32_t W = in.width;
    if (W != 4 && W != 8) return "bvh width must be 4 or 8";
    if (in.max_leaf_tris < 1 || in.max_leaf_tris > 3) return "max_leaf_tris must be 1..3";
    const uint32_t n = in.num_tris;
    out = BuildOutput();
    out.node_bytes = W == 8 ? sizeof(NodeW8) : sizeof(NodeW4);
    if (n == 0) return "scene has no triangles";

    std::vector<Aabb> boxes(n);
    std::vector<float> cent(3 * (size_t)n);
    uint64_t degenerate = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const float *v = in.tri_verts + 9 * (size_t)i;
        for (int k = 0; k < 9; ++k)
            if (!std::isfinite(v[k])) return "triangle " + std::to_string(i) + " has a non-finite vertex";
        boxes[i].grow_point(v);
        boxes[i].grow_point(v + 3);
        boxes[i].grow_point(v + 6);
        for (int a = 0; a < 3; ++a) cent[3 * (size_t)i + a] = 0.5f * (boxes[i].lo[a] + boxes[i].hi[a]);
        const double e1[3] = {(double)v[3] - v[0], (double)v[4] - v[1], (double)v[5] - v[2]};
        const double e2[3] = {(double)v[6] - v[0], (double)v[7] - v[1], (double)v[8] - v[2]};
        const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                     cz = e1[0] * e2[1] - e1[1] * e2[0];
        if (cx == 0.0 && cy == 0.0 && cz == 0.0) ++degenerate;
    }
    out.degenerate = degenerate;

    unsigned threads = in.num_threads ? in.num_threads : std::max(1u, std::thread::hardware_concurrency());
    std::vector<BinNode> bn;
    std::vector<uint32_t> ids;
    BinaryBuilder(boxes, cent, threads).build(bn, ids);

    const uint32_t L = in.max_leaf_tris;
    auto is_leaf = [&](uint32_t b) { return bn[b].count <= L; };
    auto collect = [&](uint32_t b, std::vector<uint32_t> &dst) {
        std::vector<uint32_t> st{b};
        while (!st.empty()) {
            uint32_t x = st.back();
            st.pop_back();
            if (bn[x].leaf) {
                dst.push_back(ids[bn[x].left]);
            } else {
                st.push_back(bn[x].left + 1);
                st.push_back(bn[x].left);
            }
        }
    };

    // SAH weights of the collapse: 1 per wide node visited (its eight child boxes), 3 per triangle
    // tested (fetch, watertight test and, in the path tracer, the divergence of lanes in leaves).
    // Measured on MI355X, 1080p frames of the 5.5M-triangle many-instance scene: greedy collapse
    // (open the largest child until W slots) 2.82 ms; triangle weight 0.3, 0.5 and 1 down to
    // 2.70-2.75 ms; 2 2.67-2.69; 3 2.66; 4 2.64-2.66 (with cache + NEE: 1 3.42-3.43, 3 3.32-3.37 ms).
    // Incoherent closest hit on the 2M-triangle heightfield stays at 1.88-1.89 ms for every weight.
    const Collapse collapse(bn, W, L, 1.0f, 3.0f);
    const double root_area = std::max(bn[0].box.area(), 1e-300);
    struct Pending {
        uint32_t bin;
        uint32_t depth;
    };
    std::vector<Pending> queue;
    queue.push_back({0, 1});
    out.nodes.reserve((size_t)out.node_bytes * (n / 4 + 16));
    out.tris.reserve(n);
    out.tri_input.reserve(n);
    out.num_nodes = 1;
    out.nodes.assign(out.node_bytes, 0);
    double sah = 0.0;

    for (size_t qi = 0; qi < queue.size(); ++qi) {
        const uint32_t b = queue[qi].bin;
        const uint32_t depth = queue[qi].depth;
        out.depth = std::max(out.depth, depth);
        sah += bn[b].box.area() / root_area;

        // children of this wide node
        std::vector<WideChild> ch;
        if (bn[b].leaf || (qi == 0 && is_leaf(b))) {
            ch.push_back({b, true});  // a scene of at most L triangles: one leaf under the root
        } else {
            collapse.children(b, ch);
        }
        const int k = (int)ch.size();

        // slot assignment: child c goes to the free slot whose direction best matches its offset
        int slot_of[8];
        for (int c = 0; c < 8; ++c) slot_of[c] = -1;
        {
            float pc[3];
            for (int a = 0; a < 3; ++a) pc[a] = 0.5f * (bn[b].box.lo[a] + bn[b].box.hi[a]);
            // node-local dominant axis for the 4-wide ordering
            int axis = 0;
            for (int a = 1; a < 3; ++a)
                if (bn[b].box.hi[a] - bn[b].box.lo[a] > bn[b].box.hi[axis] - bn[b].box.lo[axis]) axis = a;
            struct Cand {
                float score;
                int c, s;
            };
            std::vector<Cand> cands;
            for (int c = 0; c < k; ++c) {
                float cc[3];
                for (int a = 0; a < 3; ++a) cc[a] = 0.5f * (bn[ch[c].bin].box.lo[a] + bn[ch[c].bin].box.hi[a]) - pc[a];
                for (int s = 0; s < (int)W; ++s) {
                    float score;
                    if (W == 8) {
                        score = ((s & 1) ? cc[0] : -cc[0]) + ((s & 2) ? cc[1] : -cc[1]) + ((s & 4) ? cc[2] : -cc[2]);
                    } else {
                        score = cc[axis] * (float)(2 * s - 3);
                    }
                    cands.push_back({score, c, s});
                }
            }
            std::stable_sort(cands.begin(), cands.end(), [](const Cand &x, const Cand &y) { return x.score > y.score; });
            bool used[8] = {false};
            for (const auto &cd : cands) {
                if (slot_of[cd.c] >= 0 || used[cd.s]) continue;
                slot_of[cd.c] = cd.s;
                used[cd.s] = true;
            }
            if (W == 4) {
                // the traversal reads the ordering axis from the node
                NodeW4 *nd = reinterpret_cast<NodeW4 *>(out.nodes.data() + (size_t)qi * out.node_bytes);
                nd->pad[0] = (uint32_t)axis;
            }
        }
        int child_at[8];
        for (int s = 0; s < 8; ++s) child_at[s] = -1;
        for (int c = 0; c < k; ++c) child_at[slot_of[c]] = c;

        // quantisation over the union of child boxes
        float clo[3][8], chi[3][8];
        Aabb u;
        int kk = 0;
        int order[8];
        for (int s = 0; s < (int)W; ++s) {
            if (child_at[s] < 0) continue;
            order[kk] = s;
            const Aabb &cb = bn[ch[child_at[s]].bin].box;
            for (int a = 0; a < 3; ++a) {
                clo[a][kk] = cb.lo[a];
                chi[a][kk] = cb.hi[a];
            }
            u.grow(cb);
            ++kk;
        }
        float p[3];
        int e[3];
        uint8_t ql[3][8], qh[3][8];
        for (int a = 0; a < 3; ++a)
            if (!quantise_axis(clo[a], chi[a], kk, u.lo[a], u.hi[a], p[a], e[a], ql[a], qh[a]))
                return "quantisation failed (scene extent out of float range)";

        const uint32_t tri_base = (uint32_t)out.tris.size();
        const uint32_t child_base = out.num_nodes;
        uint8_t imask = 0, meta[8] = {0};
        uint8_t qlo[3][8] = {{0}}, qhi[3][8] = {{0}};
        uint32_t inner = 0;
        std::vector<uint32_t> leaf_tris;
        for (int j = 0; j < kk; ++j) {
            const int s = order[j];
            const WideChild &c = ch[child_at[s]];
            for (int a = 0; a < 3; ++a) {
                qlo[a][s] = ql[a][j];
                qhi[a][s] = qh[a][j];
            }
            if (c.leaf) {
                leaf_tris.clear();
                collect(c.bin, leaf_tris);
                const uint32_t off = (uint32_t)out.tris.size() - tri_base;
                meta[s] = (uint8_t)((leaf_tris.size() << 5) | off);
                sah += bn[c.bin].box.area() / root_area * (double)leaf_tris.size();
                for (uint32_t t : leaf_tris) {
                    TriRecord r;
                    const float *v = in.tri_verts + 9 * (size_t)t;
                    for (int a = 0; a < 3; ++a) {
                        r.v0[a] = v[a];
                        r.v1[a] = v[3 + a];
                        r.v2[a] = v[6 + a];
                    }
                    const uint32_t gid = in.global_ids ? in.global_ids[t] : t;
                    const uint32_t mat = in.materials ? in.materials[t] : 0u;
                    const uint32_t lid = in.local_ids ? in.local_ids[t] : t;
                    std::memcpy(&r.v0[3], &gid, 4);
                    std::memcpy(&r.v1[3], &mat, 4);
                    std::memcpy(&r.v2[3], &lid, 4);
                    out.tris.push_back(r);
                    out.tri_input.push_back(t);
                }
            } else {
                imask |= (uint8_t)(1u << s);
                queue.push_back({c.bin, depth + 1});
                ++inner;
            }
        }
        // node record for queue slot qi; children occupy [child_base, child_base + inner)
        out.num_nodes += inner;
        out.nodes.resize((size_t)out.num_nodes * out.node_bytes, 0);
        uint8_t *dst = out.nodes.data() + (size_t)qi * out.node_bytes;
        if (W == 8) {
            NodeW8 nd;
            std::memset(&nd, 0, sizeof(nd));
            nd.px = p[0];
            nd.py = p[1];
            nd.pz = p[2];
            nd.ex = (int8_t)e[0];
            nd.ey = (int8_t)e[1];
            nd.ez = (int8_t)e[2];
            nd.imask = imask;
            nd.child_base = child_base;
            nd.tri_base = tri_base;
            for (int s = 0; s < 8; ++s) {
                nd.meta[s] = meta[s];
                nd.qlo_x[s] = qlo[0][s];
                nd.qlo_y[s] = qlo[1][s];
                nd.qlo_z[s] = qlo[2][s];
                nd.qhi_x[s] = qhi[0][s];
                nd.qhi_y[s] = qhi[1][s];
                nd.qhi_z[s] = qhi[2][s];
            }
            std::memcpy(dst, &nd, sizeof(nd));
        } else {
            NodeW4 nd;
            std::memcpy(&nd, dst, sizeof(nd));  // keeps the ordering axis written above
            nd.px = p[0];
            nd.py = p[1];
            nd.pz = p[2];
            nd.ex = (int8_t)e[0];
            nd.ey = (int8_t)e[1];
            nd.ez = (int8_t)e[2];
            nd.imask = imask;
            nd.child_base = child_base;
            nd.tri_base = tri_base;
            for (int s = 0; s < 4; ++s) {
                nd.meta[s] = meta[s];
                nd.qlo_x[s] = qlo[0][s];
                nd.qlo_y[s] = qlo[1][s];
                nd.qlo_z[s] = qlo[2][s];
                nd.qhi_x[s] = qhi[0][s];
                nd.qhi_y[s] = qhi[1][s];
                nd.qhi_z[s] = qhi[2][s];
            }
            std::memcpy(dst, &nd, sizeof(nd));
        }
        if (out.num_nodes >= (1u << 24)) return "scene needs more than 2^24 nodes";
    }
    out.sah_cost = sah;
    out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return std::string();
}

std::string validate_wide_bvh(const BuildInput &in, const BuildOutput &out) {
    const uint32_t W = in.width;
    std::vector<uint8_t> seen(in.num_tris, 0);
    if (out.tris.size() != in.num_tris) return "triangle count mismatch";
    struct Item {
        uint32_t node, depth;
        Aabb box;  // intersection of the dequantised ancestor slot boxes (unbounded at the root)
    };
    std::vector<Item> st;
    Aabb all;
    for (int a = 0; a < 3; ++a) {
        all.lo[a] = -std::numeric_limits<float>::infinity();
        all.hi[a] = std::numeric_limits<float>::infinity();
    }
    st.push_back({0, 1, all});
    uint32_t maxd = 0;
    uint64_t visited = 0;
    while (!st.empty()) {
        Item it = st.back();
        st.pop_back();
        if (it.node >= out.num_nodes) return "child index out of range";
        ++visited;
        maxd = std::max(maxd, it.depth);
        const uint8_t *raw = out.nodes.data() + (size_t)it.node * out.node_bytes;
        float p[3];
        int8_t e[3];
        uint8_t imask;
        uint32_t child_base, tri_base;
        const uint8_t *meta, *q[6];
        if (W == 8) {
            const NodeW8 *nd = reinterpret_cast<const NodeW8 *>(raw);
            p[0] = nd->px, p[1] = nd->py, p[2] = nd->pz;
            e[0] = nd->ex, e[1] = nd->ey, e[2] = nd->ez;
            imask = nd->imask, child_base = nd->child_base, tri_base = nd->tri_base, meta = nd->meta;
            q[0] = nd->qlo_x, q[1] = nd->qlo_y, q[2] = nd->qlo_z, q[3] = nd->qhi_x, q[4] = nd->qhi_y, q[5] = nd->qhi_z;
        } else {
            const NodeW4 *nd = reinterpret_cast<const NodeW4 *>(raw);
            p[0] = nd->px, p[1] = nd->py, p[2] = nd->pz;
            e[0] = nd->ex, e[1] = nd->ey, e[2] = nd->ez;
            imask = nd->imask, child_base = nd->child_base, tri_base = nd->tri_base, meta = nd->meta;
            q[0] = nd->qlo_x, q[1] = nd->qlo_y, q[2] = nd->qlo_z, q[3] = nd->qhi_x, q[4] = nd->qhi_y, q[5] = nd->qhi_z;
        }
        uint32_t inner_rank = 0;
        for (uint32_t s = 0; s < W; ++s) {
            Aabb cb;
            for (int a = 0; a < 3; ++a) {
                const double step = std::ldexp(1.0, e[a]);
                cb.lo[a] = (float)((double)p[a] + q[a][s] * step);
                cb.hi[a] = (float)((double)p[a] + q[3 + a][s] * step);
            }
            if (imask & (1u << s)) {
                Aabb nb = cb;
                for (int a = 0; a < 3; ++a) {
                    nb.lo[a] = std::max(nb.lo[a], it.box.lo[a]);
                    nb.hi[a] = std::min(nb.hi[a], it.box.hi[a]);
                }
                st.push_back({child_base + inner_rank, it.depth + 1, nb});
                ++inner_rank;
                continue;
            }
            if (meta[s] == 0) continue;
            const uint32_t cnt = meta[s] >> 5, off = meta[s] & 31u;
            for (uint32_t k = 0; k < cnt; ++k) {
                const uint32_t ti = tri_base + off + k;
                if (ti >= out.tris.size()) return "triangle index out of range";
                const uint32_t src = out.tri_input[ti];
                if (seen[src]) return "triangle referenced twice";
                seen[src] = 1;
                const TriRecord &r = out.tris[ti];
                const float *vs[3] = {r.v0, r.v1, r.v2};
                for (int vtx = 0; vtx < 3; ++vtx)
                    for (int a = 0; a < 3; ++a) {
                        const double lo = (double)p[a] + q[a][s] * std::ldexp(1.0, e[a]);
                        const double hi = (double)p[a] + q[3 + a][s] * std::ldexp(1.0, e[a]);
                        if (!((double)vs[vtx][a] >= lo && (double)vs[vtx][a] <= hi))
                            return "triangle outside its quantised leaf box";
                        if (!(vs[vtx][a] >= it.box.lo[a] && vs[vtx][a] <= it.box.hi[a]))
                            return "triangle outside an ancestor box";
                    }
            }
        }
    }
    for (uint32_t i = 0; i < in.num_tris; ++i)
        if (!seen[i]) return "triangle not referenced";
    if (visited != out.num_nodes) return "unreachable nodes";
    if (maxd != out.depth) return "depth mismatch";
    return std::string();
}

}  // namespace spk
