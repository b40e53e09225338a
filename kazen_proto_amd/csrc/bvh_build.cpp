// Host BVH builder of the HIP intersector: binned-SAH BVH2, collapsed into
// 8-wide nodes with 16-bit quantised child boxes (layout in kp_layout.h).
//
// Replaces the acceleration-structure build behind Accelerator::build /
// add_trianglemesh (reference src/core/accel.h:26-32; Embree build at
// src/core/accel.cpp:185-346, in-tree median-split BVH at accel.cpp:99-132).
#include "kp_bvh.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace kp {
namespace {

struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    void grow_pt(const float* p) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    double half_area() const {
        double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
        if (!(dx >= 0 && dy >= 0 && dz >= 0)) return 0.0;
        return dx * dy + dy * dz + dz * dx;
    }
};

struct N2 {
    Box box;
    uint32_t a = 0, b = 0;    // inner: children; leaf: first, count
    bool leaf = false;
};

constexpr uint32_t kMaxBins = 64;

struct Builder {
    const float* tris;
    uint32_t n;
    uint32_t leaf_max;
    uint32_t bins;
    std::vector<Box> tb;
    std::vector<float> cen;
    std::vector<uint32_t> idx;
    std::vector<N2> nodes;
    std::atomic<uint32_t> next_node{1};
    std::atomic<int> threads_left{0};

    uint32_t alloc2() { return next_node.fetch_add(2); }

    void build(uint32_t node, uint32_t begin, uint32_t end, int depth) {
        for (;;) {
            Box nb, cb;
            for (uint32_t i = begin; i < end; ++i) {
                nb.grow(tb[idx[i]]);
                cb.grow_pt(&cen[3 * (size_t)idx[i]]);
            }
            N2& nd = nodes[node];
            nd.box = nb;
            const uint32_t cnt = end - begin;
            if (cnt <= 1) { nd.leaf = true; nd.a = begin; nd.b = cnt; return; }
            // Binned SAH over all three axes.
            double best = INFINITY;
            int best_axis = -1, best_split = -1;
            Box bb[kMaxBins];
            uint32_t bc[kMaxBins];
            double ra[kMaxBins];
            uint32_t rn[kMaxBins];
            for (int axis = 0; axis < 3; ++axis) {
                const float ext = cb.hi[axis] - cb.lo[axis];
                if (!(ext > 0.0f)) continue;
                const float scale = (float)bins / ext;
                for (uint32_t i = 0; i < bins; ++i) { bb[i] = Box{}; bc[i] = 0u; }
                for (uint32_t i = begin; i < end; ++i) {
                    uint32_t t = idx[i];
                    int bi = (int)((cen[3 * (size_t)t + axis] - cb.lo[axis]) * scale);
                    bi = std::min(std::max(bi, 0), (int)bins - 1);
                    bb[bi].grow(tb[t]);
                    bc[bi]++;
                }
                Box acc;
                uint32_t an = 0;
                for (int i = (int)bins - 1; i > 0; --i) {
                    acc.grow(bb[i]);
                    an += bc[i];
                    ra[i] = acc.half_area();
                    rn[i] = an;
                }
                Box lacc;
                uint32_t ln = 0;
                for (int i = 1; i < (int)bins; ++i) {
                    lacc.grow(bb[i - 1]);
                    ln += bc[i - 1];
                    if (ln == 0 || rn[i] == 0) continue;
                    double c = lacc.half_area() * ln + ra[i] * rn[i];
                    if (c < best) { best = c; best_axis = axis; best_split = i; }
                }
            }
            const double pa = nb.half_area();
            const double split_cost = 1.0 + (pa > 0.0 ? best / pa : (double)cnt);
            uint32_t mid;
            if (best_axis < 0 || (cnt <= leaf_max && (double)cnt <= split_cost)) {
                if (cnt <= leaf_max) { nd.leaf = true; nd.a = begin; nd.b = cnt; return; }
                // No useful split (coincident centroids): split by index.
                mid = begin + cnt / 2;
            } else {
                const int axis = best_axis;
                const float ext = cb.hi[axis] - cb.lo[axis];
                const float scale = (float)bins / ext;
                const float lo = cb.lo[axis];
                const int split = best_split;
                const int nb_ = (int)bins;
                auto pm = std::partition(idx.begin() + begin, idx.begin() + end, [&](uint32_t t) {
                    int bi = (int)((cen[3 * (size_t)t + axis] - lo) * scale);
                    bi = std::min(std::max(bi, 0), nb_ - 1);
                    return bi < split;
                });
                mid = (uint32_t)(pm - idx.begin());
                if (mid == begin || mid == end) mid = begin + cnt / 2;
            }
            const uint32_t l = alloc2();
            nd.leaf = false;
            nd.a = l;
            nd.b = l + 1;
            // Large subtrees are built on their own thread.
            if (end - mid > 65536 && threads_left.fetch_sub(1) > 0) {
                std::thread th([this, l, mid, end, depth]() { build(l + 1, mid, end, depth + 1); });
                build(l, begin, mid, depth + 1);
                th.join();
                return;
            }
            // Recurse into the smaller side and loop on the larger one, which
            // bounds the recursion depth by log2(n).
            if (mid - begin < end - mid) {
                build(l, begin, mid, depth + 1);
                node = l + 1;
                begin = mid;
            } else {
                build(l + 1, mid, end, depth + 1);
                node = l;
                end = mid;
            }
            ++depth;
        }
    }
};

}  // namespace

constexpr uint32_t kNoKid = 0xFFFFFFFFu;   // empty slot of an 8-wide node
constexpr uint32_t kQMax = 2047;            // plane grid: integers 0..2047, exact in binary16

// binary16 bit pattern of an integer 0..2047 (exact).
static uint16_t f16_int(uint32_t q) {
    if (q == 0) return 0;
    int e = 0;
    while ((q >> (e + 1)) != 0) ++e;
    const uint32_t mant = (q << (10 - e)) & 0x3FFu;
    return (uint16_t)(((uint32_t)(e + 15) << 10) | mant);
}

bool build_bvh8(const float* tris, uint32_t n_tris, const BuildOptions& opt, Bvh8Host& out,
                std::string& err) {
    out = Bvh8Host{};
    if (n_tris == 0) { err = "empty scene: no triangles"; return false; }
    if (n_tris >= (1u << 28)) { err = "too many triangles (limit 2^28)"; return false; }
    const uint32_t leaf_max = std::max(1u, std::min(opt.leaf_size ? opt.leaf_size : 4u, 4u));
    if (opt.max_children < 2 || opt.max_children > 8) { err = "node width must be 2..8"; return false; }
    const size_t max_children = opt.max_children;
    Builder b;
    b.tris = tris;
    b.n = n_tris;
    b.leaf_max = leaf_max;
    b.bins = std::max(4u, std::min(opt.sah_bins ? opt.sah_bins : 32u, kMaxBins));
    b.tb.resize(n_tris);
    b.cen.resize(3 * (size_t)n_tris);
    b.idx.resize(n_tris);
    float max_abs = 1.0f;
    for (uint32_t i = 0; i < n_tris; ++i) {
        const float* t = tris + 9 * (size_t)i;
        for (int v = 0; v < 3; ++v) {
            for (int k = 0; k < 3; ++k) {
                if (!std::isfinite(t[3 * v + k])) { err = "non-finite vertex coordinate"; return false; }
                max_abs = std::max(max_abs, std::fabs(t[3 * v + k]));
            }
            b.tb[i].grow_pt(t + 3 * v);
        }
        for (int k = 0; k < 3; ++k) b.cen[3 * (size_t)i + k] = 0.5f * (b.tb[i].lo[k] + b.tb[i].hi[k]);
        b.idx[i] = i;
    }
    // Widen every box by 2^-16 of the largest coordinate so the box test is
    // conservative with respect to the watertight triangle test.
    const float pad = std::ldexp(max_abs, -16);
    b.nodes.resize(2 * (size_t)n_tris);
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    b.threads_left = (int)std::min(hw, 64u) - 1;
    b.build(0, 0, n_tris, 0);
    const uint32_t n2 = b.next_node.load();
    b.nodes.resize(n2);
    for (auto& nd : b.nodes)
        for (int k = 0; k < 3; ++k) { nd.box.lo[k] -= pad; nd.box.hi[k] += pad; }

    // Collapse plan (cost-optimal, Ylitie, Karras, Laine 2017): for every
    // BVH2 node n and k = 1..W, cost(n, k) is the least SAH cost of n's
    // subtree as at most k roots of a wide node; a root is a leaf (at most
    // leaf_max triangles, a contiguous range of the build order) or a wide
    // node whose children are its subtree split into at most W roots.
    // Children have larger indices than their parent, so one pass from the
    // last node to the first sees every child before its parent.  Costs: a
    // node visit 1, a triangle test 0.5 (measured best on incoherent rays,
    // where a wave-wide triangle test and a node visit cost about as many
    // instructions per lane; 0.4 to 1.0 are within 1 %).  Against opening the
    // largest child first, the plan flattens small subtrees into their
    // parents: 34 % fewer nodes, 4 % fewer node visits per incoherent ray.
    const size_t W = max_children;
    const double c_node = 1.0, c_tri = 0.5;
    std::vector<uint32_t> first2, cnt2, split2;
    std::vector<float> cost2;
    std::vector<uint8_t> how2;   // k = 1: 0 leaf, 1 wide node; k > 1: 0 as k - 1, j: left j roots, right k - j
    {
        first2.resize(n2);
        cnt2.resize(n2);
        split2.resize(n2);
        cost2.resize((size_t)n2 * W);
        how2.resize((size_t)n2 * W);
        for (uint32_t n = n2; n-- > 0;) {
            const N2& nd = b.nodes[n];
            float* C = &cost2[(size_t)n * W];
            uint8_t* H = &how2[(size_t)n * W];
            const double A = nd.box.half_area();
            if (nd.leaf) {
                first2[n] = nd.a;
                cnt2[n] = nd.b;
                for (size_t k = 0; k < W; ++k) { C[k] = (float)(A * c_tri * nd.b); H[k] = 0; }
                continue;
            }
            const uint32_t l = nd.a, r = nd.b;
            first2[n] = first2[l];
            cnt2[n] = cnt2[l] + cnt2[r];
            const float* CL = &cost2[(size_t)l * W];
            const float* CR = &cost2[(size_t)r * W];
            double dist[9];
            uint32_t bj[9];
            for (size_t k = 2; k <= W; ++k) {
                dist[k] = INFINITY;
                bj[k] = 1;
                for (size_t j = 1; j < k; ++j) {
                    const double c = (double)CL[j - 1] + (double)CR[k - j - 1];
                    if (c < dist[k]) { dist[k] = c; bj[k] = (uint32_t)j; }
                }
            }
            split2[n] = bj[W];
            const double c_wide = A * c_node + dist[W];
            const double c_leaf = cnt2[n] <= leaf_max ? A * c_tri * cnt2[n] : INFINITY;
            C[0] = (float)std::min(c_wide, c_leaf);
            H[0] = c_leaf <= c_wide ? 0 : 1;
            for (size_t k = 2; k <= W; ++k) {
                if (dist[k] < (double)C[k - 2]) { C[k - 1] = (float)dist[k]; H[k - 1] = (uint8_t)bj[k]; }
                else { C[k - 1] = C[k - 2]; H[k - 1] = 0; }
            }
        }
    }
    // the roots of n's subtree as at most k children (plan above)
    std::vector<uint32_t> gather_stack;
    auto gather = [&](uint32_t n, uint32_t k, std::vector<uint32_t>& out) {
        gather_stack.clear();
        gather_stack.push_back(n);
        gather_stack.push_back(k);
        while (!gather_stack.empty()) {
            uint32_t kk = gather_stack.back();
            gather_stack.pop_back();
            const uint32_t nn = gather_stack.back();
            gather_stack.pop_back();
            while (kk > 1 && how2[(size_t)nn * W + kk - 1] == 0) --kk;
            if (kk == 1) { out.push_back(nn); continue; }
            const uint32_t j = how2[(size_t)nn * W + kk - 1];
            gather_stack.push_back(b.nodes[nn].b);
            gather_stack.push_back(kk - j);
            gather_stack.push_back(b.nodes[nn].a);
            gather_stack.push_back(j);
        }
    };
    auto as_leaf = [&](uint32_t n) { return how2[(size_t)n * W] == 0; };

    // Collapse to 8-wide nodes, breadth first.
    std::vector<KpNode8> nodes8;
    std::vector<uint32_t> tri_order;   // reordered triangle -> original index
    nodes8.reserve(n2 / 4 + 1);
    tri_order.reserve(n_tris);
    struct Pending { uint32_t n2, n8; };
    std::vector<Pending> queue;
    nodes8.emplace_back();
    queue.push_back({0, 0});
    uint32_t max_depth = 1;
    std::vector<uint32_t> depth8(1, 1);
    for (size_t qi = 0; qi < queue.size(); ++qi) {
        const Pending pend = queue[qi];
        std::vector<uint32_t> kids;
        const N2& root = b.nodes[pend.n2];
        if (root.leaf) {   // a scene of one leaf: a root node with one leaf child
            kids.push_back(pend.n2);
        } else {
            gather(root.a, split2[pend.n2], kids);
            gather(root.b, (uint32_t)W - split2[pend.n2], kids);
        }
        // Octant slot order: slot s (bit k set = high side of axis k) takes the
        // child whose centre lies most in that direction from the node centre
        // (greedy over all child/slot pairs), so a ray visiting slots in the
        // order s ^ octant(ray direction) meets children roughly near to far.
        {
            Box pb;
            for (uint32_t k : kids) pb.grow(b.nodes[k].box);
            struct Cand { double cost; int kid, slot; };
            std::vector<Cand> cand;
            for (int i = 0; i < (int)kids.size(); ++i) {
                const Box& cb = b.nodes[kids[i]].box;
                for (int s = 0; s < 8; ++s) {
                    double cost = 0.0;
                    for (int k = 0; k < 3; ++k) {
                        const double off = 0.5 * ((double)cb.lo[k] + cb.hi[k]) - 0.5 * ((double)pb.lo[k] + pb.hi[k]);
                        cost -= ((s >> k) & 1 ? 1.0 : -1.0) * off;
                    }
                    cand.push_back({cost, i, s});
                }
            }
            std::stable_sort(cand.begin(), cand.end(), [](const Cand& x, const Cand& y) { return x.cost < y.cost; });
            int slot_kid[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
            bool kid_done[8] = {false, false, false, false, false, false, false, false};
            for (const Cand& cd : cand) {
                if (kid_done[cd.kid] || slot_kid[cd.slot] >= 0) continue;
                slot_kid[cd.slot] = cd.kid;
                kid_done[cd.kid] = true;
            }
            std::vector<uint32_t> by_slot(8, kNoKid);
            for (int s = 0; s < 8; ++s)
                if (slot_kid[s] >= 0) by_slot[s] = kids[slot_kid[s]];
            kids = by_slot;
        }
        KpNode8 nd;
        std::memset(&nd, 0, sizeof(nd));
        Box nb;
        for (uint32_t k : kids)
            if (k != kNoKid) nb.grow(b.nodes[k].box);
        nd.px = nb.lo[0];
        nd.py = nb.lo[1];
        nd.pz = nb.lo[2];
        int e[3];
        for (int k = 0; k < 3; ++k) {
            double ext = (double)nb.hi[k] - (double)nb.lo[k];
            int ek = -126;
            if (ext > 0.0) ek = (int)std::ceil(std::log2(ext / (double)kQMax));
            ek = std::max(-126, std::min(127, ek));
            while (std::ldexp((double)kQMax, ek) < ext && ek < 127) ++ek;
            e[k] = ek;
        }
        nd.ex = (int8_t)e[0];
        nd.ey = (int8_t)e[1];
        nd.ez = (int8_t)e[2];
        const double p[3] = {nd.px, nd.py, nd.pz};
        uint32_t inner_rank = 0;
        const uint32_t child_base = (uint32_t)nodes8.size();
        const uint32_t prim_base = (uint32_t)tri_order.size();
        nd.child_base = child_base;
        nd.prim_base = prim_base;
        for (int s = 0; s < 8; ++s) {
            if (kids[s] == kNoKid) {
                for (int k = 0; k < 3; ++k) { nd.q[k][s] = f16_int(kQMax); nd.q[3 + k][s] = f16_int(0); }
                nd.meta[s] = 0;
                continue;
            }
            const N2& c = b.nodes[kids[s]];
            for (int k = 0; k < 3; ++k) {
                const double sc = std::ldexp(1.0, e[k]);
                double ql = std::floor(((double)c.box.lo[k] - p[k]) / sc);
                double qh = std::ceil(((double)c.box.hi[k] - p[k]) / sc);
                ql = std::max(0.0, std::min((double)kQMax, ql));
                qh = std::max(0.0, std::min((double)kQMax, qh));
                nd.q[k][s] = f16_int((uint32_t)ql);
                nd.q[3 + k][s] = f16_int((uint32_t)qh);
            }
            if (as_leaf(kids[s])) {
                const uint32_t first = first2[kids[s]], count = cnt2[kids[s]];
                const uint32_t off = (uint32_t)tri_order.size() - prim_base;
                for (uint32_t i = 0; i < count; ++i) tri_order.push_back(b.idx[first + i]);
                nd.meta[s] = (uint8_t)((count << 5) | off);
            } else {
                nd.imask |= (uint8_t)(1u << s);
                nd.meta[s] = 1;
                const uint32_t child = child_base + inner_rank++;
                queue.push_back({kids[s], child});
                (void)child;
            }
        }
        // reserve the inner children (contiguous, slot order)
        for (uint32_t i = 0; i < inner_rank; ++i) {
            nodes8.emplace_back();
            depth8.push_back(depth8[pend.n8] + 1);
            max_depth = std::max(max_depth, depth8.back());
        }
        nodes8[pend.n8] = nd;
    }
    if (nodes8.size() >= (1u << 24)) {
        err = "BVH8 exceeds 2^24 nodes (traversal stack entries hold 24-bit node indices)";
        return false;
    }
    out.nodes = std::move(nodes8);
    out.tri_order = std::move(tri_order);
    out.tris.resize(12 * (size_t)n_tris);
    for (uint32_t i = 0; i < n_tris; ++i) {
        const uint32_t src = out.tri_order[i];
        const float* t = tris + 9 * (size_t)src;
        float* d = &out.tris[12 * (size_t)i];
        d[0] = t[0]; d[1] = t[1]; d[2] = t[2];
        std::memcpy(&d[3], &src, 4);
        d[4] = t[3]; d[5] = t[4]; d[6] = t[5]; d[7] = 0.0f;
        d[8] = t[6]; d[9] = t[7]; d[10] = t[8]; d[11] = 0.0f;
    }
    out.max_depth = max_depth;
    out.n_bvh2_nodes = n2;
    out.pad = pad;
    return true;
}

}  // namespace kp
