#include "bvh.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <deque>
#include <future>
#include <stdexcept>

namespace vs {

namespace {

struct Box {
    f3 lo{FLT_MAX, FLT_MAX, FLT_MAX};
    f3 hi{-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(f3 p) {
        lo = vmin(lo, p);
        hi = vmax(hi, p);
    }
    void grow(const Box &b) {
        lo = vmin(lo, b.lo);
        hi = vmax(hi, b.hi);
    }
    bool valid() const { return lo.x <= hi.x && lo.y <= hi.y && lo.z <= hi.z; }
    float area() const {
        if (!valid()) return 0.f;
        f3 d = hi - lo;
        return 2.f * (d.x * d.y + d.y * d.z + d.z * d.x);
    }
    f3 center() const { return (lo + hi) * 0.5f; }
};

struct Node2 {
    Box box;
    uint32_t left{0}, right{0};
    uint32_t first{0}, count{0};// count > 0: leaf over idx[first, first+count)
};

struct Builder2 {
    const std::vector<Box> &tri_box;
    const std::vector<f3> &centroid;
    std::vector<uint32_t> &idx;
    std::vector<Node2> nodes;

    static constexpr int kBins = 32;

    uint32_t build(uint32_t first, uint32_t count) {
        uint32_t id = uint32_t(nodes.size());
        nodes.emplace_back();
        Box box, cbox;
        for (uint32_t i = first; i < first + count; ++i) {
            box.grow(tri_box[idx[i]]);
            cbox.grow(centroid[idx[i]]);
        }
        nodes[id].box = box;
        if (count == 1) {
            nodes[id].first = first;
            nodes[id].count = 1;
            return id;
        }
        // binned SAH over centroid bounds
        int best_axis = -1;
        int best_split = -1;
        float best_cost = FLT_MAX;
        f3 ext = cbox.hi - cbox.lo;
        for (int axis = 0; axis < 3; ++axis) {
            float e = comp(ext, axis);
            if (!(e > 0.f)) continue;
            Box bins[kBins];
            uint32_t cnt[kBins] = {};
            float lo = comp(cbox.lo, axis);
            float k = float(kBins) * (1.f - 1e-6f) / e;
            for (uint32_t i = first; i < first + count; ++i) {
                int b = std::min(kBins - 1, int((comp(centroid[idx[i]], axis) - lo) * k));
                cnt[b]++;
                bins[b].grow(tri_box[idx[i]]);
            }
            float right_area[kBins];
            uint32_t right_cnt[kBins];
            Box acc;
            uint32_t c = 0;
            for (int b = kBins - 1; b > 0; --b) {
                acc.grow(bins[b]);
                c += cnt[b];
                right_area[b] = acc.area();
                right_cnt[b] = c;
            }
            acc = Box();
            c = 0;
            for (int b = 0; b < kBins - 1; ++b) {
                acc.grow(bins[b]);
                c += cnt[b];
                if (c == 0 || right_cnt[b + 1] == 0) continue;
                float cost = acc.area() * float(c) + right_area[b + 1] * float(right_cnt[b + 1]);
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = axis;
                    best_split = b;
                }
            }
        }
        uint32_t mid;
        if (best_axis < 0) {
            // all centroids coincide: split the range in half
            mid = first + count / 2;
        } else {
            float lo = comp(cbox.lo, best_axis);
            float k = float(kBins) * (1.f - 1e-6f) / comp(ext, best_axis);
            auto it = std::partition(idx.begin() + first, idx.begin() + first + count, [&](uint32_t t) {
                int b = std::min(kBins - 1, int((comp(centroid[t], best_axis) - lo) * k));
                return b <= best_split;
            });
            mid = uint32_t(it - idx.begin());
            if (mid == first || mid == first + count) mid = first + count / 2;
        }
        uint32_t l = build(first, mid - first);
        uint32_t r = build(mid, first + count - mid);
        nodes[id].left = l;
        nodes[id].right = r;
        nodes[id].count = 0;
        return id;
    }
};

uint32_t subtree_count(const std::vector<Node2> &n, uint32_t i, std::vector<uint32_t> &memo) {
    if (memo[i] != 0) return memo[i];
    uint32_t c = n[i].count > 0 ? n[i].count : subtree_count(n, n[i].left, memo) + subtree_count(n, n[i].right, memo);
    memo[i] = c;
    return c;
}

void gather_tris(const std::vector<Node2> &n, uint32_t i, const std::vector<uint32_t> &idx, std::vector<uint32_t> &out) {
    if (n[i].count > 0) {
        for (uint32_t k = 0; k < n[i].count; ++k) out.push_back(idx[n[i].first + k]);
        return;
    }
    gather_tris(n, n[i].left, idx, out);
    gather_tris(n, n[i].right, idx, out);
}

uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

void set_byte(uint32_t *w, int word0, int slot, uint32_t v) {
    uint32_t &word = w[word0 + (slot >> 2)];
    int sh = (slot & 3) * 8;
    word = (word & ~(0xFFu << sh)) | ((v & 0xFFu) << sh);
}

}// namespace

Bvh8 build_bvh8(const std::vector<f3> &P, const std::vector<u3> &tris) {
    Bvh8 out;
    const uint32_t n = uint32_t(tris.size());
    if (n == 0) throw std::runtime_error("scene has no triangles");
    std::vector<Box> tri_box(n);
    std::vector<f3> centroid(n);
    for (uint32_t i = 0; i < n; ++i) {
        Box b;
        b.grow(P[tris[i].x]);
        b.grow(P[tris[i].y]);
        b.grow(P[tris[i].z]);
        tri_box[i] = b;
        centroid[i] = b.center();
    }
    std::vector<uint32_t> idx(n);
    for (uint32_t i = 0; i < n; ++i) idx[i] = i;
    Builder2 b2{tri_box, centroid, idx, {}};
    b2.nodes.reserve(size_t(2) * n);
    b2.build(0, n);
    const std::vector<Node2> &N2 = b2.nodes;
    out.bvh2_nodes = uint32_t(N2.size());
    std::vector<uint32_t> memo(N2.size(), 0);
    subtree_count(N2, 0, memo);

    struct Pending {
        uint32_t node8;
        uint32_t node2;
        uint32_t depth;
    };
    std::deque<Pending> queue;
    out.nodes.push_back(Bvh8Node{});
    queue.push_back({0, 0, 1});
    const float root_area = std::max(N2[0].box.area(), 1e-30f);
    double sah = 0.0;

    while (!queue.empty()) {
        Pending pend = queue.front();
        queue.pop_front();
        out.max_depth = std::max(out.max_depth, pend.depth);
        // greedy opening of the largest-area openable child
        std::vector<uint32_t> ch{pend.node2};
        while (ch.size() < 8) {
            int best = -1;
            float best_area = -1.f;
            for (size_t i = 0; i < ch.size(); ++i) {
                const Node2 &c = N2[ch[i]];
                if (c.count > 0) continue;
                float a = c.box.area();
                if (a > best_area) {
                    best_area = a;
                    best = int(i);
                }
            }
            if (best < 0) break;
            uint32_t c = ch[size_t(best)];
            ch[size_t(best)] = N2[c].left;
            ch.push_back(N2[c].right);
        }
        const Box &nb = N2[pend.node2].box;
        sah += double(nb.area() / root_area);

        // slot assignment toward octant directions
        f3 nc = nb.center();
        int slot_of[8];
        {
            bool child_used[8] = {}, slot_used[8] = {};
            struct Cand {
                float score;
                int c, s;
            };
            std::vector<Cand> cand;
            for (size_t c = 0; c < ch.size(); ++c) {
                f3 d = N2[ch[c]].box.center() - nc;
                for (int s = 0; s < 8; ++s) {
                    f3 dir = mk3((s & 1) ? 1.f : -1.f, (s & 2) ? 1.f : -1.f, (s & 4) ? 1.f : -1.f);
                    cand.push_back({dot(d, dir), int(c), s});
                }
            }
            std::stable_sort(cand.begin(), cand.end(), [](const Cand &a, const Cand &b) { return a.score > b.score; });
            for (const Cand &cd : cand) {
                if (child_used[cd.c] || slot_used[cd.s]) continue;
                child_used[cd.c] = slot_used[cd.s] = true;
                slot_of[cd.c] = cd.s;
            }
        }

        Bvh8Node node{};
        uint32_t *w = node.w;
        // quantization frame: origin padded a few ulps below the node box,
        // 253 steps cover the padded extent so every child box can be widened
        // by one step on each side (conservative under float decode)
        int ebias[3];
        double origin[3], step[3];
        for (int a = 0; a < 3; ++a) {
            double lo = comp(nb.lo, a), hi = comp(nb.hi, a);
            double mag = std::max(std::fabs(lo), std::fabs(hi));
            double margin = mag * std::ldexp(1.0, -18) + 1e-30;
            float pf = float(lo - margin);
            if (double(pf) > lo - margin) pf = std::nextafter(pf, -FLT_MAX);
            double span = (hi + margin) - double(pf);
            int e = int(std::ceil(std::log2(span / 253.0)));
            e = std::max(e, -100);
            while (std::ldexp(253.0, e) < span) ++e;
            if (e > 120) throw std::runtime_error("bvh: scene extent too large to quantize");
            ebias[a] = e + 127;
            origin[a] = pf;
            step[a] = std::ldexp(1.0, e);
            w[a] = fbits(pf);
        }
        w[3] = uint32_t(ebias[0]) | (uint32_t(ebias[1]) << 8) | (uint32_t(ebias[2]) << 16);
        // empty slots: inverted box (qlo=255 > qhi=0) never hits
        for (int s = 0; s < 8; ++s) {
            set_byte(w, 8, s, 255);
            set_byte(w, 10, s, 255);
            set_byte(w, 12, s, 255);
            set_byte(w, 14, s, 0);
            set_byte(w, 16, s, 0);
            set_byte(w, 18, s, 0);
        }
        uint32_t imask = 0;
        // internal children first, in slot order, get consecutive node ids
        uint32_t child_base = uint32_t(out.nodes.size());
        uint32_t tri_base = uint32_t(out.tris.size());
        w[4] = child_base;
        w[5] = tri_base;
        std::vector<int> by_slot(8, -1);
        for (size_t c = 0; c < ch.size(); ++c) by_slot[size_t(slot_of[c])] = int(c);
        uint32_t tri_off = 0;
        for (int s = 0; s < 8; ++s) {
            int c = by_slot[size_t(s)];
            if (c < 0) continue;
            const Node2 &cn = N2[ch[size_t(c)]];
            for (int a = 0; a < 3; ++a) {
                double lo = comp(cn.box.lo, a), hi = comp(cn.box.hi, a);
                double ql = std::floor((lo - origin[a]) / step[a]) - 1.0;
                double qh = std::ceil((hi - origin[a]) / step[a]) + 1.0;
                ql = std::max(0.0, ql);
                qh = std::min(255.0, qh);
                set_byte(w, 8 + 2 * a, s, uint32_t(ql));
                set_byte(w, 14 + 2 * a, s, uint32_t(qh));
            }
            uint32_t cnt = memo[ch[size_t(c)]];
            if (cnt <= 3) {
                std::vector<uint32_t> lt;
                gather_tris(N2, ch[size_t(c)], idx, lt);
                set_byte(w, 6, s, (cnt << 5) | tri_off);
                for (uint32_t t : lt) {
                    LeafTri rec{};
                    f3 v0 = P[tris[t].x], v1 = P[tris[t].y], v2 = P[tris[t].z];
                    rec.v0[0] = v0.x; rec.v0[1] = v0.y; rec.v0[2] = v0.z;
                    rec.v1[0] = v1.x; rec.v1[1] = v1.y; rec.v1[2] = v1.z;
                    rec.v2[0] = v2.x; rec.v2[1] = v2.y; rec.v2[2] = v2.z;
                    rec.id = t;
                    out.tris.push_back(rec);
                }
                tri_off += cnt;
                sah += double(cn.box.area() / root_area) * cnt;
            } else {
                imask |= 1u << s;
                set_byte(w, 6, s, 1);
                uint32_t id = uint32_t(out.nodes.size());
                out.nodes.push_back(Bvh8Node{});
                queue.push_back({id, ch[size_t(c)], pend.depth + 1});
            }
        }
        if (tri_off > 24) throw std::runtime_error("bvh: leaf triangle budget exceeded");
        w[3] |= imask << 24;
        out.nodes[pend.node8] = node;
    }
    out.sah_cost = sah;
    return out;
}

}// namespace vs
