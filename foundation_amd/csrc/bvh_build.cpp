// bvh_build.cpp — binned-SAH binary build, 4-wide collapse, 8-bit quantisation.
//
// Algorithms (the reference has no BVH): binned SAH after Wald, "On fast
// Construction of SAH-based Bounding Volume Hierarchies", IEEE RT 2007; the
// wide collapse greedily opens the child of largest surface area until a node
// has 4 slots (as in Wald et al., "Getting Rid of Packets", IEEE RT 2008);
// quantisation after Ylitie et al., HPG 2017, with outward rounding.
#include "bvh_build.hpp"
#include "host_pool.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>

namespace fnd {
namespace {

constexpr int kBins = 32;

struct BNode {
    Box box;
    int32_t left = -1;   // internal: children left, right
    int32_t right = -1;
    uint32_t first = 0;  // leaf: range in idx
    uint32_t count = 0;
};

inline void box_empty(Box& b)
{
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = std::numeric_limits<float>::infinity();
        b.hi[a] = -std::numeric_limits<float>::infinity();
    }
}
inline void box_grow(Box& b, const Box& o)
{
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = std::min(b.lo[a], o.lo[a]);
        b.hi[a] = std::max(b.hi[a], o.hi[a]);
    }
}
inline double box_area(const Box& b)
{
    const double dx = double(b.hi[0]) - b.lo[0], dy = double(b.hi[1]) - b.lo[1], dz = double(b.hi[2]) - b.lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0;
    return 2.0 * (dx * dy + dy * dz + dz * dx);
}

struct Task {
    uint32_t node, begin, end;
};

class BinaryBuilder {
public:
    BinaryBuilder(const Box* boxes, uint32_t n, uint32_t leaf_max)
        : boxes_(boxes), n_(n), leaf_max_(std::max(1u, leaf_max)), cent_(3ull * n), idx_(n), nodes_(2ull * n + 1)
    {
        for (uint32_t i = 0; i < n; ++i) {
            idx_[i] = i;
            for (int a = 0; a < 3; ++a) cent_[3ull * i + a] = 0.5f * (boxes[i].lo[a] + boxes[i].hi[a]);
        }
        count_.store(1);
    }

    // Splits breadth-first until there are enough independent subtrees, then
    // builds them in parallel.  Each split decision depends only on the
    // primitives of its range, so the tree is the same for any thread count.
    void build(unsigned threads)
    {
        std::vector<Task> frontier{{0, 0, n_}};
        const uint32_t par_min = 32768;
        if (threads > 1) {
            for (int round = 0; round < 8; ++round) {
                std::vector<Task> next;
                bool split_any = false;
                for (const Task& t : frontier) {
                    if (t.end - t.begin < par_min || next.size() + frontier.size() > 8u * threads) {
                        next.push_back(t);
                        continue;
                    }
                    Task l, r;
                    if (split(t, l, r)) {
                        next.push_back(l);
                        next.push_back(r);
                        split_any = true;
                    }
                }
                frontier.swap(next);
                if (!split_any || frontier.size() >= 4u * threads) break;
            }
        }
        std::sort(frontier.begin(), frontier.end(),
                  [](const Task& a, const Task& b) { return (a.end - a.begin) > (b.end - b.begin); });
        std::atomic<size_t> next{0};
        auto worker = [&]() {
            std::vector<Task> stack;
            for (;;) {
                const size_t i = next.fetch_add(1);
                if (i >= frontier.size()) break;
                stack.push_back(frontier[i]);
                while (!stack.empty()) {
                    Task t = stack.back();
                    stack.pop_back();
                    Task l, r;
                    if (split(t, l, r)) {
                        stack.push_back(r);
                        stack.push_back(l);
                    }
                }
            }
        };
        const unsigned nt = std::max(1u, std::min<unsigned>(threads, unsigned(frontier.size())));
        run_pool(nt, worker);
    }

    const std::vector<BNode>& nodes() const { return nodes_; }
    const std::vector<uint32_t>& idx() const { return idx_; }

private:
    // Fills node t.node; returns true with the two child tasks if it split.
    bool split(const Task& t, Task& l, Task& r)
    {
        BNode& nd = nodes_[t.node];
        Box bb, cb;
        box_empty(bb);
        box_empty(cb);
        for (uint32_t i = t.begin; i < t.end; ++i) {
            const uint32_t p = idx_[i];
            box_grow(bb, boxes_[p]);
            for (int a = 0; a < 3; ++a) {
                cb.lo[a] = std::min(cb.lo[a], cent_[3ull * p + a]);
                cb.hi[a] = std::max(cb.hi[a], cent_[3ull * p + a]);
            }
        }
        nd.box = bb;
        const uint32_t cnt = t.end - t.begin;
        if (cnt <= leaf_max_) {
            nd.first = t.begin;
            nd.count = cnt;
            return false;
        }
        int best_axis = -1, best_split = -1;
        double best_cost = std::numeric_limits<double>::infinity();
        Box bin_box[kBins];
        uint32_t bin_cnt[kBins];
        Box right_box[kBins];
        uint32_t right_cnt[kBins];
        for (int a = 0; a < 3; ++a) {
            const float ext = cb.hi[a] - cb.lo[a];
            if (!(ext > 0.0f)) continue;
            const float scale = float(kBins) / ext;
            for (int k = 0; k < kBins; ++k) {
                box_empty(bin_box[k]);
                bin_cnt[k] = 0;
            }
            for (uint32_t i = t.begin; i < t.end; ++i) {
                const uint32_t p = idx_[i];
                int k = int((cent_[3ull * p + a] - cb.lo[a]) * scale);
                k = std::min(std::max(k, 0), kBins - 1);
                box_grow(bin_box[k], boxes_[p]);
                bin_cnt[k]++;
            }
            Box acc;
            box_empty(acc);
            uint32_t accn = 0;
            for (int k = kBins - 1; k > 0; --k) {
                box_grow(acc, bin_box[k]);
                accn += bin_cnt[k];
                right_box[k] = acc;
                right_cnt[k] = accn;
            }
            box_empty(acc);
            accn = 0;
            for (int k = 0; k < kBins - 1; ++k) {
                box_grow(acc, bin_box[k]);
                accn += bin_cnt[k];
                if (!accn || !right_cnt[k + 1]) continue;
                const double cost = box_area(acc) * accn + box_area(right_box[k + 1]) * right_cnt[k + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = a;
                    best_split = k;
                }
            }
        }
        uint32_t mid = t.begin + cnt / 2;
        if (best_axis >= 0) {
            const float ext = cb.hi[best_axis] - cb.lo[best_axis];
            const float scale = float(kBins) / ext;
            auto it = std::partition(idx_.begin() + t.begin, idx_.begin() + t.end, [&](uint32_t p) {
                int k = int((cent_[3ull * p + best_axis] - cb.lo[best_axis]) * scale);
                k = std::min(std::max(k, 0), kBins - 1);
                return k <= best_split;
            });
            mid = uint32_t(it - idx_.begin());
            if (mid == t.begin || mid == t.end) mid = t.begin + cnt / 2;
        }
        const uint32_t li = count_.fetch_add(2);
        nd.left = int32_t(li);
        nd.right = int32_t(li + 1);
        nd.count = 0;
        l = {li, t.begin, mid};
        r = {li + 1, mid, t.end};
        return true;
    }

    const Box* boxes_;
    uint32_t n_;
    uint32_t leaf_max_;
    std::vector<float> cent_;
    std::vector<uint32_t> idx_;
    std::vector<BNode> nodes_;
    std::atomic<uint32_t> count_{0};
};

// Smallest e with 255 * 2^e >= ext, clamped to a float-normal range.
int8_t choose_exponent(float ext)
{
    if (!(ext > 0.0f)) return -100;
    const double q = double(ext) / 255.0;
    int e;
    std::frexp(q, &e);  // q = m * 2^e, m in [0.5, 1)
    int ee = (std::ldexp(1.0, e - 1) >= q) ? e - 1 : e;
    ee = std::max(-100, std::min(127, ee));
    return int8_t(ee);
}

void encode_node(WideNode& w, const Box* child, const bool* internal, int nc)
{
    std::memset(&w, 0, sizeof(w));
    Box nb;
    box_empty(nb);
    for (int k = 0; k < nc; ++k) box_grow(nb, child[k]);
    const float p[3] = {nb.lo[0], nb.lo[1], nb.lo[2]};
    int8_t e[3];
    for (int a = 0; a < 3; ++a) e[a] = choose_exponent(nb.hi[a] - nb.lo[a]);
    w.px = p[0];
    w.py = p[1];
    w.pz = p[2];
    w.ex = e[0];
    w.ey = e[1];
    w.ez = e[2];
    for (int a = 0; a < 3; ++a) {
        uint32_t lo_word = 0, hi_word = 0;
        const double s = std::ldexp(1.0, e[a]);
        for (int k = 0; k < 4; ++k) {
            uint32_t ql = 255, qh = 0;  // empty slot
            if (k < nc) {
                const double lo = std::floor((double(child[k].lo[a]) - double(p[a])) / s);
                const double hi = std::ceil((double(child[k].hi[a]) - double(p[a])) / s);
                ql = uint32_t(std::max(0.0, std::min(255.0, lo)));
                qh = uint32_t(std::max(0.0, std::min(255.0, hi)));
            }
            lo_word |= ql << (8 * k);
            hi_word |= qh << (8 * k);
        }
        w.qlo[a] = lo_word;
        w.qhi[a] = hi_word;
    }
    uint8_t imask = 0;
    for (int k = 0; k < nc; ++k) {
        imask |= uint8_t(0x10u << k);  // slot k holds a child
        if (internal[k]) imask |= uint8_t(1u << k);
    }
    w.imask = imask;
}

}  // namespace

bool build_wide_tree(const Box* boxes, uint32_t n, uint32_t leaf_max, unsigned threads, WideTree& out,
                     std::string& err)
{
    out = WideTree{};
    if (n == 0) {
        err = "empty primitive set";
        return false;
    }
    if (leaf_max < 1 || leaf_max > 7) {
        err = "leaf_max must be in [1, 7]";
        return false;
    }
    BinaryBuilder bb(boxes, n, leaf_max);
    bb.build(std::max(1u, threads));
    const auto& bn = bb.nodes();
    const auto& idx = bb.idx();
    out.bounds = bn[0].box;
    out.order.reserve(n);

    struct Item {
        uint32_t wide;
        int32_t bin;
        uint32_t depth;
    };
    std::vector<Item> stack;
    out.nodes.emplace_back();
    stack.push_back({0, 0, 1});
    while (!stack.empty()) {
        const Item it = stack.back();
        stack.pop_back();
        out.max_depth = std::max(out.max_depth, it.depth);
        int32_t ch[4];
        int nc = 0;
        const BNode& b = bn[it.bin];
        if (b.count) {
            ch[nc++] = it.bin;  // a root that is already a leaf
        } else {
            ch[nc++] = b.left;
            ch[nc++] = b.right;
            while (nc < 4) {
                int pick = -1;
                double best = -1.0;
                for (int k = 0; k < nc; ++k) {
                    if (bn[ch[k]].count) continue;
                    const double a = box_area(bn[ch[k]].box);
                    if (a > best) {
                        best = a;
                        pick = k;
                    }
                }
                if (pick < 0) break;
                const BNode& o = bn[ch[pick]];
                ch[pick] = o.left;
                ch[nc++] = o.right;
            }
        }
        Box cbox[4];
        bool internal[4];
        uint32_t entry[4] = {0, 0, 0, 0};
        int n_int = 0;
        for (int k = 0; k < nc; ++k) {
            cbox[k] = bn[ch[k]].box;
            internal[k] = bn[ch[k]].count == 0;
            n_int += internal[k] ? 1 : 0;
        }
        const uint32_t child_base = uint32_t(out.nodes.size());
        const uint32_t prim_base = uint32_t(out.order.size());
        for (int k = 0; k < nc; ++k) {
            if (internal[k]) continue;
            const BNode& lf = bn[ch[k]];
            entry[k] = kEntryTri | (uint32_t(out.order.size()) << 3) | lf.count;
            for (uint32_t j = 0; j < lf.count; ++j) out.order.push_back(idx[lf.first + j]);
        }
        for (int k = 0, rank = 0; k < nc; ++k)
            if (internal[k]) entry[k] = child_base + uint32_t(rank++);
        out.nodes.resize(out.nodes.size() + n_int);
        WideNode& w = out.nodes[it.wide];
        encode_node(w, cbox, internal, nc);
        for (int k = 0; k < 4; ++k) w.entry[k] = entry[k];
        w.child_base = child_base;
        w.prim_base = prim_base;
        int rank = n_int;
        for (int k = nc - 1; k >= 0; --k) {
            if (!internal[k]) continue;
            --rank;
            stack.push_back({child_base + uint32_t(rank), ch[k], it.depth + 1});
        }
    }
    if (out.order.size() != n) {
        err = "internal error: leaf payload count mismatch";
        return false;
    }
    return true;
}

namespace {

void widen_by_mapped_corners(const double* blo, const double* bhi, const float* m, double* lo, double* hi)
{
    for (int c = 0; c < 8; ++c) {
        const double p[3] = {(c & 1) ? bhi[0] : blo[0], (c & 2) ? bhi[1] : blo[1], (c & 4) ? bhi[2] : blo[2]};
        for (int i = 0; i < 3; ++i) {
            const double w = double(m[4 * i]) * p[0] + double(m[4 * i + 1]) * p[1] + double(m[4 * i + 2]) * p[2] +
                             double(m[4 * i + 3]);
            lo[i] = std::min(lo[i], w);
            hi[i] = std::max(hi[i], w);
        }
    }
}

Box padded_outward(const double* lo, const double* hi)
{
    Box out;
    for (int i = 0; i < 3; ++i) {
        const double pad = (hi[i] - lo[i]) * 0x1p-10 + 1e-30;
        out.lo[i] = std::nextafter(float(lo[i] - pad), -INFINITY);
        out.hi[i] = std::nextafter(float(hi[i] + pad), INFINITY);
    }
    return out;
}

void collect_cover(const WideTree& t, uint32_t node, int depth, std::vector<DBox>& out)
{
    const WideNode& w = t.nodes[node];
    const int8_t e[3] = {w.ex, w.ey, w.ez};
    const float p[3] = {w.px, w.py, w.pz};
    for (int k = 0; k < 4; ++k) {
        if (!((w.imask >> (4 + k)) & 1u)) continue;
        if (((w.imask >> k) & 1u) && depth > 1) {
            collect_cover(t, w.entry[k], depth - 1, out);
            continue;
        }
        // p + q * 2^e is exact in double and bounds the slot's content (the
        // encoder rounded the planes outward)
        DBox b;
        for (int a = 0; a < 3; ++a) {
            const double s = std::ldexp(1.0, e[a]);
            b.lo[a] = double(p[a]) + double((w.qlo[a] >> (8 * k)) & 0xffu) * s;
            b.hi[a] = double(p[a]) + double((w.qhi[a] >> (8 * k)) & 0xffu) * s;
        }
        out.push_back(b);
    }
}

}  // namespace

std::vector<DBox> cover_boxes(const WideTree& t, int depth)
{
    std::vector<DBox> out;
    if (!t.nodes.empty()) collect_cover(t, 0, std::max(1, depth), out);
    return out;
}

Box transform_cover_conservative(const std::vector<DBox>& cover, const float* m)
{
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (const DBox& b : cover) widen_by_mapped_corners(b.lo, b.hi, m, lo, hi);
    return padded_outward(lo, hi);
}

}  // namespace fnd
