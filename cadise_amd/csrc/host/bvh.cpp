// Host BVH build: the reference's binary SAH tree, then a wide collapse.
#include "bvh.hpp"

#include <algorithm>
#include <cmath>
#include <cfloat>
#include <cstring>
#include <limits>

namespace mpt {

namespace {

// AABB3R. The reference's empty box initialises max to
// numeric_limits<float>::min() (+1.2e-38, tAabb.ipp:13-17), so a box of
// geometry with negative coordinates always reaches the origin; with
// `quirk` false the empty max is -FLT_MAX (a true empty box). Unions use
// std::min / std::max as TVector::min / max do.
struct Box {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3];
    explicit Box(bool quirk = true) {
        const float e = quirk ? FLT_MIN : -FLT_MAX;
        hi[0] = hi[1] = hi[2] = e;
    }
    void add(const Box& o) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], o.lo[k]);
            hi[k] = std::max(hi[k], o.hi[k]);
        }
    }
    void add(const float p[3]) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    float area() const {
        const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
        const float half = ex * ey + ex * ez + ey * ez;
        return 2.f * half;
    }
    // TVector::maxDimension on the extent: first strictly larger wins
    int max_axis() const {
        float e[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
        int d = 0;
        float m = e[0];
        if (e[1] > m) { m = e[1]; d = 1; }
        if (e[2] > m) { m = e[2]; d = 2; }
        return d;
    }
};

struct Info {
    Box     bound;
    float   centroid[3];
    int64_t index;
};

// Triangle / Rectangle / Sphere::evaluateBound: the vertices' box (the
// sphere's centre -+ radius) padded by 1e-4 (triangle.cpp:45-52,
// rectangle.cpp:39-47, sphere.cpp:33-40).
Box prim_bound(const HostScene& s, int64_t i) {
    const float* v = &s.prim_v[9 * i];
    Box b(false);
    if (s.prim_type[i] == PRIM_SPHERE) {
        for (int k = 0; k < 3; ++k) {
            b.lo[k] = (v[k] - v[3]) - 0.0001f;
            b.hi[k] = (v[k] + v[3]) + 0.0001f;
        }
        return b;
    }
    for (int k = 0; k < 3; ++k) b.lo[k] = b.hi[k] = v[k];
    b.add(v + 3);
    b.add(v + 6);
    if (s.prim_type[i] == PRIM_RECTANGLE) {
        // vD = vB + eA + eB with eA = vA - vB, eB = vC - vB (rectangle.cpp:29-32)
        float d[3];
        for (int k = 0; k < 3; ++k) {
            const float ea = v[k] - v[3 + k], eb = v[6 + k] - v[3 + k];
            d[k] = (v[3 + k] + ea) + eb;
        }
        b.add(d);
    }
    for (int k = 0; k < 3; ++k) {
        b.lo[k] -= 0.0001f;
        b.hi[k] += 0.0001f;
    }
    return b;
}

}  // namespace

void reference_root_bound(const HostScene& s, float lo[3], float hi[3]) {
    Box b(true);
    for (int64_t i = 0; i < s.nprims(); ++i) b.add(prim_bound(s, i));
    for (int k = 0; k < 3; ++k) {
        lo[k] = b.lo[k];
        hi[k] = b.hi[k];
    }
}

namespace {

struct Builder {
    std::vector<Info>& infos;
    BinaryBvh&         out;
    bool               sah;
    bool               quirk;

    static size_t bucket(float c, float mn, float inv) {
        const float offset = (c - mn) * inv;
        const size_t idx = static_cast<size_t>(offset * static_cast<float>(12));
        return idx == 12 ? 11 : idx;
    }

    // _canSplitWithEqual: nth_element at (n + 1) / 2 on centroid[axis]
    int64_t split_equal(int64_t begin, int64_t n, int axis) {
        const int64_t size = (n + 1) / 2;
        std::nth_element(infos.begin() + begin, infos.begin() + begin + size, infos.begin() + begin + n,
                         [axis](const Info& a, const Info& b) { return a.centroid[axis] < b.centroid[axis]; });
        return size;
    }

    // _canSplitWithSah: 12 buckets, cost 0.125 + sum(area_i / area * n_i); -1 = make a leaf
    int64_t split_sah(int64_t begin, int64_t n, int axis, const Box& node, const Box& cb) {
        const float mn = cb.lo[axis];
        const float inv = 1.f / (cb.hi[axis] - cb.lo[axis]);
        if (n <= 2) return split_equal(begin, n, axis);
        Box bb[12] = {Box(quirk), Box(quirk), Box(quirk), Box(quirk), Box(quirk), Box(quirk),
                      Box(quirk), Box(quirk), Box(quirk), Box(quirk), Box(quirk), Box(quirk)};
        size_t bc[12] = {0};
        for (int64_t i = begin; i < begin + n; ++i) {
            const size_t k = bucket(infos[i].centroid[axis], mn, inv);
            bb[k].add(infos[i].bound);
            bc[k] += 1;
        }
        const float no_split = 1.f * static_cast<float>(static_cast<size_t>(n));
        const float inv_area = 1.f / node.area();
        size_t best = std::numeric_limits<size_t>::max();
        float best_cost = std::numeric_limits<float>::max();
        for (size_t split = 0; split < 11; ++split) {
            Box a(quirk), b(quirk);
            size_t ca = 0, cbn = 0;
            for (size_t i = 0; i <= split; ++i) { a.add(bb[i]); ca += bc[i]; }
            for (size_t i = split + 1; i < 12; ++i) { b.add(bb[i]); cbn += bc[i]; }
            const float pa = a.area() * inv_area;
            const float pb = b.area() * inv_area;
            const float cost = 0.125f + 1.f * (pa * static_cast<float>(ca) + pb * static_cast<float>(cbn));
            if (cost < best_cost) {
                best_cost = cost;
                best = split;
            }
        }
        if (best_cost >= no_split) return -1;
        auto mid = std::partition(infos.begin() + begin, infos.begin() + begin + n, [&](const Info& x) {
            return bucket(x.centroid[axis], mn, inv) <= best;
        });
        return static_cast<int64_t>(mid - (infos.begin() + begin));
    }

    void leaf(int64_t node, int64_t begin, int64_t n, const Box& b) {
        BinaryNode& nd = out.nodes[node];
        std::memcpy(nd.bmin, b.lo, sizeof(nd.bmin));
        std::memcpy(nd.bmax, b.hi, sizeof(nd.bmax));
        nd.leaf = 1;
        nd.a = static_cast<int64_t>(out.ordered.size());
        nd.b = n;
        for (int64_t i = begin; i < begin + n; ++i) out.ordered.push_back(static_cast<int32_t>(infos[i].index));
    }

    // _buildBinaryNodesRecursively + _buildLinearNodesRecursively in one pass
    void build(int64_t begin, int64_t n, int32_t depth) {
        out.depth = std::max(out.depth, depth);
        const int64_t node = static_cast<int64_t>(out.nodes.size());
        out.nodes.emplace_back();
        Box nb(quirk);
        for (int64_t i = begin; i < begin + n; ++i) nb.add(infos[i].bound);
        if (n <= 1) return leaf(node, begin, n, nb);
        Box cb(quirk);
        for (int64_t i = begin; i < begin + n; ++i) cb.add(infos[i].centroid);
        const int axis = cb.max_axis();
        if (cb.lo[axis] == cb.hi[axis]) return leaf(node, begin, n, nb);
        const int64_t mid = sah ? split_sah(begin, n, axis, nb, cb) : split_equal(begin, n, axis);
        if (mid < 0) return leaf(node, begin, n, nb);
        {
            BinaryNode& nd = out.nodes[node];
            std::memcpy(nd.bmin, nb.lo, sizeof(nd.bmin));
            std::memcpy(nd.bmax, nb.hi, sizeof(nd.bmax));
            nd.leaf = 0;
            nd.b = axis;
        }
        build(begin, mid, depth + 1);
        out.nodes[node].a = static_cast<int64_t>(out.nodes.size());
        build(begin + mid, n - mid, depth + 1);
    }
};

}  // namespace

void build_binary(const HostScene& s, bool reference_quirk, BinaryBvh* out) {
    const int64_t n = s.nprims();
    std::vector<Info> infos(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
        Info& f = infos[i];
        f.bound = prim_bound(s, i);
        for (int k = 0; k < 3; ++k) f.centroid[k] = (f.bound.lo[k] + f.bound.hi[k]) * 0.5f;
        f.index = i;
    }
    out->nodes.clear();
    out->ordered.clear();
    out->nodes.reserve(static_cast<size_t>(2 * n + 1));
    out->ordered.reserve(static_cast<size_t>(n));
    out->depth = 0;
    Builder b{infos, *out, s.split_sah != 0, reference_quirk};
    b.build(0, n, 0);
}

// ---- wide collapse -------------------------------------------------------------

namespace {

struct Tight {
    float lo[3], hi[3];
};

// SAH-optimal collapse (Ylitie, Karras and Laine, "Efficient Incoherent Ray
// Traversal on GPUs Through Compressed Wide BVHs", HPG 2017, section 3.1):
// cost(n, i) is the least SAH cost of binary subtree n spread over at most i
// child slots of its wide parent. cost(n, 1) = min(leaf, A(n) c_node +
// dist(n, W)), a leaf only for at most leaf_max primitives (cost A(n) N c_prim)
// or when n is a binary leaf (kept whole, whatever its size);
// cost(n, i > 1) = min(cost(n, 1), dist(n, i)); dist(n, j) = min over k of
// cost(left, k) + cost(right, j - k). The wide tree is rebuilt top-down from
// the recorded choices. Leaves stay contiguous binary subtrees, so primitive
// order is unchanged.
template <int W>
struct SahCollapser {
    const BinaryBvh&            bin;
    const std::vector<Tight>&   tight;
    const std::vector<int64_t>& first;
    const std::vector<int64_t>& count;
    int    leaf_max;
    double c_node, c_prim;
    std::vector<double> cost;     // [n * W + i - 1]
    std::vector<int8_t> split;    // [n * W + i - 1]: 0 = one slot (cost(n, 1)), k = k slots to the left child
    std::vector<uint8_t> as_leaf; // per n: cost(n, 1) is the leaf
    std::vector<WideNode<W>> nodes;
    int32_t depth = 0;

    double area(int64_t n) const {
        const Tight& t = tight[n];
        const double ex = static_cast<double>(t.hi[0]) - t.lo[0], ey = static_cast<double>(t.hi[1]) - t.lo[1],
                     ez = static_cast<double>(t.hi[2]) - t.lo[2];
        return std::max(0.0, ex * ey + ex * ez + ey * ez);
    }
    void solve() {
        const int64_t nn = static_cast<int64_t>(bin.nodes.size());
        cost.assign(static_cast<size_t>(nn) * W, 0.0);
        split.assign(static_cast<size_t>(nn) * W, 0);
        as_leaf.assign(static_cast<size_t>(nn), 0);
        const double inf = std::numeric_limits<double>::infinity();
        for (int64_t n = nn - 1; n >= 0; --n) {   // pre-order: children after their parent
            const double A = area(n);
            const double leaf_cost = A * c_prim * static_cast<double>(count[n]);
            double* c = &cost[static_cast<size_t>(n) * W];
            int8_t* sp = &split[static_cast<size_t>(n) * W];
            if (bin.nodes[n].leaf) {   // a binary leaf cannot be split: a leaf whatever its size
                as_leaf[n] = 1;
                for (int i = 0; i < W; ++i) c[i] = leaf_cost;
                continue;
            }
            const double leaf = count[n] <= leaf_max ? leaf_cost : inf;
            const double* cl = &cost[static_cast<size_t>(n + 1) * W];
            const double* cr = &cost[static_cast<size_t>(bin.nodes[n].a) * W];
            double dist[W + 1];
            int8_t dk[W + 1];
            for (int j = 2; j <= W; ++j) {
                dist[j] = inf;
                dk[j] = 1;
                for (int k = 1; k < j; ++k) {
                    const double v = cl[k - 1] + cr[j - k - 1];
                    if (v < dist[j]) { dist[j] = v; dk[j] = static_cast<int8_t>(k); }
                }
            }
            const double node = A * c_node + dist[W];   // finite: binary leaves cost finitely
            as_leaf[n] = leaf < node ? 1 : 0;
            c[0] = std::min(leaf, node);
            sp[0] = 0;
            for (int i = 2; i <= W; ++i) {
                if (dist[i] < c[0]) { c[i - 1] = dist[i]; sp[i - 1] = dk[i]; }
                else { c[i - 1] = c[0]; sp[i - 1] = 0; }
            }
        }
    }
    void gather(int64_t n, int i, std::vector<int64_t>* out) const {
        const int k = split[static_cast<size_t>(n) * W + i - 1];
        if (k == 0) {
            out->push_back(n);
            return;
        }
        gather(n + 1, k, out);
        gather(bin.nodes[n].a, i - k, out);
    }
    // the wide node of binary node n (n interior, opened over W slots)
    int32_t make(int64_t n, int32_t d) {
        depth = std::max(depth, d);
        std::vector<int64_t> kids;
        if (bin.nodes[n].leaf) {
            kids.push_back(n);
        } else {
            const double* cl = &cost[static_cast<size_t>(n + 1) * W];
            const double* cr = &cost[static_cast<size_t>(bin.nodes[n].a) * W];
            int bk = 1;
            double best = std::numeric_limits<double>::infinity();
            for (int k = 1; k < W; ++k) {
                const double v = cl[k - 1] + cr[W - k - 1];
                if (v < best) { best = v; bk = k; }
            }
            gather(n + 1, bk, &kids);
            gather(bin.nodes[n].a, W - bk, &kids);
        }
        const int32_t idx = static_cast<int32_t>(nodes.size());
        nodes.emplace_back();
        for (int i = 0; i < W; ++i) {
            WideNode<W>& nd = nodes[idx];
            nd.lo_x[i] = nd.lo_y[i] = nd.lo_z[i] = FLT_MAX;
            nd.hi_x[i] = nd.hi_y[i] = nd.hi_z[i] = -FLT_MAX;
            nd.child[i] = -1;
            nd.count[i] = 0;
        }
        for (int i = 0; i < static_cast<int>(kids.size()); ++i) {
            const int64_t c = kids[i];
            int32_t ch, cnt;
            if (as_leaf[c]) {
                ch = static_cast<int32_t>(first[c]);
                cnt = static_cast<int32_t>(count[c]);
                if (cnt == 0) ch = -1;
            } else {
                ch = make(c, d + 1);
                cnt = 0;
            }
            WideNode<W>& nd = nodes[idx];   // re-fetch: make() may reallocate
            const Tight& t = tight[c];
            nd.lo_x[i] = t.lo[0]; nd.lo_y[i] = t.lo[1]; nd.lo_z[i] = t.lo[2];
            nd.hi_x[i] = t.hi[0]; nd.hi_y[i] = t.hi[1]; nd.hi_z[i] = t.hi[2];
            if (cnt == 0 && ch < 0) {
                nd.lo_x[i] = nd.lo_y[i] = nd.lo_z[i] = FLT_MAX;
                nd.hi_x[i] = nd.hi_y[i] = nd.hi_z[i] = -FLT_MAX;
            }
            nd.child[i] = ch;
            nd.count[i] = cnt;
        }
        return idx;
    }
};

// Quantize one float node: origin = the children's minimum corner, the
// smallest power-of-two step per axis that spans the extent in 255 steps, then
// q_lo / q_hi moved outwards until fma(q, step, origin) encloses the tight box
// exactly as the device evaluates it (a coarser step when 255 do not suffice).
template <int W>
bool quantize_node(const WideNode<W>& f, QNode<W>* out) {
    QNode<W> q{};
    const float* lo[3] = {f.lo_x, f.lo_y, f.lo_z};
    const float* hi[3] = {f.hi_x, f.hi_y, f.hi_z};
    uint32_t exps = 0;
    for (int a = 0; a < 3; ++a) {
        float o = FLT_MAX, top = -FLT_MAX;
        for (int c = 0; c < W; ++c) {
            if (f.child[c] < 0) continue;
            o = std::min(o, lo[a][c]);
            top = std::max(top, hi[a][c]);
        }
        if (o == FLT_MAX) o = 0.f, top = 0.f;
        q.origin[a] = o;
        const double ext = static_cast<double>(top) - static_cast<double>(o);
        int e = -126;
        if (ext > 0.0) {
            int x = 0;
            std::frexp(ext / 255.0, &x);   // ext / 255 in [2^(x-1), 2^x)
            e = std::max(-126, x - 1);
            while (e < 127 && std::ldexp(255.0, e) < ext) ++e;
        }
        for (;; ++e) {
            if (e > 127) return false;
            const float step = std::ldexp(1.0f, e);
            bool ok = true;
            for (int c = 0; c < W && ok; ++c) {
                if (f.child[c] < 0) {
                    q.q[a][c] = 255;
                    q.q[3 + a][c] = 0;
                    continue;
                }
                double dl = std::floor((static_cast<double>(lo[a][c]) - o) / step);
                int ql = static_cast<int>(std::max(0.0, std::min(255.0, dl)));
                while (ql > 0 && std::fma(static_cast<float>(ql), step, o) > lo[a][c]) --ql;
                double dh = std::ceil((static_cast<double>(hi[a][c]) - o) / step);
                int qh = static_cast<int>(std::max(0.0, std::min(255.0, dh)));
                while (qh < 255 && std::fma(static_cast<float>(qh), step, o) < hi[a][c]) ++qh;
                if (std::fma(static_cast<float>(ql), step, o) > lo[a][c] ||
                    std::fma(static_cast<float>(qh), step, o) < hi[a][c]) {
                    ok = false;
                    break;
                }
                q.q[a][c] = static_cast<uint8_t>(ql);
                q.q[3 + a][c] = static_cast<uint8_t>(qh);
            }
            if (ok) break;
        }
        exps |= static_cast<uint32_t>(e + 127) << (8 * a);
    }
    q.exps = exps;
    for (int c = 0; c < W; ++c) {
        q.child[c] = f.child[c];
        if (f.count[c] > 65535) return false;
        q.count[c] = static_cast<uint16_t>(f.count[c]);
    }
    *out = q;
    return true;
}

}  // namespace

// SAH costs of the collapse: a wide node visit (W slab tests, sort, push)
// against one primitive test, in units of a node visit.
constexpr double kSahNode = 1.0, kSahPrim = 0.3;

void collapse_wide(const HostScene& s, const BinaryBvh& bin, int width, int leaf_max, WideBvh* out) {
    const int64_t nn = static_cast<int64_t>(bin.nodes.size());
    std::vector<Tight> tight(static_cast<size_t>(nn));
    std::vector<int64_t> first(static_cast<size_t>(nn)), count(static_cast<size_t>(nn));
    // bottom-up: children have larger pre-order indices than their parent
    for (int64_t i = nn - 1; i >= 0; --i) {
        const BinaryNode& n = bin.nodes[i];
        Tight t;
        for (int k = 0; k < 3; ++k) { t.lo[k] = FLT_MAX; t.hi[k] = -FLT_MAX; }
        if (n.leaf) {
            first[i] = n.a;
            count[i] = n.b;
            for (int64_t j = n.a; j < n.a + n.b; ++j) {
                const int64_t p = bin.ordered[j];
                const float* v = &s.prim_v[9 * p];
                if (s.prim_type[p] == PRIM_SPHERE) {
                    for (int k = 0; k < 3; ++k) {
                        t.lo[k] = std::min(t.lo[k], (v[k] - v[3]) - 0.0001f);
                        t.hi[k] = std::max(t.hi[k], (v[k] + v[3]) + 0.0001f);
                    }
                    continue;
                }
                float pts[4][3];
                int np = 3;
                for (int q = 0; q < 3; ++q)
                    for (int k = 0; k < 3; ++k) pts[q][k] = v[3 * q + k];
                if (s.prim_type[p] == PRIM_RECTANGLE) {
                    for (int k = 0; k < 3; ++k) pts[3][k] = (v[3 + k] + (v[k] - v[3 + k])) + (v[6 + k] - v[3 + k]);
                    np = 4;
                }
                for (int q = 0; q < np; ++q)
                    for (int k = 0; k < 3; ++k) {
                        t.lo[k] = std::min(t.lo[k], pts[q][k] - 0.0001f);
                        t.hi[k] = std::max(t.hi[k], pts[q][k] + 0.0001f);
                    }
            }
        } else {
            const int64_t l = i + 1, r = n.a;
            first[i] = first[l];
            count[i] = count[l] + count[r];
            for (int k = 0; k < 3; ++k) {
                t.lo[k] = std::min(tight[l].lo[k], tight[r].lo[k]);
                t.hi[k] = std::max(tight[l].hi[k], tight[r].hi[k]);
            }
        }
        tight[i] = t;
    }
    out->width = width;
    if (width == 8) {
        SahCollapser<8> c{bin, tight, first, count, leaf_max, kSahNode, kSahPrim};
        if (nn > 0) {
            c.solve();
            c.make(0, 0);
        }
        out->nnodes = static_cast<int64_t>(c.nodes.size());
        out->depth = c.depth;
        out->nodes.resize(c.nodes.size() * sizeof(QNode<8>));
        QNode<8>* qn = reinterpret_cast<QNode<8>*>(out->nodes.data());
        for (size_t i = 0; i < c.nodes.size(); ++i)
            if (!quantize_node<8>(c.nodes[i], &qn[i])) out->ok = false;
    } else {
        SahCollapser<4> c{bin, tight, first, count, leaf_max, kSahNode, kSahPrim};
        if (nn > 0) {
            c.solve();
            c.make(0, 0);
        }
        out->nnodes = static_cast<int64_t>(c.nodes.size());
        out->depth = c.depth;
        out->nodes.resize(c.nodes.size() * sizeof(QNode<4>));
        QNode<4>* qn = reinterpret_cast<QNode<4>*>(out->nodes.data());
        for (size_t i = 0; i < c.nodes.size(); ++i)
            if (!quantize_node<4>(c.nodes[i], &qn[i])) out->ok = false;
        out->width = 4;
    }
    out->max_stack = (out->width - 1) * (out->depth + 1) + 1;
}

}  // namespace mpt
