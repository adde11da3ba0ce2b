// Oracle — TEST INFRASTRUCTURE ONLY (see vec.hpp for the rules).
//
// Restatement of RAPTor's AAC ("approximate agglomerative clustering") BVH
// builder, ssd/bvh_builder.{h,cc}. The packed lower-triangular cost matrix,
// its in-place compaction and the lane-ordered min_element are kept as the
// reference has them because they decide which clusters merge on ties, and so
// the tree the CPU baseline traverses.
#include <cassert>
#include <cmath>

#include "oracle.hpp"

namespace oracle {

namespace {

enum class Axis { x, y, z };

struct Bound {
    V3 low, high;
};

class AacBuilder {
  public:
    AacBuilder(Scene *s, const AacParams &p) : s_(s), p_(p) {}

    int build();

  private:
    int &ind(int i) { return s_->indirect[i]; }
    std::vector<BvhNode> &nodes() { return s_->nodes; }

    // ssd/bvh_builder.cc:181-185
    int reduction_function(int n) const {
        const float r = std::max(1.0f, (std::pow(p_.delta, 0.5f + eps_) * std::pow(static_cast<float>(n), p_.alpha) * 0.5f));
        return static_cast<int>(r);
    }
    // ssd/bvh_builder.cc:190-193
    float cost_function(const BvhNode &l, const BvhNode &r) const { return combined_surface_area(l, r); }
    // ssd/bvh_builder.cc:195-206
    float cost_function(const V3 &low, const V3 &high, float count, int tri_idx) {
        const Bound &b = bounds_[ind(tri_idx)];
        const V3 bl = vmin(low, b.low);
        const V3 tr = vmax(high, b.high);
        const V3 dist = tr - bl;
        return (((dist.x * dist.y) + (dist.x * dist.z)) + (dist.y * dist.z)) * count;
    }
    // ssd/bvh_builder.cc:208-239: split the current axis at its midpoint and
    // move to the next axis in z -> y -> x -> z order.
    static Axis divide_spatial_bounds(V3 *bm, V3 *tm, const V3 &bl, const V3 &tr, Axis axis) {
        switch (axis) {
            case Axis::x: {
                const float split = (tr.x + bl.x) * 0.5f;
                bm->x = split;
                tm->x = split;
                return Axis::z;
            }
            case Axis::y: {
                const float split = (tr.y + bl.y) * 0.5f;
                bm->y = split;
                tm->y = split;
                return Axis::x;
            }
            default: {
                const float split = (tr.z + bl.z) * 0.5f;
                bm->z = split;
                tm->z = split;
                return Axis::y;
            }
        }
    }

    int build_leaf_node(int *cost_b, int *cost_e, int b, int e);
    int build_layer(int *cost_b, int *cost_e, const V3 &bl, const V3 &tr, int b, int e, Axis axis);
    int build_layer_primitive(int *cost_b, int *cost_e, const V3 &bl, const V3 &tr, int b, int e, Axis axis);
    void combine_nodes(int *cost_b, int *cost_e, int layer_size, int cost_start_idx, int cost_begin, int cost_m,
                       int cost_end);

    Scene *s_;
    AacParams p_;
    const float eps_ = 1e-8f;
    float max_leaf_sah_ = 0.0f;
    int depth_ = 0;
    int next_node_ = 0;
    std::vector<Bound> bounds_;
    std::vector<int> code_buffer_;
    std::vector<float> cost_matrix_;
    std::vector<int> cost_addrs_;
};

// ssd/bvh_builder.cc:241-290
int AacBuilder::build_leaf_node(int *cost_b, int *cost_e, int b, int e) {
    *cost_b = static_cast<int>(cost_addrs_.size());
    const int cost_begin = *cost_b;
    for (int i = b; i < e;) {
        int node_end = i + 1;
        int layer_top = e;
        V3 low = bounds_[ind(i)].low;
        V3 high = bounds_[ind(i)].high;
        while (node_end < layer_top) {
            if (cost_function(low, high, static_cast<float>(node_end - i + 1), node_end) > max_leaf_sah_) {
                std::swap(ind(node_end), ind(--layer_top));
            } else {
                low = vmin(low, bounds_[ind(node_end)].low);
                high = vmax(high, bounds_[ind(node_end)].high);
                ++node_end;
            }
        }
        const int node_idx = next_node_++;
        cost_addrs_.push_back(node_idx);
        BvhNode &n = nodes()[node_idx];
        n.low = low;
        n.high = high;
        n.left_b = i;
        n.right_e = node_end;
        n.leaf = true;
        i = node_end;
    }
    *cost_e = static_cast<int>(cost_addrs_.size());
    const int cost_end = *cost_e;

    const int cost_start_idx = static_cast<int>(cost_matrix_.size());
    for (int i = cost_begin + 1; i < cost_end; ++i) {
        for (int j = cost_begin; j < i; ++j) {
            cost_matrix_.push_back(cost_function(nodes()[cost_addrs_[i]], nodes()[cost_addrs_[j]]));
        }
    }
    combine_nodes(cost_b, cost_e, reduction_function(std::max(cost_end - cost_begin, static_cast<int>(p_.delta))),
                  cost_start_idx, cost_begin, cost_end, cost_end);
    return cost_start_idx;
}

// ssd/bvh_builder.cc:292-405
void AacBuilder::combine_nodes(int *cost_b, int *cost_e, int layer_size, int cost_start_idx, int cost_begin,
                               int cost_m, int cost_end) {
    // Merge the left triangle [cost_begin, cost_m) and the right triangle
    // [cost_m, cost_end), which lie back to back, into one triangle for
    // [cost_begin, cost_end), computing the new cross costs.
    {
        const int cost_m_base = cost_m - cost_begin;
        const int cost_end_base = cost_end - cost_begin;
        const int cost_m_idx = cost_start_idx + (((cost_m_base - 1) * cost_m_base) >> 1);
        const int cost_end_idx = cost_start_idx + (((cost_end_base - 1) * cost_end_base) >> 1);
        const int tri_1_rows = cost_end - cost_m;
        cost_matrix_.resize(cost_end_idx);
        int to_idx = cost_end_idx - 1;
        int right_idx = cost_m_idx + (((tri_1_rows - 1) * tri_1_rows) >> 1) - 1;
        for (int i = cost_end - 1; i >= cost_m; --i) {
            for (int j = i; j > cost_m; --j) cost_matrix_[to_idx--] = cost_matrix_[right_idx--];
            for (int j = cost_m - 1; j >= cost_begin; --j) {
                cost_matrix_[to_idx--] = cost_function(nodes()[cost_addrs_[i]], nodes()[cost_addrs_[j]]);
            }
        }
    }

    // Greedy merging of the cheapest pair until layer_size clusters remain.
    while ((cost_end - cost_begin) > layer_size) {
        const int cost_end_m1 = cost_end - 1;
        const int ceb = cost_end_m1 - cost_begin;
        const int cost_end_idx = cost_start_idx + ((ceb * (ceb + 1)) >> 1);
        float lowest_cost = std::numeric_limits<float>::max();
        const long addr = min_element_range(cost_matrix_.data(), &lowest_cost, cost_start_idx, cost_end_idx) -
                          cost_start_idx;
        const int addr_i = (static_cast<int>(std::sqrt(static_cast<double>(1 + (8 * addr)))) - 1) >> 1;
        const int lowest_i_base = addr_i + 1;
        const int lowest_j_base = static_cast<int>(addr - ((addr_i * (addr_i + 1)) >> 1));
        const int lowest_i = cost_begin + lowest_i_base;
        const int lowest_j = cost_begin + lowest_j_base;

        // ssd/bvh_node.h:38-46 create_generic_node(nodes, left, right).
        const int node_idx = next_node_++;
        {
            BvhNode &n = nodes()[node_idx];
            const BvhNode &l = nodes()[cost_addrs_[lowest_i]];
            const BvhNode &r = nodes()[cost_addrs_[lowest_j]];
            n.low = vmin(l.low, r.low);
            n.high = vmax(l.high, r.high);
            n.left_b = cost_addrs_[lowest_i];
            n.right_e = cost_addrs_[lowest_j];
            n.leaf = false;
        }
        cost_addrs_[lowest_i] = cost_addrs_.back();
        cost_addrs_[lowest_j] = node_idx;
        cost_addrs_.pop_back();
        cost_end = cost_end_m1;

        // Costs against the new node: its row, then its column.
        const int lowest_j_m1 = lowest_j_base - 1;
        int row_idx = cost_start_idx + ((lowest_j_m1 * lowest_j_base) >> 1);
        for (int i = cost_begin; i < lowest_j; ++i) {
            cost_matrix_[row_idx++] = cost_function(nodes()[node_idx], nodes()[cost_addrs_[i]]);
        }
        row_idx += (lowest_j - cost_begin);
        for (int i = lowest_j + 1; i < cost_end; ++i) {
            cost_matrix_[row_idx] = cost_function(nodes()[node_idx], nodes()[cost_addrs_[i]]);
            row_idx += (i - cost_begin);
        }

        // Move the last row/column into the space of lowest_i.
        const int erase_idx = cost_start_idx + (((ceb - 1) * ceb) >> 1);
        if (lowest_i < cost_end) {
            const int size = static_cast<int>(cost_matrix_.size());
            row_idx = cost_start_idx + ((addr_i * lowest_i_base) >> 1);
            int top_row = erase_idx;
            for (int i = cost_begin; i < lowest_i; ++i) {
                if (i != lowest_j) cost_matrix_[row_idx] = cost_matrix_[top_row];
                ++row_idx;
                ++top_row;
            }
            ++top_row;
            row_idx += (lowest_i - cost_begin);
            for (int i = lowest_i + 1; i < (cost_end + 1); ++i) {
                // The final step targets the row being erased; the reference
                // reads one element past the matrix there. Skip that read.
                if (row_idx < erase_idx && top_row < size) cost_matrix_[row_idx] = cost_matrix_[top_row];
                ++top_row;
                row_idx += (i - cost_begin);
            }
        }
        cost_matrix_.resize(erase_idx);
    }
    *cost_e = cost_end;
    *cost_b = cost_begin;
}

// ssd/bvh_builder.cc:407-476
int AacBuilder::build_layer(int *cost_b, int *cost_e, const V3 &bl, const V3 &tr, int b, int e, Axis axis) {
    assert(depth_ < p_.max_down_phase_depth);
    ++depth_;
    const int node_size = e - b;
    if ((node_size < p_.delta) || (depth_ == p_.max_down_phase_depth)) {
        const int r = build_leaf_node(cost_b, cost_e, b, e);
        --depth_;
        return r;
    } else if (*cost_b == (*cost_e - 1)) {
        const int r = build_layer_primitive(cost_b, cost_e, bl, tr, b, e, axis);
        --depth_;
        return r;
    }

    const int morton_middle = *cost_b + ((*cost_e - *cost_b) >> 1);
    const int middle_idx =
        static_cast<int>(std::lower_bound(code_buffer_.begin() + b, code_buffer_.begin() + e, morton_middle) -
                         code_buffer_.begin());

    V3 bm = bl, tm = tr;
    axis = divide_spatial_bounds(&bm, &tm, bl, tr, axis);

    int cost_start_idx = -1;
    int cost_begin = *cost_b;
    int cost_m = morton_middle;
    if (b != middle_idx) cost_start_idx = build_layer(&cost_begin, &cost_m, bl, tm, b, middle_idx, axis);

    int cost_end = *cost_e;
    int cost_mr = morton_middle;
    if (middle_idx != e) {
        const int right_start_idx = build_layer(&cost_mr, &cost_end, bm, tr, middle_idx, e, axis);
        if (b == middle_idx) cost_start_idx = right_start_idx;
    } else {
        cost_end = cost_m;
    }
    if (b == middle_idx) {
        cost_m = cost_mr;
        cost_begin = cost_mr;
    }
    combine_nodes(cost_b, cost_e, reduction_function(node_size), cost_start_idx, cost_begin, cost_m, cost_end);
    --depth_;
    return cost_start_idx;
}

// ssd/bvh_builder.cc:478-642 — partition about the spatial midpoint of the
// current axis once the Morton codes are exhausted.
int AacBuilder::build_layer_primitive(int *cost_b, int *cost_e, const V3 &bl, const V3 &tr, int b, int e,
                                      Axis axis) {
    assert(depth_ < p_.max_down_phase_depth);
    ++depth_;
    const int node_size = e - b;
    if ((node_size < p_.delta) || (depth_ == p_.max_down_phase_depth)) {
        const int r = build_leaf_node(cost_b, cost_e, b, e);
        --depth_;
        return r;
    }

    const int ax = (axis == Axis::x) ? 0 : ((axis == Axis::y) ? 1 : 2);
    auto centre2 = [&](int i) {
        const Bound &bb = bounds_[ind(i)];
        return bb.high[ax] + bb.low[ax];
    };
    const float dbl_split = tr[ax] + bl[ax];
    int top = e - 1;
    int bottom = b;
    while (bottom < top) {
        while (dbl_split < centre2(top) && bottom < top) --top;
        while (dbl_split >= centre2(bottom) && bottom < top) ++bottom;
        if (bottom < top) {
            std::swap(ind(bottom), ind(top));
            --top;
            ++bottom;
        }
    }
    int middle_idx = bottom;
    if ((top == bottom) && (dbl_split >= centre2(top))) ++middle_idx;

    V3 bm = bl, tm = tr;
    axis = divide_spatial_bounds(&bm, &tm, bl, tr, axis);

    int cost_start_idx = -1;
    int cost_begin = *cost_b;
    int cost_m = middle_idx;
    if (b != middle_idx) cost_start_idx = build_layer_primitive(&cost_begin, &cost_m, bl, tm, b, middle_idx, axis);

    int cost_end = *cost_e;
    int cost_mr = middle_idx;
    if (middle_idx != e) {
        const int right_start_idx = build_layer_primitive(&cost_mr, &cost_end, bm, tr, middle_idx, e, axis);
        if (b == middle_idx) cost_start_idx = right_start_idx;
    } else {
        cost_end = cost_m;
    }
    if (b == middle_idx) {
        cost_m = cost_mr;
        cost_begin = cost_mr;
    }
    combine_nodes(cost_b, cost_e, reduction_function(node_size), cost_start_idx, cost_begin, cost_m, cost_end);
    --depth_;
    return cost_start_idx;
}

// ssd/bvh_builder.cc:18-49: triangle bounds, scene box, per-axis cell
// widths and the Morton code of every box centroid.
void aac_codes(const Scene &s, std::vector<Bound> *bounds, V3 *lo, V3 *hi, std::vector<int> *codes) {
    const int n = static_cast<int>(s.tris.size());
    V3 b(kMaxDist, kMaxDist, kMaxDist), t(-kMaxDist, -kMaxDist, -kMaxDist);
    bounds->resize(n);
    for (int i = 0; i < n; ++i) {
        const Triangle &tri = s.tris[i];
        // rt/triangle.h:56-57
        (*bounds)[i].low = vmin(tri.a, vmin(tri.b, tri.c));
        (*bounds)[i].high = vmax(tri.a, vmax(tri.b, tri.c));
        b = vmin(b, (*bounds)[i].low);
        t = vmax(t, (*bounds)[i].high);
    }
    *lo = b;
    *hi = t;
    const V3 scene_width = t - b;
    const V3 width = scene_width * (1.00001f / 1024.0f);
    const V3 width_inv = 1.0f / width;
    codes->resize(n);
    for (int i = 0; i < n; ++i) {
        const V3 c = (((*bounds)[i].high + (*bounds)[i].low) * 0.5f) - b;
        (*codes)[i] = morton_code(c.x, c.y, c.z, width_inv.x, width_inv.y, width_inv.z);
    }
}

// ssd/bvh_builder.cc:18-178
int AacBuilder::build() {
    const int n = static_cast<int>(s_->tris.size());
    nodes().assign(std::max(1, n << 1), BvhNode());
    s_->indirect.resize(n);
    for (int i = 0; i < n; ++i) s_->indirect[i] = i;

    V3 b, t;
    std::vector<int> codes;
    aac_codes(*s_, &bounds_, &b, &t, &codes);
    s_->scene_lo = b;
    s_->scene_hi = t;
    const V3 scene_width = t - b;
    max_leaf_sah_ = ((((scene_width.x * scene_width.y) + (scene_width.x * scene_width.z)) +
                      (scene_width.y * scene_width.z)) *
                     static_cast<float>(n)) *
                    p_.max_leaf_sah_factor;

    std::vector<int> order;
    morton_radix_sort(codes, &code_buffer_, &order);
    s_->indirect = order;

    int cost_b = 0x0;
    int cost_e = 0x40000000;
    build_layer(&cost_b, &cost_e, b, t, 0, n, Axis::z);
    combine_nodes(&cost_b, &cost_e, 1, 0, cost_b, cost_e, cost_e);
    s_->root = next_node_ - 1;
    return s_->root;
}

}  // namespace

// ssd/bvh_builder.cc:50-158 — histograms of the three 10-bit digits, exclusive
// prefix sums, then three stable scatter passes.
void morton_radix_sort(const std::vector<int> &codes, std::vector<int> *sorted_codes, std::vector<int> *order) {
    const int n = static_cast<int>(codes.size());
    const int hs = 1024;
    std::vector<unsigned> hist(3 * hs, 0u);
    for (int i = 0; i < n; ++i) {
        const unsigned c = static_cast<unsigned>(codes[i]);
        ++hist[c & 0x3ff];
        ++hist[hs + ((c >> 10) & 0x3ff)];
        ++hist[2 * hs + (c >> 20)];
    }
    for (int k = 0; k < 3; ++k) {
        unsigned sum = 0;
        for (int i = 0; i < hs; ++i) {
            const unsigned tmp = hist[k * hs + i] + sum;
            hist[k * hs + i] = sum;
            sum = tmp;
        }
    }
    std::vector<unsigned> c0(n), c1(n);
    std::vector<int> p0(n), p1(n);
    for (int i = 0; i < n; ++i) {
        const unsigned c = static_cast<unsigned>(codes[i]);
        const unsigned pos = hist[c & 0x3ff]++;
        c0[pos] = c;
        p0[pos] = i;
    }
    for (int i = 0; i < n; ++i) {
        const unsigned pos = hist[hs + ((c0[i] >> 10) & 0x3ff)]++;
        c1[pos] = c0[i];
        p1[pos] = p0[i];
    }
    sorted_codes->resize(n);
    order->resize(n);
    for (int i = 0; i < n; ++i) {
        const unsigned pos = hist[2 * hs + (c1[i] >> 20)]++;
        (*sorted_codes)[pos] = static_cast<int>(c1[i]);
        (*order)[pos] = p1[i];
    }
}

void aac_morton(const Scene &s, std::vector<int> *codes, std::vector<int> *order) {
    std::vector<Bound> bounds;
    V3 lo, hi;
    std::vector<int> sorted;
    aac_codes(s, &bounds, &lo, &hi, codes);
    morton_radix_sort(*codes, &sorted, order);
}

int build_aac(Scene *scene, const AacParams &p) {
    if (scene->tris.empty()) {
        scene->nodes.assign(1, BvhNode());
        scene->indirect.clear();
        scene->root = -1;
        return -1;
    }
    AacBuilder builder(scene, p);
    return builder.build();
}

}  // namespace oracle
