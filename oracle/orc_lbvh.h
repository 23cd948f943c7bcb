// Oracle: linear BVH over faces or edges (Morton codes, sort, Karras build, refit, query).
//
// TEST INFRASTRUCTURE ONLY (see orc_math.h). Restates, on the CPU:
//   Solver/CollisionDetector/lbvh.cpp:84-135   expand_bits / make_morton32 / make_morton64
//   Solver/CollisionDetector/lbvh.cpp:155-291  clz, determineRange, findSplit
//   Solver/CollisionDetector/lbvh.cpp:293-360  query_template2 (32-entry stack, overflow ends the query)
//   Solver/CollisionDetector/lbvh.cpp:449-596  leaf centres, global AABB, mortons, apply_sorted, inner nodes
//   Solver/CollisionDetector/lbvh.cpp:617-754  leaf AABB update and bottom-up refit
//   Solver/CollisionDetector/lbvh.cpp:954-998  reduce_*_tree_aabb sizing and construct_tree (host sort)
//   Solver/CollisionDetector/aabb.h:68-100     add_thickness / add_aabb / strict overlap
#pragma once
#include <vector>
#include "orc_math.h"

namespace orc
{

enum TreeType
{
    TreeFace = 0,
    TreeEdge = 1
};

inline uint32_t expand_bits(uint32_t b)
{
    b = (b | (b << 16)) & 0x030000FFu;
    b = (b | (b << 8)) & 0x0300F00Fu;
    b = (b | (b << 4)) & 0x030C30C3u;
    return (b | (b << 2)) & 0x09249249u;
}

// lbvh.cpp:108-121: each axis is clamp(norm * 1024, 0, 1023) truncated to an integer.
template <class T>
inline uint32_t morton32(V3<T> p)
{
    const T range = T(1024), hi = T(1023), lo = T(0);
    T x = clampv<T>(p.x * range, lo, hi);
    T y = clampv<T>(p.y * range, lo, hi);
    T z = clampv<T>(p.z * range, lo, hi);
    uint32_t xx = expand_bits((uint32_t)x), yy = expand_bits((uint32_t)y), zz = expand_bits((uint32_t)z);
    return (xx << 2) | (yy << 1) | zz;
}
inline uint64_t morton64(uint32_t m32, uint32_t id) { return ((uint64_t)m32 << 32) | (uint64_t)id; }
inline int clz64(uint64_t x) { return x == 0 ? 64 : __builtin_clzll(x); }

template <class T>
struct AABB3
{
    V3<T> lo, hi;
};
template <class T> inline AABB3<T> aabb_union(const AABB3<T>& a, const AABB3<T>& b) { return {vmin(a.lo, b.lo), vmax(a.hi, b.hi)}; }
template <class T> inline AABB3<T> aabb_pad(const AABB3<T>& a, T t) { return {a.lo - V3<T>(t, t, t), a.hi + V3<T>(t, t, t)}; }
// aabb.h:96-100: strict on every axis
template <class T> inline bool aabb_overlap(const AABB3<T>& l, const AABB3<T>& r)
{
    return (l.lo.x < r.hi.x && l.lo.y < r.hi.y && l.lo.z < r.hi.z) && (l.hi.x > r.lo.x && l.hi.y > r.lo.y && l.hi.z > r.lo.z);
}

template <class T>
struct Lbvh
{
    int                   type      = TreeFace;
    uint32_t              n_leaves  = 0;
    std::vector<uint32_t> prims;          // 3 (face) or 2 (edge) vertex ids per leaf
    std::vector<uint64_t> morton;         // unsorted, per original leaf id
    std::vector<uint64_t> morton_sorted;  // sorted keys
    std::vector<uint32_t> sorted_to_orig; // sa_sorted_get_original
    std::vector<uint32_t> parent;         // 2n-1, parent[0] = 0xffffffff
    std::vector<uint32_t> children;       // 2 * (2n-1)
    std::vector<uint32_t> object_idx;     // 2n-1, inner = 0xffffffff
    std::vector<T>        node_aabb;      // 6 * (2n-1): lo.xyz, hi.xyz
    V3<T>                 global_lo, global_hi;
    bool                  healthy = true;

    int arity() const { return type == TreeFace ? 3 : 2; }

    void init(int tree_type, const uint32_t* p, uint32_t n)
    {
        type     = tree_type;
        n_leaves = n;
        prims.assign(p, p + (size_t)n * arity());
        morton.assign(n, 0);
        morton_sorted.assign(n, 0);
        sorted_to_orig.assign(n, 0);
        parent.assign(2 * (size_t)n - 1, 0xffffffffu);
        children.assign(2 * (2 * (size_t)n - 1), 0u);
        object_idx.assign(2 * (size_t)n - 1, 0xffffffffu);
        node_aabb.assign(6 * (2 * (size_t)n - 1), T(0));
    }

    // lbvh.cpp:954-998: leaf centres and a global AABB whose second pass covers only the
    // first ceil(n_verts / 256) blocks of 256 leaves, then mortons, sort, Karras build.
    void construct(const V3<T>* x, uint32_t n_verts)
    {
        const int                 k = arity();
        std::vector<V3<T>>        center(n_leaves);
        const uint64_t            n_blocks_pass2 = (n_verts + 255u) / 256u;
        const uint64_t            covered        = std::min<uint64_t>(n_leaves, n_blocks_pass2 * 256u);
        V3<T>                     lo(T(1e30), T(1e30), T(1e30)), hi(T(-1e30), T(-1e30), T(-1e30));
        for (uint32_t l = 0; l < n_leaves; l++)
        {
            const uint32_t* v = &prims[(size_t)l * k];
            AABB3<T>        b;
            if (k == 3)
            {
                V3<T> p0 = x[v[0]], p1 = x[v[1]], p2 = x[v[2]];
                center[l] = T(0.333333f) * ((p0 + p1) + p2);
                b         = {vmin(p0, vmin(p1, p2)), vmax(p0, vmax(p1, p2))};
            }
            else
            {
                V3<T> p0 = x[v[0]], p1 = x[v[1]];
                center[l] = T(0.5f) * (p0 + p1);
                b         = {vmin(p0, p1), vmax(p0, p1)};
            }
            if (l < covered)
            {
                lo = vmin(lo, b.lo);
                hi = vmax(hi, b.hi);
            }
        }
        global_lo = lo;
        global_hi = hi;
        V3<T> ext = vmax(hi - lo, V3<T>(T(1e-6f), T(1e-6f), T(1e-6f)));
        V3<T> inv(T(1) / ext.x, T(1) / ext.y, T(1) / ext.z);
        for (uint32_t l = 0; l < n_leaves; l++)
        {
            V3<T> c = center[l] - lo;
            V3<T> nrm(c.x * inv.x, c.y * inv.y, c.z * inv.z);
            morton[l] = morton64(morton32(nrm), l);
        }
        for (uint32_t l = 0; l < n_leaves; l++) sorted_to_orig[l] = l;
        std::sort(sorted_to_orig.begin(), sorted_to_orig.end(), [&](uint32_t a, uint32_t b) { return morton[a] < morton[b]; });
        const uint32_t n_inner = n_leaves - 1;
        std::fill(object_idx.begin(), object_idx.end(), 0xffffffffu);
        parent[0] = 0xffffffffu;
        for (uint32_t l = 0; l < n_leaves; l++)
        {
            uint32_t o          = sorted_to_orig[l];
            morton_sorted[l]    = morton[o];
            children[2 * (n_inner + l) + 0] = o;
            children[2 * (n_inner + l) + 1] = o;
            object_idx[n_inner + l]         = o;
        }
        const int n = (int)n_leaves;
        auto      cp = [&](uint64_t mi, int j) -> int { return (j >= 0 && j < n) ? clz64(mi ^ morton_sorted[j]) : -1; };
        for (int nid = 0; nid < (int)n_inner; nid++)
        {
            int i = 0, j = (int)n_inner;
            if (nid != 0)
            {
                // determineRange, lbvh.cpp:204-244
                i              = nid;
                uint64_t mi    = morton_sorted[i];
                int      cl    = clz64(mi ^ morton_sorted[i - 1]);
                int      cr    = clz64(mi ^ morton_sorted[i + 1]);
                int      d     = cl < cr ? 1 : -1;
                int      cpmin = std::min(cl, cr);
                int      lmax  = 2;
                while (cp(mi, i + lmax * d) > cpmin) lmax <<= 1;
                int l = 0;
                for (int t = lmax >> 1; t >= 1; t >>= 1)
                    if (cp(mi, i + (l + t) * d) > cpmin) l += t;
                j = i + l * d;
            }
            // findSplit, lbvh.cpp:246-291
            int a = i, b = j;
            if (a > b) std::swap(a, b);
            uint64_t ma = morton_sorted[a], mb = morton_sorted[b];
            int      split;
            if (ma == mb)
                split = (a + b) >> 1;
            else
            {
                int cpn = clz64(ma ^ mb);
                int t   = b - a;
                split   = a;
                while (true)
                {
                    t          = (t + 1) >> 1;
                    int ns     = split + t;
                    if (ns < b && clz64(ma ^ morton_sorted[ns]) > cpn) split = ns;
                    if (!(t > 1)) break;
                }
            }
            // lbvh.cpp:575-595
            int cl_ = (std::min(i, j) == split) ? (int)n_inner + split : split;
            int cr_ = (std::max(i, j) == split + 1) ? (int)n_inner + split + 1 : split + 1;
            if (cr_ >= (int)n_inner) std::swap(cl_, cr_);
            parent[cl_]              = (uint32_t)nid;
            parent[cr_]              = (uint32_t)nid;
            children[2 * nid + 0]    = (uint32_t)cl_;
            children[2 * nid + 1]    = (uint32_t)cr_;
        }
        // fn_check_construction, lbvh.cpp:598-611
        healthy = true;
        for (uint32_t nid = 0; nid < n_inner; nid++)
            if (parent[children[2 * nid]] != nid || parent[children[2 * nid + 1]] != nid) healthy = false;
    }

    AABB3<T> node(uint32_t i) const
    {
        const T* a = &node_aabb[6 * (size_t)i];
        return {V3<T>(a[0], a[1], a[2]), V3<T>(a[3], a[4], a[5])};
    }
    void set_node(uint32_t i, const AABB3<T>& b)
    {
        T* a = &node_aabb[6 * (size_t)i];
        a[0] = b.lo.x; a[1] = b.lo.y; a[2] = b.lo.z; a[3] = b.hi.x; a[4] = b.hi.y; a[5] = b.hi.z;
    }

    // lbvh.cpp:636-683 leaves (swept over start/end, padded by the first vertex's offset),
    // lbvh.cpp:702-754 inner nodes = union of children (order-free, so done top-down here).
    void refit(const V3<T>* x0, const V3<T>* x1, const T* thickness)
    {
        const int      k       = arity();
        const uint32_t n_inner = n_leaves - 1;
        for (uint32_t l = 0; l < n_leaves; l++)
        {
            uint32_t        o = sorted_to_orig[l];
            const uint32_t* v = &prims[(size_t)o * k];
            AABB3<T>        b;
            if (k == 3)
            {
                AABB3<T> s = {vmin(x0[v[0]], vmin(x0[v[1]], x0[v[2]])), vmax(x0[v[0]], vmax(x0[v[1]], x0[v[2]]))};
                AABB3<T> e = {vmin(x1[v[0]], vmin(x1[v[1]], x1[v[2]])), vmax(x1[v[0]], vmax(x1[v[1]], x1[v[2]]))};
                b          = aabb_union(s, e);
            }
            else
            {
                b = {vmin(vmin(x0[v[0]], x0[v[1]]), vmin(x1[v[0]], x1[v[1]])),
                     vmax(vmax(x0[v[0]], x0[v[1]]), vmax(x1[v[0]], x1[v[1]]))};
            }
            set_node(n_inner + l, aabb_pad(b, thickness[v[0]]));
        }
        // children of an inner node always have larger depth; a post-order walk fills parents after children
        std::vector<uint32_t> order;
        order.reserve(n_inner);
        std::vector<uint32_t> stack;
        if (n_inner > 0) stack.push_back(0);
        while (!stack.empty())
        {
            if (order.size() > n_inner) return;  // not a tree (construct() was not run)
            uint32_t nd = stack.back();
            stack.pop_back();
            order.push_back(nd);
            for (int c = 0; c < 2; c++)
            {
                uint32_t ch = children[2 * nd + c];
                if (ch < n_inner) stack.push_back(ch);
            }
        }
        for (size_t i = order.size(); i-- > 0;)
        {
            uint32_t nd = order[i];
            set_node(nd, aabb_union(node(children[2 * nd]), node(children[2 * nd + 1])));
        }
    }

    // query_template2, lbvh.cpp:293-360. Returns false if a 32-entry stack overflow ended
    // the traversal early (the pairs found up to that point are kept, as in the reference).
    template <class Emit>
    bool query(const AABB3<T>& box, uint32_t self, bool edge_filter, Emit&& emit) const
    {
        uint32_t stack[32];
        int      sp  = 0;
        stack[sp++]  = 0;
        bool ok      = true;
        while (sp > 0)
        {
            uint32_t nd = stack[--sp];
            bool     brk = false;
            for (int c = 0; c < 2; c++)
            {
                uint32_t ch = children[2 * nd + c];
                if (aabb_overlap(node(ch), box))
                {
                    uint32_t obj = object_idx[ch];
                    if (obj != 0xffffffffu)
                    {
                        if (!edge_filter || self < obj) emit(self, obj);
                    }
                    else
                    {
                        if (sp < 32)
                            stack[sp++] = ch;
                        else
                        {
                            brk = true;
                            ok  = false;
                            break;
                        }
                    }
                }
            }
            if (brk) break;
        }
        return ok;
    }
};

}  // namespace orc
