// Linear BVH on gfx950: global AABB, Morton codes, stable LSD radix sort, Karras build,
// bottom-up refit and stack traversal with wave-aggregated pair emission.
//
// Semantics follow reference Solver/CollisionDetector/lbvh.cpp (file:line per kernel below);
// the data path is MI355X-first: the sort runs on the device (the reference sorts on the host
// and copies keys twice per tree and frame), refit uses one atomic ticket per inner node
// instead of a float-flag spin protocol, and pair lists are appended with one atomic per wave.
#include "lcsx_internal.h"
#include "lcsx_physics.h"

namespace lcsx
{

namespace
{
constexpr int kBlock     = 256;
constexpr int kSortItems = 16;  // keys per thread per sort pass
constexpr int kSortTile  = kBlock * kSortItems;

LCSX_HD uint32_t expand_bits(uint32_t b)
{
    b = (b | (b << 16)) & 0x030000FFu;
    b = (b | (b << 8)) & 0x0300F00Fu;
    b = (b | (b << 4)) & 0x030C30C3u;
    return (b | (b << 2)) & 0x09249249u;
}
// lbvh.cpp:108-121
LCSX_HD uint32_t morton32(f3 p)
{
    float x = clampf(p.x * 1024.0f, 0.0f, 1023.0f);
    float y = clampf(p.y * 1024.0f, 0.0f, 1023.0f);
    float z = clampf(p.z * 1024.0f, 0.0f, 1023.0f);
    return (expand_bits((uint32_t)x) << 2) | (expand_bits((uint32_t)y) << 1) | expand_bits((uint32_t)z);
}

// Leaf centres (lbvh.cpp:449-491) and the global box over leaves [0, covered): the reference's
// second reduction pass is dispatched over ceil(n_verts / 256) blocks of leaves (lbvh.cpp:958, 968).
__global__ __launch_bounds__(kBlock) void k_leaf_centers(uint32_t n, int arity, const uint32_t* __restrict__ prims,
                                                          const float* __restrict__ x, uint32_t covered,
                                                          float4* __restrict__ center, uint32_t* __restrict__ gbox)
{
    uint32_t l   = blockIdx.x * kBlock + threadIdx.x;
    f3       lo  = mk(3.0e38f, 3.0e38f, 3.0e38f), hi = mk(-3.0e38f, -3.0e38f, -3.0e38f);
    if (l < n)
    {
        f3 c, blo, bhi;
        if (arity == 3)
        {
            f3 p0 = ld3(x, prims[3 * l]), p1 = ld3(x, prims[3 * l + 1]), p2 = ld3(x, prims[3 * l + 2]);
            c   = 0.333333f * ((p0 + p1) + p2);
            blo = vmin(p0, vmin(p1, p2));
            bhi = vmax(p0, vmax(p1, p2));
        }
        else
        {
            f3 p0 = ld3(x, prims[2 * l]), p1 = ld3(x, prims[2 * l + 1]);
            c   = 0.5f * (p0 + p1);
            blo = vmin(p0, p1);
            bhi = vmax(p0, p1);
        }
        center[l] = make_float4(c.x, c.y, c.z, 0.0f);
        if (l < covered)
        {
            lo = blo;
            hi = bhi;
        }
    }
    float v[6] = {lo.x, lo.y, lo.z, hi.x, hi.y, hi.z};
    for (int k = 0; k < 3; k++) v[k] = wave_min(v[k]);
    for (int k = 3; k < 6; k++) v[k] = wave_max(v[k]);
    __shared__ float s[4][6];
    int              w = threadIdx.x >> 6;
    if (lane_id() == 0)
        for (int k = 0; k < 6; k++) s[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 6)
    {
        int   k = threadIdx.x;
        float r = s[0][k];
        for (int ww = 1; ww < 4; ww++) r = k < 3 ? smin(r, s[ww][k]) : smax(r, s[ww][k]);
        if (blockIdx.x * kBlock < covered)
        {
            if (k < 3) atomicMin(&gbox[k], f2ord(r));
            else atomicMax(&gbox[k], f2ord(r));
        }
    }
}

__global__ void k_init_gbox(uint32_t* gbox)
{
    int k = threadIdx.x;
    if (k < 3) gbox[k] = 0xffffffffu;
    else if (k < 6) gbox[k] = 0u;
}

// fn_compute_mortons, lbvh.cpp:507-523: key = morton32 of the normalised centre, value = leaf id.
__global__ __launch_bounds__(kBlock) void k_morton(uint32_t n, const float4* __restrict__ center,
                                                    const uint32_t* __restrict__ gbox, uint32_t* __restrict__ keys,
                                                    uint32_t* __restrict__ vals)
{
    uint32_t l = blockIdx.x * kBlock + threadIdx.x;
    if (l >= n) return;
    f3     lo  = mk(ord2f(gbox[0]), ord2f(gbox[1]), ord2f(gbox[2]));
    f3     hi  = mk(ord2f(gbox[3]), ord2f(gbox[4]), ord2f(gbox[5]));
    f3     ext = vmax(hi - lo, mk(1e-6f, 1e-6f, 1e-6f));
    f3     inv = mk(1.0f / ext.x, 1.0f / ext.y, 1.0f / ext.z);
    float4 c4  = center[l];
    f3     c   = mk(c4.x, c4.y, c4.z) - lo;
    keys[l]    = morton32(mk(c.x * inv.x, c.y * inv.y, c.z * inv.z));
    vals[l]    = l;
}

// ---------------------------------------------------------------- stable LSD radix sort
__global__ __launch_bounds__(kBlock) void k_radix_hist(const uint32_t* __restrict__ keys, uint32_t n, int shift,
                                                        uint32_t* __restrict__ hist, uint32_t nblocks)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * kSortTile;
    for (int i = 0; i < kSortItems; i++)
    {
        uint32_t idx = base + i * kBlock + threadIdx.x;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of `total` counters in one workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void k_scan_single(uint32_t* __restrict__ a, uint32_t total)
{
    __shared__ uint32_t part[1024];
    uint32_t            t     = threadIdx.x;
    uint32_t            chunk = (total + 1023) / 1024;
    uint32_t            b = t * chunk, e = min(total, b + chunk);
    uint32_t            sum = 0;
    for (uint32_t i = b; i < e; i++) sum += a[i];
    part[t] = sum;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1)
    {
        uint32_t v = t >= o ? part[t - o] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    for (uint32_t i = b; i < e; i++)
    {
        uint32_t v = a[i];
        a[i]       = run;
        run += v;
    }
}

__global__ __launch_bounds__(kBlock) void k_radix_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                           uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                           uint32_t n, int shift, const uint32_t* __restrict__ hist,
                                                           uint32_t nblocks)
{
    __shared__ uint32_t s_base[256];
    __shared__ uint32_t s_cnt[4][256];
    const uint32_t      t = threadIdx.x, w = t >> 6, lane = t & 63;
    s_base[t]              = hist[t * nblocks + blockIdx.x];
    const uint32_t base    = blockIdx.x * kSortTile;
    for (int i = 0; i < kSortItems; i++)
    {
        for (int k = 0; k < 4; k++) s_cnt[w][lane + 64 * k] = 0;
        __syncthreads();
        uint32_t idx   = base + i * kBlock + t;
        bool     valid = idx < n;
        uint32_t key   = valid ? kin[idx] : 0u;
        uint32_t d     = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < 8; b++)
        {
            bool     bit = (d >> b) & 1u;
            uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        uint32_t rank = (uint32_t)__popcll(peers & lanemask_lt());
        if (valid && rank == 0) s_cnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid)
        {
            uint32_t off = s_base[d] + rank;
            for (uint32_t ww = 0; ww < w; ww++) off += s_cnt[ww][d];
            kout[off] = key;
            vout[off] = vin[idx];
        }
        __syncthreads();
        s_base[t] += s_cnt[0][t] + s_cnt[1][t] + s_cnt[2][t] + s_cnt[3][t];
        __syncthreads();
    }
}

// fn_apply_sorted, lbvh.cpp:525-539 (+ the 64-bit keys the build reads)
__global__ __launch_bounds__(kBlock) void k_leaves(uint32_t n, const uint32_t* __restrict__ skeys,
                                                    const uint32_t* __restrict__ svals, uint64_t* __restrict__ morton_sorted,
                                                    uint2* __restrict__ children, uint32_t* __restrict__ object_idx,
                                                    uint32_t* __restrict__ parent)
{
    uint32_t l = blockIdx.x * kBlock + threadIdx.x;
    if (l >= n) return;
    uint32_t o       = svals[l];
    morton_sorted[l] = ((uint64_t)skeys[l] << 32) | (uint64_t)o;
    uint32_t ni      = n - 1;
    children[ni + l] = make_uint2(o, o);
    object_idx[ni + l] = o;
    if (l == 0) parent[0] = kInvalid;
}

LCSX_D int clz64(uint64_t v) { return v == 0 ? 64 : __clzll((long long)v); }

// fn_build_inner_nodes with determineRange / findSplit, lbvh.cpp:204-291, 541-596
__global__ __launch_bounds__(kBlock) void k_karras(uint32_t n, const uint64_t* __restrict__ m,
                                                    uint32_t* __restrict__ parent, uint2* __restrict__ children,
                                                    uint32_t* __restrict__ object_idx)
{
    int nid = blockIdx.x * kBlock + threadIdx.x;
    int ni  = (int)n - 1;
    if (nid >= ni) return;
    int  N  = (int)n;
    auto cp = [&](uint64_t mi, int j) -> int { return (j >= 0 && j < N) ? clz64(mi ^ m[j]) : -1; };
    int  i = 0, j = ni;
    if (nid != 0)
    {
        i            = nid;
        uint64_t mi  = m[i];
        int      cl  = clz64(mi ^ m[i - 1]);
        int      cr  = clz64(mi ^ m[i + 1]);
        int      d   = cl < cr ? 1 : -1;
        int      cpm = cl < cr ? cl : cr;
        int      lmax = 2;
        while (cp(mi, i + lmax * d) > cpm) lmax <<= 1;
        int l = 0;
        for (int t = lmax >> 1; t >= 1; t >>= 1)
            if (cp(mi, i + (l + t) * d) > cpm) l += t;
        j = i + l * d;
    }
    int      a = i < j ? i : j, b = i < j ? j : i;
    uint64_t ma = m[a], mb = m[b];
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
            t      = (t + 1) >> 1;
            int ns = split + t;
            if (ns < b && clz64(ma ^ m[ns]) > cpn) split = ns;
            if (!(t > 1)) break;
        }
    }
    int cl_ = (a == split) ? ni + split : split;
    int cr_ = (b == split + 1) ? ni + split + 1 : split + 1;
    if (cr_ >= ni)
    {
        int tmp = cl_;
        cl_     = cr_;
        cr_     = tmp;
    }
    parent[cl_]     = (uint32_t)nid;
    parent[cr_]     = (uint32_t)nid;
    children[nid]   = make_uint2((uint32_t)cl_, (uint32_t)cr_);
    object_idx[nid] = kInvalid;
}

// update_*_leave_aabb (lbvh.cpp:636-683) + refit (lbvh.cpp:702-754): the second child to arrive
// at an inner node (atomic ticket) computes its box; fences publish boxes across XCD L2s.
__global__ __launch_bounds__(kBlock) void k_refit(uint32_t n, int arity, const uint32_t* __restrict__ prims,
                                                   const uint32_t* __restrict__ sorted_to_orig,
                                                   const float* __restrict__ x0, const float* __restrict__ x1,
                                                   const float* __restrict__ pad, const uint32_t* __restrict__ parent,
                                                   const uint2* __restrict__ children, NodeBox* nodes,
                                                   float4* __restrict__ packed, uint32_t* counters)
{
    uint32_t l = blockIdx.x * kBlock + threadIdx.x;
    if (l >= n) return;
    uint32_t o = sorted_to_orig[l];
    f3       lo, hi;
    float    th;
    if (arity == 3)
    {
        uint32_t a = prims[3 * o], b = prims[3 * o + 1], c = prims[3 * o + 2];
        f3       s0 = ld3(x0, a), s1 = ld3(x0, b), s2 = ld3(x0, c);
        f3       e0 = ld3(x1, a), e1 = ld3(x1, b), e2 = ld3(x1, c);
        lo = vmin(vmin(s0, vmin(s1, s2)), vmin(e0, vmin(e1, e2)));
        hi = vmax(vmax(s0, vmax(s1, s2)), vmax(e0, vmax(e1, e2)));
        th = pad[a];
    }
    else
    {
        uint32_t a = prims[2 * o], b = prims[2 * o + 1];
        f3       s0 = ld3(x0, a), s1 = ld3(x0, b), e0 = ld3(x1, a), e1 = ld3(x1, b);
        lo = vmin(vmin(s0, s1), vmin(e0, e1));
        hi = vmax(vmax(s0, s1), vmax(e0, e1));
        th = pad[a];
    }
    lo              = lo - mk(th, th, th);
    hi              = hi + mk(th, th, th);
    uint32_t ni     = n - 1;
    uint32_t cur    = ni + l;
    NodeBox  box    = {make_float4(lo.x, lo.y, lo.z, 0.0f), make_float4(hi.x, hi.y, hi.z, 0.0f)};
    nodes[cur]      = box;
    // a tree of n leaves has depth < n; the bound only guards against a corrupted parent array
    for (uint32_t depth = 0; depth < n; depth++)
    {
        uint32_t p = parent[cur];
        if (p == kInvalid || p >= ni) break;
        __threadfence();
        if (atomicAdd(&counters[p], 1u) == 0u) break;
        __threadfence();
        uint2   c = children[p];
        NodeBox A = nodes[c.x], B = nodes[c.y];
        NodeBox U;
        U.lo     = make_float4(smin(A.lo.x, B.lo.x), smin(A.lo.y, B.lo.y), smin(A.lo.z, B.lo.z), 0.0f);
        U.hi     = make_float4(smax(A.hi.x, B.hi.x), smax(A.hi.y, B.hi.y), smax(A.hi.z, B.hi.z), 0.0f);
        nodes[p] = U;
        // child codes: leaves keep (element, element) in children[]
        uint32_t ca = c.x >= ni ? (0x80000000u | children[c.x].x) : c.x;
        uint32_t cb = c.y >= ni ? (0x80000000u | children[c.y].x) : c.y;
        float4*  P  = packed + 4 * (size_t)p;
        P[0] = make_float4(A.lo.x, A.lo.y, A.lo.z, __uint_as_float(ca));
        P[1] = make_float4(A.hi.x, A.hi.y, A.hi.z, 0.0f);
        P[2] = make_float4(B.lo.x, B.lo.y, B.lo.z, __uint_as_float(cb));
        P[3] = make_float4(B.hi.x, B.hi.y, B.hi.z, 0.0f);
        cur  = p;
    }
}

__global__ void k_fill_u32(uint32_t* a, uint32_t n, uint32_t v)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) a[i] = v;
}

// ---------------------------------------------------------------- traversal
constexpr int kQueryBlock = 128;
constexpr int kStack      = 32;

// query_template2, lbvh.cpp:293-360: depth-first with a 32-entry stack staged in LDS; a full
// stack ends the query (the reference's silent truncation), reported through *overflow.
LCSX_D bool overlap4(float4 lo4, float4 hi4, f3 lo, f3 hi)
{
    return (lo4.x < hi.x && lo4.y < hi.y && lo4.z < hi.z) && (hi4.x > lo.x && hi4.y > lo.y && hi4.z > lo.z);
}

// query_template2, lbvh.cpp:293-360: depth-first with a 32-entry stack staged in LDS; a full
// stack ends the query (the reference's silent truncation), reported through *overflow. Each
// step loads one packed inner node (both child boxes and codes); children are tested and inner
// hits pushed in the reference's order, so pairs and overflow behave as the reference's.
__global__ __launch_bounds__(kQueryBlock) void k_query(QueryArgs a, const float4* __restrict__ packed, uint32_t n_inner)
{
    __shared__ uint32_t stack[kStack][kQueryBlock];
    uint32_t            q = blockIdx.x * kQueryBlock + threadIdx.x;
    if (q >= a.n_queries) return;
    f3   lo, hi;
    bool filt = false;
    if (a.mode == QueryVerts)
    {
        f3    p0 = ld3(a.x0, q), p1 = ld3(a.x1, q);
        float t  = a.offset[q] + a.dhat[q];
        lo       = vmin(p0, p1) - mk(t, t, t);
        hi       = vmax(p0, p1) + mk(t, t, t);
    }
    else if (a.mode == QueryEdges)
    {
        uint32_t e0 = a.edges[2 * q], e1 = a.edges[2 * q + 1];
        f3       s0 = ld3(a.x0, e0), s1 = ld3(a.x0, e1), t0 = ld3(a.x1, e0), t1 = ld3(a.x1, e1);
        float    t  = a.offset[e0] + a.dhat[e0];
        lo          = vmin(vmin(s0, s1), vmin(t0, t1)) - mk(t, t, t);
        hi          = vmax(vmax(s0, s1), vmax(t0, t1)) + mk(t, t, t);
        filt        = true;
    }
    else
    {
        lo   = ld3(a.boxes, 2 * q);
        hi   = ld3(a.boxes, 2 * q + 1);
        filt = a.edge_filter != 0;
    }
    const uint32_t tid = threadIdx.x;
    int            sp  = 0;
    stack[sp++][tid]   = 0u;
    // each inner node is visited at most once per query in a well-formed tree; the cap (as the
    // reference's 100000-iteration guard) only bounds a corrupted one
    uint32_t visits = 0;
    while (sp > 0)
    {
        if (++visits > n_inner + 1u)
        {
            if (a.overflow) atomicOr(a.overflow, 2u);
            break;
        }
        uint32_t      nd = stack[--sp][tid];
        const float4* P  = packed + 4 * (size_t)nd;
        float4        r0 = P[0], r1 = P[1], r2 = P[2], r3 = P[3];
        bool          brk = false;
#pragma unroll
        for (int ii = 0; ii < 2; ii++)
        {
            float4   blo  = ii == 0 ? r0 : r2;
            float4   bhi  = ii == 0 ? r1 : r3;
            uint32_t code = __float_as_uint(blo.w);
            bool     hit  = overlap4(blo, bhi, lo, hi);
            bool     leaf = (code & 0x80000000u) != 0u;
            uint32_t obj  = code & 0x7fffffffu;
            // a full stack at the first child ends the query before the second is tested
            hit           = hit && !brk;
            bool     emit = hit && leaf && (!filt || q < obj);
            uint32_t slot = wave_append(a.count, emit);
            if (emit && slot < a.capacity)
            {
                a.pairs[2 * (size_t)slot]     = q;
                a.pairs[2 * (size_t)slot + 1] = obj;
            }
            if (hit && !leaf)
            {
                if (sp < kStack)
                    stack[sp++][tid] = code;
                else
                {
                    brk = true;
                    if (a.overflow) atomicOr(a.overflow, 1u);
                }
            }
        }
        if (brk) break;
    }
}
}  // namespace

void Lbvh::init(int arity_, const uint32_t* device_prims, uint32_t n_leaves)
{
    arity = arity_;
    n     = n_leaves;
    prims = device_prims;
    if (n == 0) return;
    keys[0].alloc(n); keys[1].alloc(n); vals[0].alloc(n); vals[1].alloc(n);
    uint32_t nb = (n + kSortTile - 1) / kSortTile;
    hist.alloc(256 * (size_t)nb);
    center.alloc(n);
    morton_sorted.alloc(n);
    parent.alloc(n_nodes());
    object_idx.alloc(n_nodes());
    children.alloc(n_nodes());
    nodes.alloc(n_nodes());
    packed.alloc(4 * (size_t)(n_inner() ? n_inner() : 1));
    counters.alloc(n_inner() ? n_inner() : 1);
    gbox.alloc(6);
    sorted_to_orig = vals[0].p;
}

void Lbvh::construct(const float* x, uint32_t n_verts_for_reduce, hipStream_t s)
{
    if (n == 0) return;
    uint64_t covered = std::min<uint64_t>(n, ((uint64_t)n_verts_for_reduce + 255u) / 256u * 256u);
    k_init_gbox<<<1, 64, 0, s>>>(gbox.p);
    k_leaf_centers<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, arity, prims, x, (uint32_t)covered, center.p, gbox.p);
    k_morton<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, center.p, gbox.p, keys[0].p, vals[0].p);
    uint32_t nb  = (n + kSortTile - 1) / kSortTile;
    int      cur = 0;
    for (int shift = 0; shift < 32; shift += 8)
    {
        k_radix_hist<<<nb, kBlock, 0, s>>>(keys[cur].p, n, shift, hist.p, nb);
        k_scan_single<<<1, 1024, 0, s>>>(hist.p, 256 * nb);
        k_radix_scatter<<<nb, kBlock, 0, s>>>(keys[cur].p, vals[cur].p, keys[1 - cur].p, vals[1 - cur].p, n, shift,
                                               hist.p, nb);
        cur = 1 - cur;
    }
    sorted_to_orig = vals[cur].p;
    k_leaves<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, keys[cur].p, vals[cur].p, morton_sorted.p, children.p,
                                                    object_idx.p, parent.p);
    if (n > 1) k_karras<<<grid_for(n - 1, kBlock), kBlock, 0, s>>>(n, morton_sorted.p, parent.p, children.p, object_idx.p);
    LCSX_HIP(hipGetLastError());
}

void Lbvh::refit(const float* x0, const float* x1, const float* pad, hipStream_t s)
{
    if (n == 0) return;
    k_fill_u32<<<grid_for((uint32_t)counters.n, kBlock), kBlock, 0, s>>>(counters.p, (uint32_t)counters.n, 0u);
    k_refit<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, arity, prims, sorted_to_orig, x0, x1, pad, parent.p, children.p,
                                                   nodes.p, packed.p, counters.p);
    LCSX_HIP(hipGetLastError());
}

void lbvh_query(const Lbvh& t, const QueryArgs& a, hipStream_t s)
{
    if (t.n == 0 || a.n_queries == 0) return;
    if (t.n < 2) return;  // a single leaf has no inner node to start from (the solver never queries one)
    k_query<<<grid_for(a.n_queries, kQueryBlock), kQueryBlock, 0, s>>>(a, t.packed.p, t.n_inner());
    LCSX_HIP(hipGetLastError());
}

}  // namespace lcsx
