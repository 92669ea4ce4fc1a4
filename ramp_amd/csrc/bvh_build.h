/* bvh_build.h - linear BVH over the triangles of a GMesh, built on the GPU.
 *
 * Method (Karras, "Maximizing parallelism in the construction of BVHs,
 * octrees and k-d trees", HPG 2012): 30-bit Morton codes of triangle
 * centroids in the mesh's bounding box, radix sort of (code, index) pairs
 * (hipcub), one thread per internal node determines its key range and split
 * from longest common prefixes (equal codes ordered by sorted position),
 * then leaf boxes and subtree sizes propagate to the root bottom-up with one
 * atomic flag per internal node.  In that build layout internal nodes are
 * 0..n-2, leaf k is node n-1+k and refers to sorted position k; order[k] is
 * the triangle index.
 *
 * A last pass writes the tree in depth-first preorder as 32-byte BvhFlat
 * nodes: the first-visited child of an internal node is the next node, and
 * `esc` is the node after the subtree (the layout's start + preorder index
 * + subtree size), so traversal needs neither child pointers nor a stack.
 * The same pass writes the tree of the host SAH builder (bvh_sah.h), whose
 * leaves may hold two triangles; the leaves' triangles are copied in leaf
 * order into BvhTri records (k_bvh_leaf_tris here, on the host there).  A node's preorder
 * index is its depth plus, over its ancestors whose second-visited subtree
 * holds it, the sizes of their first-visited subtrees.  Six such layouts
 * are written, one per ray direction class (+x, -x, +y, -y, +z, -z): in
 * layout 2a + s the first-visited child is the one whose box centre lies
 * nearer along axis a for a ray running in that direction, so a ray
 * traversed in the layout of its dominant direction meets nearer triangles
 * first and its nearest-hit bounds prune the farther subtrees early.
 *
 * Every leaf box is the triangle's box widened by a margin of 1e-3 of the
 * mesh extent (and at least 1e-6 m), far above the rounding of both the
 * slab test and the Moller-Trumbore test, so traversal never skips a
 * triangle the reference's brute-force loop would accept.
 */
#ifndef RAMP_BVH_BUILD_H
#define RAMP_BVH_BUILD_H

#include <hipcub/hipcub.hpp>

__device__ static inline uint32_t bvh_expand10(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__global__ void k_bvh_morton(const float *pts, uint32_t n, float3 lo, float3 inv_ext, uint32_t *keys,
                             uint32_t *vals)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float *P = pts + 9 * (size_t)t;
    float c[3];
    float l[3] = {lo.x, lo.y, lo.z}, ie[3] = {inv_ext.x, inv_ext.y, inv_ext.z};
    uint32_t q[3];
    for (int a = 0; a < 3; a++) {
        c[a] = (P[a] + P[3 + a] + P[6 + a]) * (1.0f / 3.0f);
        float u = (c[a] - l[a]) * ie[a];
        u = fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f);
        q[a] = (uint32_t)u;
    }
    keys[t] = (bvh_expand10(q[0]) << 2) | (bvh_expand10(q[1]) << 1) | bvh_expand10(q[2]);
    vals[t] = t;
}

__device__ static inline int bvh_delta(const uint32_t *codes, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    uint32_t a = codes[i], b = codes[j];
    if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clz(a ^ b);
}

__global__ void k_bvh_tree(const uint32_t *codes, int n, BvhNode *nodes, int *parent)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (bvh_delta(codes, n, i, i + 1) - bvh_delta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = bvh_delta(codes, n, i, i - d);
    int lmax = 2;
    while (bvh_delta(codes, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (bvh_delta(codes, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = bvh_delta(codes, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (bvh_delta(codes, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + (d < 0 ? -1 : 0);
    int lo = i < j ? i : j, hi = i < j ? j : i;
    int left = (lo == gamma) ? (n - 1 + gamma) : gamma;
    int right = (hi == gamma + 1) ? (n - 1 + gamma + 1) : (gamma + 1);
    nodes[i].left = left;
    nodes[i].right = right;
    nodes[i].count = 0;
    parent[left] = i;
    parent[right] = i;
}

__global__ void k_bvh_boxes(const float *pts, const uint32_t *order, int n, float margin,
                            BvhNode *nodes, const int *parent, int *flags)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float *P = pts + 9 * (size_t)order[k];
    int node = n - 1 + k;
    BvhNode &lf = nodes[node];
    for (int a = 0; a < 3; a++) {
        lf.lo[a] = fminf(P[a], fminf(P[3 + a], P[6 + a])) - margin;
        lf.hi[a] = fmaxf(P[a], fmaxf(P[3 + a], P[6 + a])) + margin;
    }
    lf.left = -1;
    lf.right = k;
    lf.count = 1;               /* subtree size */
    lf.ntri = 1;
    if (n == 1) return;
    __threadfence();
    int p = parent[node];
    /* a tree of n leaves is at most 30 + log2(n) + 1 deep for 30-bit keys
     * with index tie-breaks; the bound only guards against a malformed tree */
    for (int guard = 0; p >= 0 && guard < 96; guard++) {
        if (atomicAdd(&flags[p], 1) == 0) return; /* the sibling finishes this node */
        __threadfence();
        /* the sibling's box may have been written by another CU: read the
         * children through agent-scope atomic loads, not the CU cache */
        int cl = __hip_atomic_load(&nodes[p].left, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int cr = __hip_atomic_load(&nodes[p].right, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int sl = __hip_atomic_load(&nodes[cl].count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int sr = __hip_atomic_load(&nodes[cr].count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        nodes[p].count = sl + sr + 1;
        for (int ax = 0; ax < 3; ax++) {
            float alo = __hip_atomic_load(&nodes[cl].lo[ax], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            float blo = __hip_atomic_load(&nodes[cr].lo[ax], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            float ahi = __hip_atomic_load(&nodes[cl].hi[ax], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            float bhi = __hip_atomic_load(&nodes[cr].hi[ax], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            nodes[p].lo[ax] = fminf(alo, blo);
            nodes[p].hi[ax] = fmaxf(ahi, bhi);
        }
        __threadfence();
        p = parent[p];
    }
}

/* the child of internal node p visited first in layout `lay` (axis lay/2,
 * positive direction for even lay): the one whose box centre is nearer */
__device__ static inline int bvh_first_child(const BvhNode *nodes, int p, int lay)
{
    const int a = lay >> 1;
    const int l = nodes[p].left, r = nodes[p].right;
    const float cl = nodes[l].lo[a] + nodes[l].hi[a], cr = nodes[r].lo[a] + nodes[r].hi[a];
    if (lay & 1) return cl >= cr ? l : r;
    return cl <= cr ? l : r;
}

/* Preorder flattening of layout `lay`, one thread per build-layout node
 * (nn of them).  Launched after k_bvh_boxes has completed, so plain loads
 * see its writes. */
__global__ void k_bvh_flatten(const BvhNode *nodes, const int *parent, int nn, int lay, BvhFlat *flat)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nn) return;
    int pre = 0;
    int y = x;
    /* depth is at most 30 Morton bits + 32 index bits of splits */
    for (int guard = 0; guard < 96; guard++) {
        int p = parent[y];
        if (p < 0) break;
        int f = bvh_first_child(nodes, p, lay);
        pre += (y == f) ? 1 : 1 + nodes[f].count;
        y = p;
    }
    const BvhNode &nd = nodes[x];
    BvhFlat f;
    for (int a = 0; a < 3; a++) { f.lo[a] = nd.lo[a]; f.hi[a] = nd.hi[a]; }
    f.esc = lay * nn + pre + nd.count;   /* absolute in the six layouts */
    const int cnt = nd.ntri > 0 ? nd.ntri : 1;
    f.tri = (nd.left < 0) ? (int32_t)((uint32_t)nd.right | ((uint32_t)(cnt - 1) << BVH_LEAF_POS_BITS)) : -1;
    flat[pre] = f;
}

/* a leaf record of triangle t (vertices P), host and device */
__host__ __device__ inline BvhTri bvh_tri_record(const float *P, uint32_t t)
{
    BvhTri r;
    for (int i = 0; i < 9; i++) r.v[i] = P[i];
    r.idx = t;
    r.pad[0] = r.pad[1] = 0u;
    return r;
}

/* the leaf-order triangle records: lt[k] holds triangle order[k] */
__global__ void k_bvh_leaf_tris(const float *pts, const uint32_t *order, int n, BvhTri *lt)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t t = order[k];
    lt[k] = bvh_tri_record(pts + 9 * (size_t)t, t);
}

#endif
