/* bvh_sah.h - binned surface-area-heuristic BVH over a GMesh's triangles,
 * built on the host when the plan is created.
 *
 * Method (Wald, "On fast construction of SAH-based bounding volume
 * hierarchies", IRT 2007): top-down; at every node the triangle centroids
 * are binned into SAH_BINS equal bins along each axis of the centroid box,
 * and the split plane between bins minimising
 *     n_left * area(box_left) + n_right * area(box_right)
 * is taken (boxes of the triangles, widened by the leaf margin).  A node
 * whose centroids coincide, or whose best split leaves a side empty, is cut
 * at the middle of its range.  Below depth SAH_MEDIAN_DEPTH every node is
 * cut at the middle, so the depth stays under the 96 levels the preorder
 * pass walks.  A range of at most SAH_LEAF_TRIS triangles is a leaf.
 *
 * The output is the build layout of bvh_build.h (nodes, parents, leaf
 * order), which k_bvh_flatten writes as the six direction-sorted preorder
 * layouts; only the tree's shape differs from the LBVH, so traversal and
 * its equality with the brute-force loop are unchanged.  On the imaging
 * workload's 50k-triangle sphere a ray that meets the mesh box visits 82
 * nodes and tests 15 triangles, against the one-triangle-leaf LBVH's 125
 * nodes and 13 triangles.
 */
#ifndef RAMP_BVH_SAH_H
#define RAMP_BVH_SAH_H

#include <algorithm>
#include <vector>

#define SAH_BINS 32
#define SAH_MEDIAN_DEPTH 48
#define SAH_LEAF_TRIS 2

struct SahBuilder {
    const float *P;
    int n;
    float margin;
    std::vector<float> cen, tlo, thi;   /* per triangle: centroid and widened box */
    std::vector<uint32_t> idx;          /* triangles, partitioned in place */
    std::vector<BvhNode> nodes;
    std::vector<int> parent;
    std::vector<uint32_t> order;

    static float area(const float *lo, const float *hi)
    {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return dx * dy + dy * dz + dz * dx;
    }

    void init()
    {
        cen.resize(3 * (size_t)n);
        tlo.resize(3 * (size_t)n);
        thi.resize(3 * (size_t)n);
        idx.resize(n);
        for (int t = 0; t < n; t++) {
            const float *p = P + 9 * (size_t)t;
            idx[t] = (uint32_t)t;
            for (int a = 0; a < 3; a++) {
                cen[3 * (size_t)t + a] = (p[a] + p[3 + a] + p[6 + a]) * (1.0f / 3.0f);
                tlo[3 * (size_t)t + a] = std::min(p[a], std::min(p[3 + a], p[6 + a])) - margin;
                thi[3 * (size_t)t + a] = std::max(p[a], std::max(p[3 + a], p[6 + a])) + margin;
            }
        }
        nodes.assign(2 * (size_t)n - 1, BvhNode());
        parent.assign(2 * (size_t)n - 1, -1);
        order.assign(n, 0u);
    }

    /* the split position in [b, e) for node range [b, e), b + 1 < e */
    int split(int b, int e, int depth)
    {
        const int mid = b + (e - b) / 2;
        if (depth >= SAH_MEDIAN_DEPTH) return mid;
        float clo[3] = {3e38f, 3e38f, 3e38f}, chi[3] = {-3e38f, -3e38f, -3e38f};
        for (int i = b; i < e; i++)
            for (int a = 0; a < 3; a++) {
                clo[a] = std::min(clo[a], cen[3 * (size_t)idx[i] + a]);
                chi[a] = std::max(chi[a], cen[3 * (size_t)idx[i] + a]);
            }
        float best = 3e38f;
        int best_axis = -1, best_bin = -1;
        for (int a = 0; a < 3; a++) {
            if (!(chi[a] > clo[a])) continue;
            const float scale = (float)SAH_BINS / (chi[a] - clo[a]);
            float blo[SAH_BINS][3], bhi[SAH_BINS][3];
            int bc[SAH_BINS] = {0};
            for (int k = 0; k < SAH_BINS; k++)
                for (int q = 0; q < 3; q++) { blo[k][q] = 3e38f; bhi[k][q] = -3e38f; }
            for (int i = b; i < e; i++) {
                const size_t t = idx[i];
                const int k = std::min(SAH_BINS - 1, (int)((cen[3 * t + a] - clo[a]) * scale));
                bc[k]++;
                for (int q = 0; q < 3; q++) {
                    blo[k][q] = std::min(blo[k][q], tlo[3 * t + q]);
                    bhi[k][q] = std::max(bhi[k][q], thi[3 * t + q]);
                }
            }
            /* right-hand sweeps: area and count of bins k..SAH_BINS-1 */
            float rarea[SAH_BINS];
            int rcount[SAH_BINS];
            float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
            int c = 0;
            for (int k = SAH_BINS - 1; k > 0; k--) {
                for (int q = 0; q < 3; q++) { lo[q] = std::min(lo[q], blo[k][q]); hi[q] = std::max(hi[q], bhi[k][q]); }
                c += bc[k];
                rcount[k] = c;
                rarea[k] = c ? area(lo, hi) : 0.0f;
            }
            float llo[3] = {3e38f, 3e38f, 3e38f}, lhi[3] = {-3e38f, -3e38f, -3e38f};
            int lc = 0;
            for (int k = 0; k < SAH_BINS - 1; k++) {
                for (int q = 0; q < 3; q++) { llo[q] = std::min(llo[q], blo[k][q]); lhi[q] = std::max(lhi[q], bhi[k][q]); }
                lc += bc[k];
                if (lc == 0 || rcount[k + 1] == 0) continue;
                const float cost = (float)lc * area(llo, lhi) + (float)rcount[k + 1] * rarea[k + 1];
                if (cost < best) { best = cost; best_axis = a; best_bin = k; }
            }
        }
        if (best_axis < 0) return mid;
        const int a = best_axis;
        const float scale = (float)SAH_BINS / (chi[a] - clo[a]);
        uint32_t *m = std::partition(idx.data() + b, idx.data() + e, [&](uint32_t t) {
            return std::min(SAH_BINS - 1, (int)((cen[3 * (size_t)t + a] - clo[a]) * scale)) <= best_bin;
        });
        const int s = (int)(m - idx.data());
        return (s == b || s == e) ? mid : s;
    }

    /* Build the subtree of range [b, e) under `par`; returns its node and
     * takes nodes from `next` on (a subtree of m leaves takes 2m - 1). */
    int build(int b, int e, int par, int &next, int depth)
    {
        const int me = next++;
        parent[me] = par;
        BvhNode &nd = nodes[me];
        if (e - b <= SAH_LEAF_TRIS) {
            nd.left = -1;
            nd.right = b;
            nd.count = 1;
            nd.ntri = e - b;
            for (int a = 0; a < 3; a++) { nd.lo[a] = 3e38f; nd.hi[a] = -3e38f; }
            for (int i = b; i < e; i++) {
                const size_t t = idx[i];
                order[i] = (uint32_t)t;
                for (int a = 0; a < 3; a++) {
                    nd.lo[a] = std::min(nd.lo[a], tlo[3 * t + a]);
                    nd.hi[a] = std::max(nd.hi[a], thi[3 * t + a]);
                }
            }
            return me;
        }
        const int s = split(b, e, depth);
        const int l = build(b, s, me, next, depth + 1);
        const int r = build(s, e, me, next, depth + 1);
        BvhNode &m = nodes[me];
        m.left = l;
        m.right = r;
        m.count = 1 + nodes[l].count + nodes[r].count;
        for (int a = 0; a < 3; a++) {
            m.lo[a] = std::min(nodes[l].lo[a], nodes[r].lo[a]);
            m.hi[a] = std::max(nodes[l].hi[a], nodes[r].hi[a]);
        }
        return me;
    }
};

#endif
