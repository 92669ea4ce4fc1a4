/* dev_mesh.h - triangle-mesh geometry (geom/mesh.cl:3-130) on the GPU.
 *
 * The reference tests the ray against every triangle.  Here the triangles
 * are visited through a bounding-volume hierarchy built on the GPU
 * (bvh_build.h) and stored in depth-first preorder, traversed without a
 * stack: on a box hit an internal node continues with the next node (its
 * left child), a leaf tests its triangle; a miss or a finished leaf jumps to
 * the node after the subtree.  A node is skipped only when the ray provably
 * cannot hit any of its triangles with t > eps (slab test against the node
 * box, widened by a relative margin so that no rounding can reject a
 * triangle the brute-force loop would accept).  Every triangle that is
 * visited runs the reference's exact Moller-Trumbore arithmetic; ties in t
 * are broken by the lowest triangle index, as the reference's in-order
 * strict '<' does, and the crossing count for 'interior' meshes counts every
 * triangle with t > eps, so results equal the brute-force loop.
 */
#ifndef RAMP_DEV_MESH_H
#define RAMP_DEV_MESH_H

struct BvhNode {
    float lo[3];
    int32_t left;    /* >= 0: inner node, index of the left child; < 0: leaf */
    float hi[3];
    int32_t right;   /* inner: right child; leaf: first primitive */
    int32_t count;   /* leaf: number of primitives */
    int32_t pad[3];
};

/* Traversal layout: 32 bytes, preorder; esc = preorder index of the node
 * after this subtree; tri = triangle index of a leaf, -1 for an inner node. */
struct BvhFlat {
    float lo[3];
    int32_t esc;
    float hi[3];
    int32_t tri;
};

/* One Moller-Trumbore test of the reference (mesh.cl:54-111). */
DEV void mesh_tri(const float *P, rf3 pos, rf3 vel, float *out, uint32_t *out_idx,
                  uint32_t tri, uint32_t &nint)
{
    const float eps = 0.000000001f;
    rf3 v0 = rf3_make(P[0], P[1], P[2]), v1 = rf3_make(P[3], P[4], P[5]);
    rf3 v2 = rf3_make(P[6], P[7], P[8]);
    rf3 e1 = v3_sub(v1, v0), e2 = v3_sub(v2, v0);
    rf3 h = v3_cross(vel, e2);
    float a = v3_dot(e1, h);
    if (rm_fabs(a) < eps) return;
    float f = 1.0f / a;
    rf3 s = v3_sub(pos, v0);
    float u = f * v3_dot(s, h);
    if (u < 0.0f || u > 1.0f) return;
    rf3 q = v3_cross(s, e1);
    float v = f * v3_dot(vel, q);
    if (v < 0.0f || u + v > 1.0f) return;
    float t = f * v3_dot(e2, q);
    if (t > eps) nint++;
    rf3 nrm = v3_cross(e1, e2);
    float nv = v3_dot(nrm, vel);
    /* strict '<' in index order: an equal t wins only with a lower index */
    if ((t > eps) && (nv < 0.0f) && (t < out[3] || (t == out[3] && tri < out_idx[0]))) {
        rf3 hp = v3_add(pos, v3_mulv(vel, t));
        out[0] = hp.x; out[1] = hp.y; out[2] = hp.z; out[3] = t;
        out_idx[0] = tri;
    }
    if ((t > eps) && (nv > 0.0f) && (t < out[7] || (t == out[7] && tri < out_idx[1]))) {
        rf3 hp = v3_add(pos, v3_mulv(vel, t));
        out[4] = hp.x; out[5] = hp.y; out[6] = hp.z; out[7] = t;
        out_idx[1] = tri;
    }
}

/* conservative slab test: may the ray meet the box at some t > 0 that
 * could improve a hit bounded by tmax_need (or any t when counting)? */
DEV bool bvh_box_hit(const BvhFlat &nd, rf3 pos, rf3 inv, float tlimit)
{
    float tx1 = (nd.lo[0] - pos.x) * inv.x, tx2 = (nd.hi[0] - pos.x) * inv.x;
    float ty1 = (nd.lo[1] - pos.y) * inv.y, ty2 = (nd.hi[1] - pos.y) * inv.y;
    float tz1 = (nd.lo[2] - pos.z) * inv.z, tz2 = (nd.hi[2] - pos.z) * inv.z;
    float tmin = rm_fmax(rm_fmax(rm_fmin(tx1, tx2), rm_fmin(ty1, ty2)), rm_fmin(tz1, tz2));
    float tmax = rm_fmin(rm_fmin(rm_fmax(tx1, tx2), rm_fmax(ty1, ty2)), rm_fmax(tz1, tz2));
    /* NaN (0*inf on a slab face) keeps the node: fmin/fmax drop NaN */
    return tmax >= 0.0f && tmin <= tmax && tmin <= tlimit;
}

/* geom/mesh.cl:3-130 */
DEV void g_mesh(NS &n, const DPart &g, uint32_t ci)
{
    rf3 pos = npos(n), vel = nvel(n);
    /* reference bounding-box precheck (mesh.cl:23-42) */
    float t1 = (g.p[GP_XMIN] - pos.x) / vel.x, t2 = (g.p[GP_XMAX] - pos.x) / vel.x;
    float t3 = (g.p[GP_YMIN] - pos.y) / vel.y, t4 = (g.p[GP_YMAX] - pos.y) / vel.y;
    float t5 = (g.p[GP_ZMIN] - pos.z) / vel.z, t6 = (g.p[GP_ZMAX] - pos.z) / vel.z;
    float tmin = ocl_max(ocl_max(ocl_min(t1, t2), ocl_min(t3, t4)), ocl_min(t5, t6));
    float tmax = ocl_min(ocl_min(ocl_max(t1, t2), ocl_max(t3, t4)), ocl_max(t5, t6));
    if (tmax < 0.0f || tmin > tmax) return;
    const float *pts = g.tab[GT_POINTS];
    uint32_t ntri = g.ni[GI_NTRI];
    uint32_t interior = g.ni[GI_INTERIOR];
    float out[8] = {0.0f, 0.0f, 0.0f, 1e8f, 0.0f, 0.0f, 0.0f, 1e8f};
    uint32_t oidx[2] = {0xffffffffu, 0xffffffffu};
    uint32_t nint = 0;
    const BvhFlat *nodes = (const BvhFlat *)g.aux;
    if (nodes == nullptr) {
        for (uint32_t i = 0; i < ntri; i++) mesh_tri(pts + 9 * (size_t)i, pos, vel, out, oidx, i, nint);
    } else {
        rf3 inv = rf3_make(1.0f / vel.x, 1.0f / vel.y, 1.0f / vel.z);
        const int end = 2 * (int)ntri - 1;
        int cur = 0;
        /* cur strictly increases, so the loop ends after at most 2n-1 nodes */
        while (cur < end) {
            const BvhFlat nd = nodes[cur];
            float lim = interior ? 3.0e38f : rm_fmax(out[3], out[7]) * 1.0001f + 1e-30f;
            int next = nd.esc;
            if (bvh_box_hit(nd, pos, inv, lim)) {
                if ((uint32_t)nd.tri < ntri)
                    mesh_tri(pts + 9 * (size_t)nd.tri, pos, vel, out, oidx, (uint32_t)nd.tri, nint);
                else if (nd.tri < 0)
                    next = cur + 1;
            }
            cur = next > cur ? next : cur + 1;
        }
    }
    if (interior == 1 && (nint % 2) == 1) {
        out[4] = out[0]; out[5] = out[1]; out[6] = out[2]; out[7] = out[3];
    } else if (interior == 1 && (nint % 2) == 0) {
        return;
    }
    if (out[3] < n.is[3]) {
        for (int i = 0; i < 8; i++) n.is[i] = out[i];
        n.iidx = ci;
    }
}

#endif
