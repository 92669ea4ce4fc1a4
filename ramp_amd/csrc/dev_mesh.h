/* dev_mesh.h - triangle-mesh geometry (geom/mesh.cl:3-130) on the GPU.
 *
 * The reference tests the ray against every triangle.  Here the triangles
 * are visited through a bounding-volume hierarchy (a binned-SAH tree built
 * on the host, bvh_sah.h, or an LBVH built on the GPU, bvh_build.h) stored
 * in depth-first preorder and traversed without a stack: on a box hit an
 * internal node continues with the next node (its first-visited child, the
 * nearer one along the ray's dominant direction in the layout chosen for
 * the ray), a leaf tests its one or two triangles; a miss or a finished
 * leaf jumps to the node after the subtree.  A node is skipped only when
 * the ray provably cannot hit any of its triangles with t > eps (slab test
 * against the node box, widened by a relative margin so that no rounding
 * can reject a triangle the brute-force loop would accept).  Every triangle
 * that is visited runs the reference's exact Moller-Trumbore arithmetic;
 * ties in t are broken by the lowest triangle index, as the reference's
 * in-order strict '<' does, and the crossing count for 'interior' meshes
 * counts every triangle with t > eps, so results equal the brute-force
 * loop.
 */
#ifndef RAMP_DEV_MESH_H
#define RAMP_DEV_MESH_H

#include "bvh_types.h"

/* One Moller-Trumbore test of the reference (mesh.cl:54-111), from the
 * triangle's first vertex v0, edges e1 = v1 - v0, e2 = v2 - v0 and normal
 * nrm = e1 x e2.  Only the
 * distances of the nearest entering (t[0]) and exiting (t[1]) hits and
 * their triangles are kept: a hit point is pos + vel t, formed from them
 * when the intersection is written (mesh_commit). */
DEV void mesh_tri_core(rf3 v0, rf3 e1, rf3 e2, rf3 nrm, rf3 pos, rf3 vel, float *tb, uint32_t *out_idx,
                       uint32_t tri, uint32_t &nint)
{
    const float eps = 0.000000001f;
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
    float nv = v3_dot(nrm, vel);
    /* strict '<' in index order: an equal t wins only with a lower index */
    if ((t > eps) && (nv < 0.0f) && (t < tb[0] || (t == tb[0] && tri < out_idx[0]))) {
        tb[0] = t;
        out_idx[0] = tri;
    }
    if ((t > eps) && (nv > 0.0f) && (t < tb[1] || (t == tb[1] && tri < out_idx[1]))) {
        tb[1] = t;
        out_idx[1] = tri;
    }
}

/* the test from the triangle's three vertices P[0..8] */
DEV void mesh_tri(const float *P, rf3 pos, rf3 vel, float *tb, uint32_t *out_idx, uint32_t tri, uint32_t &nint)
{
    rf3 v0 = rf3_make(P[0], P[1], P[2]), v1 = rf3_make(P[3], P[4], P[5]);
    rf3 v2 = rf3_make(P[6], P[7], P[8]);
    rf3 e1 = v3_sub(v1, v0), e2 = v3_sub(v2, v0);
    mesh_tri_core(v0, e1, e2, v3_cross(e1, e2), pos, vel, tb, out_idx, tri, nint);
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

/* reference bounding-box precheck (mesh.cl:23-42): may the ray meet the
 * mesh's box at all? */
DEV bool mesh_box_precheck(const DPart &g, rf3 pos, rf3 vel)
{
    float t1 = (g.p[GP_XMIN] - pos.x) / vel.x, t2 = (g.p[GP_XMAX] - pos.x) / vel.x;
    float t3 = (g.p[GP_YMIN] - pos.y) / vel.y, t4 = (g.p[GP_YMAX] - pos.y) / vel.y;
    float t5 = (g.p[GP_ZMIN] - pos.z) / vel.z, t6 = (g.p[GP_ZMAX] - pos.z) / vel.z;
    float tmin = ocl_max(ocl_max(ocl_min(t1, t2), ocl_min(t3, t4)), ocl_min(t5, t6));
    float tmax = ocl_min(ocl_min(ocl_max(t1, t2), ocl_max(t3, t4)), ocl_max(t5, t6));
    return !(tmax < 0.0f || tmin > tmax);
}

/* the state of one ray's walk over the BVH (or the brute-force loop) */
struct MeshWalk {
    const BvhFlat *nodes;   /* the six layouts; null: brute force */
    rf3 pos, vel, inv;
    int cur, end;           /* next node (or triangle) and the end */
    float t[2];             /* nearest entering and exiting hit distances */
    uint32_t oidx[2];
    uint32_t nint;
};

/* start a walk with inv = 1/vel, componentwise */
DEV void mesh_walk_start(MeshWalk &w, const DPart &g, rf3 pos, rf3 vel, rf3 inv)
{
    w.pos = pos;
    w.vel = vel;
    w.t[0] = w.t[1] = 1e8f;
    w.oidx[0] = w.oidx[1] = 0xffffffffu;
    w.nint = 0;
    w.cur = 0;
    w.nodes = (const BvhFlat *)g.aux;
    if (w.nodes == nullptr) {
        w.end = (int)g.ni[GI_NTRI];
        return;
    }
    w.inv = inv;
    /* the layout of the ray's dominant direction: nodes [cur, end) */
    const int nn = (int)g.ni[GI_NNODES];
    const float ax = rm_fabs(vel.x), ay = rm_fabs(vel.y), az = rm_fabs(vel.z);
    const int lay = (ax >= ay && ax >= az) ? (vel.x < 0.0f ? 1 : 0)
                    : (ay >= az) ? (vel.y < 0.0f ? 3 : 2) : (vel.z < 0.0f ? 5 : 4);
    w.cur = lay * nn;
    w.end = w.cur + nn;
}

DEV void mesh_walk_init(MeshWalk &w, const DPart &g, rf3 pos, rf3 vel)
{
    mesh_walk_start(w, g, pos, vel, rf3_make(1.0f / vel.x, 1.0f / vel.y, 1.0f / vel.z));
}

/* One box step of the walk at node nd (= nodes[cur]): the code of the leaf
 * whose box the ray meets, for mesh_walk_leaf, or -1 (an inner node or a
 * miss).  cur strictly increases, so a walk ends after at most GI_NNODES
 * steps.  Testing a leaf's triangles later than its box step only leaves
 * the pruning limit larger meanwhile: more nodes are visited, none that
 * could improve a hit is skipped, and the result is the same. */
DEV int32_t mesh_walk_node(MeshWalk &w, const BvhFlat &nd, uint32_t ntri, bool interior)
{
    const float lim = interior ? 3.0e38f : rm_fmax(w.t[0], w.t[1]) * 1.0001f + 1e-30f;
    int next = nd.esc;
    int32_t code = -1;
    if (bvh_box_hit(nd, w.pos, w.inv, lim)) {
        if (nd.tri < 0) next = w.cur + 1;
        else if (((uint32_t)nd.tri & BVH_LEAF_POS_MASK) < ntri) code = nd.tri;
    }
    w.cur = next > w.cur ? next : w.cur + 1;
    return code;
}

/* the triangles of a leaf (its code from mesh_walk_node) */
DEV void mesh_walk_leaf(MeshWalk &w, const BvhTri *lt, int32_t code)
{
    const uint32_t first = (uint32_t)code & BVH_LEAF_POS_MASK;
    const uint32_t cnt = ((uint32_t)code >> BVH_LEAF_POS_BITS) + 1u;
    for (uint32_t k = 0; k < cnt; k++) {
        const BvhTri &r = lt[first + k];
        mesh_tri(r.v, w.pos, w.vel, w.t, w.oidx, r.idx, w.nint);
    }
}

/* one triangle of the brute-force loop */
DEV void mesh_walk_brute(MeshWalk &w, const float *pts, uint32_t tri)
{
    mesh_tri(pts + 9 * (size_t)tri, w.pos, w.vel, w.t, w.oidx, tri, w.nint);
}

/* One step of the walk with its triangle tests; true once the walk is over. */
DEV bool mesh_walk_step(MeshWalk &w, const DPart &g)
{
    if (w.cur >= w.end) return true;
    if (w.nodes == nullptr) {
        mesh_walk_brute(w, g.tab[GT_POINTS], (uint32_t)w.cur);
        w.cur++;
    } else {
        const int32_t code = mesh_walk_node(w, w.nodes[w.cur], g.ni[GI_NTRI], g.ni[GI_INTERIOR] != 0);
        if (code >= 0) mesh_walk_leaf(w, (const BvhTri *)g.aux2, code);
    }
    return w.cur >= w.end;
}

/* the walk's outcome (mesh.cl:114-123): 0 no intersection is written (an
 * 'interior' mesh crossed an even number of times), 1 the nearest entering
 * and exiting hits, 2 the entering hit as both (an odd count) */
DEV uint32_t mesh_walk_code(const MeshWalk &w, uint32_t interior)
{
    if (interior == 1) return (w.nint % 2) == 1 ? 2u : 0u;
    return 1u;
}

/* write the intersection from the hit distances t1 (entering) and t2
 * (exiting) and the outcome code when it is nearer than the current one;
 * a hit point is pos + vel t, as the walk formed it (an unset hit keeps
 * (0, 0, 0, 1e8)) */
DEV void mesh_commit(NS &n, rf3 pos, rf3 vel, float t1, float t2, uint32_t code, uint32_t ci)
{
    if (code == 0u) return;
    float o[8];
    if (t1 < 1e8f) {
        rf3 hp = v3_add(pos, v3_mulv(vel, t1));
        o[0] = hp.x; o[1] = hp.y; o[2] = hp.z;
    } else {
        o[0] = o[1] = o[2] = 0.0f;
    }
    o[3] = t1;
    if (code == 2u) {
        o[4] = o[0]; o[5] = o[1]; o[6] = o[2]; o[7] = t1;
    } else if (t2 < 1e8f) {
        rf3 hp = v3_add(pos, v3_mulv(vel, t2));
        o[4] = hp.x; o[5] = hp.y; o[6] = hp.z; o[7] = t2;
    } else {
        o[4] = o[5] = o[6] = 0.0f; o[7] = t2;
    }
    if (o[3] < n.is[3]) {
        for (int i = 0; i < 8; i++) n.is[i] = o[i];
        n.iidx = ci;
    }
}

/* geom/mesh.cl:3-130 */
DEV void g_mesh(NS &n, const DPart &g, uint32_t ci)
{
    rf3 pos = npos(n), vel = nvel(n);
    if (!mesh_box_precheck(g, pos, vel)) return;
    MeshWalk w;
    mesh_walk_init(w, g, pos, vel);
    while (!mesh_walk_step(w, g)) {
    }
    mesh_commit(n, pos, vel, w.t[0], w.t[1], mesh_walk_code(w, g.ni[GI_INTERIOR]), ci);
}

#endif
