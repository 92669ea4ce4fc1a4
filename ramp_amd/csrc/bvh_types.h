/* bvh_types.h - the node records of a GMesh's BVH (host and device). */
#ifndef RAMP_BVH_TYPES_H
#define RAMP_BVH_TYPES_H

#include <stdint.h>

/* Build layout: an inner node's children, a leaf's position in the leaf
 * order, and the subtree's size in nodes. */
struct BvhNode {
    float lo[3];
    int32_t left;    /* >= 0: inner node, index of the left child; < 0: leaf */
    float hi[3];
    int32_t right;   /* inner: right child; leaf: its position in the leaf order */
    int32_t count;   /* the subtree's size in nodes (1 for a leaf) */
    int32_t ntri;    /* leaf: its triangles (0 counts as 1) */
    int32_t pad[2];
};

/* A leaf's triangles in leaf order: the vertices and the triangle's index
 * in the mesh (48 bytes, three 16-byte loads). */
struct BvhTri {
    float v[9];
    uint32_t idx;
    uint32_t pad[2];
};

/* a leaf's code in BvhFlat.tri: its first BvhTri and its count (1-4) */
#define BVH_LEAF_POS_BITS 28
#define BVH_LEAF_POS_MASK ((1u << BVH_LEAF_POS_BITS) - 1u)
#define BVH_LEAF_MAX_TRIS 4

/* Traversal layout: 32 bytes, preorder; esc = index of the node after this
 * subtree (counted from the first layout's start); tri = a leaf's code (first BvhTri | (count - 1) <<
 * BVH_LEAF_POS_BITS), -1 for an inner node.  BVH_LAYOUTS preorders of
 * GI_NNODES nodes (2 leaves - 1) follow each other, layout 2a + s
 * visiting the nearer child first for rays along axis a, positive (s = 0)
 * or negative (s = 1). */
#define BVH_LAYOUTS 6
struct BvhFlat {
    float lo[3];
    int32_t esc;
    float hi[3];
    int32_t tri;
};

#endif
