/* ORACLE -- test infrastructure only (see oracle_math.h).
 * Binary BVH over world-space triangles for the oracle's own traversal: a
 * binned-SAH build written independently of the product's GPU LBVH. */
#ifndef ORACLE_BVH_H
#define ORACLE_BVH_H

#include <stdint.h>

typedef struct ONode {
    float lo[3], hi[3];
    int32_t left; /* internal: index of the left child (right = left + 1); leaf: -1 */
    uint32_t first, count;
} ONode;

typedef struct OBvh {
    ONode* nodes;
    uint32_t* order;
    uint32_t nodeCount;
} OBvh;

OBvh* oracle_bvh_build(const float* tris9, uint32_t triCount);
void oracle_bvh_free(OBvh* bvh);
uint32_t oracle_bvh_node_count(const OBvh* bvh);

#endif
