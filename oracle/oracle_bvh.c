/* ORACLE -- test infrastructure only (see oracle_math.h).
 * Binned-SAH BVH build (16 bins per axis, leaves of at most 4 triangles) for
 * the oracle's traversal, plus the oracle's restatement of the reference's
 * host alias-table construction (src/core/scene/lighting.c:108-164). */
#include "oracle_bvh.h"

#include <float.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    float lo[3], hi[3];
} Box;

static void boxEmpty(Box* b) {
    for (int a = 0; a < 3; a++) {
        b->lo[a] = FLT_MAX;
        b->hi[a] = -FLT_MAX;
    }
}
static void boxGrowPt(Box* b, const float* p) {
    for (int a = 0; a < 3; a++) {
        if (p[a] < b->lo[a]) b->lo[a] = p[a];
        if (p[a] > b->hi[a]) b->hi[a] = p[a];
    }
}
static void boxGrow(Box* b, const Box* o) {
    for (int a = 0; a < 3; a++) {
        if (o->lo[a] < b->lo[a]) b->lo[a] = o->lo[a];
        if (o->hi[a] > b->hi[a]) b->hi[a] = o->hi[a];
    }
}
static double boxArea(const Box* b) {
    double d[3];
    for (int a = 0; a < 3; a++) d[a] = b->hi[a] > b->lo[a] ? (double)b->hi[a] - b->lo[a] : 0.0;
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
}

typedef struct {
    const float* tris;
    Box* triBox;
    float* cent;
    uint32_t* order;
    ONode* nodes;
    uint32_t count;
} Builder;

#define NBINS 16
#define LEAF_MAX 4

static void buildRange(Builder* bd, uint32_t nodeIdx, uint32_t first, uint32_t n) {
    ONode* node = &bd->nodes[nodeIdx];
    Box b, cb;
    boxEmpty(&b);
    boxEmpty(&cb);
    for (uint32_t i = first; i < first + n; i++) {
        boxGrow(&b, &bd->triBox[bd->order[i]]);
        boxGrowPt(&cb, &bd->cent[bd->order[i] * 3]);
    }
    memcpy(node->lo, b.lo, sizeof(b.lo));
    memcpy(node->hi, b.hi, sizeof(b.hi));
    node->left = -1;
    node->first = first;
    node->count = n;
    if (n <= LEAF_MAX) return;
    double bestCost = DBL_MAX;
    int bestAxis = -1, bestSplit = -1;
    for (int a = 0; a < 3; a++) {
        float ext = cb.hi[a] - cb.lo[a];
        if (!(ext > 0.0f)) continue;
        Box bins[NBINS];
        uint32_t cnt[NBINS];
        for (int k = 0; k < NBINS; k++) {
            boxEmpty(&bins[k]);
            cnt[k] = 0;
        }
        float scale = NBINS / ext;
        for (uint32_t i = first; i < first + n; i++) {
            uint32_t t = bd->order[i];
            int k = (int)((bd->cent[t * 3 + a] - cb.lo[a]) * scale);
            if (k >= NBINS) k = NBINS - 1;
            if (k < 0) k = 0;
            cnt[k]++;
            boxGrow(&bins[k], &bd->triBox[t]);
        }
        double rightArea[NBINS];
        uint32_t rightCnt[NBINS];
        Box acc;
        boxEmpty(&acc);
        uint32_t c = 0;
        for (int k = NBINS - 1; k > 0; k--) {
            boxGrow(&acc, &bins[k]);
            c += cnt[k];
            rightArea[k] = boxArea(&acc);
            rightCnt[k] = c;
        }
        boxEmpty(&acc);
        c = 0;
        for (int k = 0; k < NBINS - 1; k++) {
            boxGrow(&acc, &bins[k]);
            c += cnt[k];
            if (c == 0 || rightCnt[k + 1] == 0) continue;
            double cost = boxArea(&acc) * c + rightArea[k + 1] * rightCnt[k + 1];
            if (cost < bestCost) {
                bestCost = cost;
                bestAxis = a;
                bestSplit = k;
            }
        }
    }
    uint32_t mid;
    if (bestAxis < 0) {
        mid = first + n / 2; /* all centroids coincide: split by count */
    } else {
        float scale = NBINS / (cb.hi[bestAxis] - cb.lo[bestAxis]);
        uint32_t i = first, j = first + n;
        while (i < j) {
            uint32_t t = bd->order[i];
            int k = (int)((bd->cent[t * 3 + bestAxis] - cb.lo[bestAxis]) * scale);
            if (k >= NBINS) k = NBINS - 1;
            if (k <= bestSplit) {
                i++;
            } else {
                j--;
                uint32_t tmp = bd->order[i];
                bd->order[i] = bd->order[j];
                bd->order[j] = tmp;
            }
        }
        mid = i;
        if (mid == first || mid == first + n) mid = first + n / 2;
    }
    uint32_t left = bd->count;
    bd->count += 2;
    node->left = (int32_t)left;
    node->count = 0;
    buildRange(bd, left, first, mid - first);
    buildRange(bd, left + 1, mid, first + n - mid);
}

OBvh* oracle_bvh_build(const float* tris9, uint32_t triCount) {
    if (!tris9 || triCount == 0) return NULL;
    Builder bd;
    memset(&bd, 0, sizeof(bd));
    bd.tris = tris9;
    bd.triBox = (Box*)malloc(sizeof(Box) * triCount);
    bd.cent = (float*)malloc(sizeof(float) * 3 * triCount);
    bd.order = (uint32_t*)malloc(sizeof(uint32_t) * triCount);
    bd.nodes = (ONode*)malloc(sizeof(ONode) * (2u * triCount + 1u));
    if (!bd.triBox || !bd.cent || !bd.order || !bd.nodes) {
        free(bd.triBox);
        free(bd.cent);
        free(bd.order);
        free(bd.nodes);
        return NULL;
    }
    for (uint32_t t = 0; t < triCount; t++) {
        boxEmpty(&bd.triBox[t]);
        for (int v = 0; v < 3; v++) boxGrowPt(&bd.triBox[t], &tris9[(size_t)t * 9 + v * 3]);
        for (int a = 0; a < 3; a++) bd.cent[t * 3 + a] = 0.5f * (bd.triBox[t].lo[a] + bd.triBox[t].hi[a]);
        bd.order[t] = t;
    }
    bd.count = 1;
    buildRange(&bd, 0, 0, triCount);
    free(bd.triBox);
    free(bd.cent);
    OBvh* b = (OBvh*)malloc(sizeof(OBvh));
    if (!b) {
        free(bd.order);
        free(bd.nodes);
        return NULL;
    }
    b->nodes = bd.nodes;
    b->order = bd.order;
    b->nodeCount = bd.count;
    return b;
}

void oracle_bvh_free(OBvh* b) {
    if (!b) return;
    free(b->nodes);
    free(b->order);
    free(b);
}

uint32_t oracle_bvh_node_count(const OBvh* b) { return b ? b->nodeCount : 0u; }

/* Restatement of src/core/scene/lighting.c:108-164 (Vose alias method in fp32
 * with LIFO small/large work lists; leftover entries get q = 1). */
int oracle_build_alias(const float* pmf, uint32_t n, float* q, uint32_t* alias) {
    if (n == 0) return 1;
    float* p = (float*)malloc(sizeof(float) * n);
    uint32_t* sm = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* lg = (uint32_t*)malloc(sizeof(uint32_t) * n);
    if (!p || !sm || !lg) {
        free(p);
        free(sm);
        free(lg);
        return 0;
    }
    uint32_t ns = 0, nl = 0;
    for (uint32_t i = 0; i < n; i++) {
        p[i] = pmf[i] * (float)n;
        if (p[i] < 1.0f)
            sm[ns++] = i;
        else
            lg[nl++] = i;
    }
    while (ns && nl) {
        uint32_t a = sm[--ns], g = lg[--nl];
        q[a] = p[a];
        alias[a] = g;
        p[g] = (p[g] + p[a]) - 1.0f;
        if (p[g] < 1.0f)
            sm[ns++] = g;
        else
            lg[nl++] = g;
    }
    while (nl) {
        uint32_t g = lg[--nl];
        q[g] = 1.0f;
        alias[g] = g;
    }
    while (ns) {
        uint32_t a = sm[--ns];
        q[a] = 1.0f;
        alias[a] = a;
    }
    free(p);
    free(sm);
    free(lg);
    return 1;
}
