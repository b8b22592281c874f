/*
 * Camera and mesh-transform math for the host side.
 *
 * Camera: the reference builds a right-handed look-at view and an OpenGL-style
 * (depth -1..1) perspective with cglm, flips proj[1][1] for Vulkan's y-down
 * clip space and uploads both inverses (src/core/scene/camera.c:128-143). The
 * kernels turn those inverses into primary rays exactly as
 * src/shaders/camera/ray.slang:15-30. Matrices are column-major m[col][row].
 *
 * Mesh transforms: world = T * Rz * Ry * Rx * S with Euler angles in degrees
 * (src/core/scene/transform.c:26-35); the shading code rotates X, then Y,
 * then Z (src/shaders/geometry/surface/transform.slang:27-30).
 */
#include "internal.h"

#include <math.h>
#include <string.h>

static void mat4Identity(float m[4][4]) {
    memset(m, 0, sizeof(float) * 16);
    m[0][0] = m[1][1] = m[2][2] = m[3][3] = 1.0f;
}

static void mat4Mul(const float a[4][4], const float b[4][4], float out[4][4]) {
    float r[4][4];
    for (int c = 0; c < 4; c++) {
        for (int row = 0; row < 4; row++) {
            r[c][row] = a[0][row] * b[c][0] + a[1][row] * b[c][1] + a[2][row] * b[c][2] + a[3][row] * b[c][3];
        }
    }
    memcpy(out, r, sizeof(r));
}

/* General 4x4 inverse by cofactors; returns 0 for a singular matrix. */
static int mat4Inverse(const float m[4][4], float out[4][4]) {
    double a[16], inv[16];
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) a[c * 4 + r] = m[c][r];
    inv[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] + a[9] * a[7] * a[14] +
             a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
    inv[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] - a[8] * a[7] * a[14] -
             a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
    inv[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] + a[8] * a[7] * a[13] +
             a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
    inv[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] - a[8] * a[6] * a[13] -
              a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
    inv[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] - a[9] * a[3] * a[14] -
             a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
    inv[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] + a[8] * a[3] * a[14] +
             a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
    inv[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] - a[8] * a[3] * a[13] -
             a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
    inv[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] + a[8] * a[2] * a[13] +
              a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
    inv[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] + a[5] * a[3] * a[14] +
             a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
    inv[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] - a[4] * a[3] * a[14] -
             a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
    inv[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] + a[4] * a[3] * a[13] +
              a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
    inv[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] - a[4] * a[2] * a[13] -
              a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
    inv[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] - a[5] * a[3] * a[10] -
             a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
    inv[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] + a[4] * a[3] * a[10] +
             a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
    inv[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] - a[4] * a[3] * a[9] -
              a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
    inv[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] + a[4] * a[2] * a[9] +
              a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
    double det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
    if (det == 0.0) return 0;
    double id = 1.0 / det;
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) out[c][r] = (float)(inv[c * 4 + r] * id);
    return 1;
}

static void normalize3(float v[3]) {
    float l = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (l > 0.0f) {
        v[0] /= l;
        v[1] /= l;
        v[2] /= l;
    }
}

static void cross3(const float a[3], const float b[3], float o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

static float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* Right-handed look-at (same convention as cglm glm_lookat_rh). */
static void lookAt(const float eye[3], const float center[3], const float up[3], float m[4][4]) {
    float f[3] = {center[0] - eye[0], center[1] - eye[1], center[2] - eye[2]};
    normalize3(f);
    float s[3];
    cross3(f, up, s);
    normalize3(s);
    float u[3];
    cross3(s, f, u);
    mat4Identity(m);
    m[0][0] = s[0];
    m[0][1] = u[0];
    m[0][2] = -f[0];
    m[1][0] = s[1];
    m[1][1] = u[1];
    m[1][2] = -f[1];
    m[2][0] = s[2];
    m[2][1] = u[2];
    m[2][2] = -f[2];
    m[3][0] = -dot3(s, eye);
    m[3][1] = -dot3(u, eye);
    m[3][2] = dot3(f, eye);
}

/* Right-handed perspective with clip depth -1..1 (cglm glm_perspective_rh_no). */
static void perspective(float fovy, float aspect, float nearZ, float farZ, float m[4][4]) {
    memset(m, 0, sizeof(float) * 16);
    float f = 1.0f / tanf(fovy * 0.5f);
    float fn = 1.0f / (nearZ - farZ);
    m[0][0] = f / aspect;
    m[1][1] = f;
    m[2][2] = (nearZ + farZ) * fn;
    m[2][3] = -1.0f;
    m[3][2] = 2.0f * nearZ * farZ * fn;
}

void vkrtSyncCameraMatrices(VKRT* vkrt) {
    float view[4][4], proj[4][4];
    const Camera* cam = &vkrt->settings.camera;
    uint32_t w = vkrt->sceneData.viewportRect[2] > 0u ? vkrt->sceneData.viewportRect[2] : 1u;
    uint32_t h = vkrt->sceneData.viewportRect[3] > 0u ? vkrt->sceneData.viewportRect[3] : 1u;
    lookAt(cam->pos, cam->target, cam->up, view);
    perspective(cam->vfov * (float)(M_PI / 180.0), (float)w / (float)h, cam->nearZ, cam->farZ, proj);
    proj[1][1] *= -1.0f;
    float vi[4][4], pi[4][4];
    if (!mat4Inverse(view, vi)) mat4Identity(vi);
    if (!mat4Inverse(proj, pi)) mat4Identity(pi);
    memcpy(vkrt->sceneData.viewInverse, vi, sizeof(vi));
    memcpy(vkrt->sceneData.projInverse, pi, sizeof(pi));
}

static void rotationAxisMatrix(int axis, float radians, float m[4][4]) {
    mat4Identity(m);
    float c = cosf(radians), s = sinf(radians);
    int a = (axis + 1) % 3, b = (axis + 2) % 3;
    m[a][a] = c;
    m[b][b] = c;
    m[a][b] = s;  /* column a, row b */
    m[b][a] = -s; /* column b, row a */
}

void VKRT_buildMeshTransformMatrix(const float position[3], const float rotationDegrees[3], const float scale[3],
                                   float outMatrix[4][4]) {
    if (!position || !rotationDegrees || !scale || !outMatrix) return;
    const float d2r = (float)(M_PI / 180.0);
    float t[4][4], rx[4][4], ry[4][4], rz[4][4], s[4][4], tmp[4][4];
    mat4Identity(t);
    t[3][0] = position[0];
    t[3][1] = position[1];
    t[3][2] = position[2];
    rotationAxisMatrix(0, rotationDegrees[0] * d2r, rx);
    rotationAxisMatrix(1, rotationDegrees[1] * d2r, ry);
    rotationAxisMatrix(2, rotationDegrees[2] * d2r, rz);
    mat4Identity(s);
    s[0][0] = scale[0];
    s[1][1] = scale[1];
    s[2][2] = scale[2];
    mat4Mul(t, rz, tmp);
    mat4Mul(tmp, ry, tmp);
    mat4Mul(tmp, rx, tmp);
    mat4Mul(tmp, s, outMatrix);
}

static void eulerFromRotation(const float r[4][4], float outDegrees[3]) {
    /* R = Rz(g) Ry(b) Rx(a); element (row, col) = r[col][row] */
    float r20 = r[0][2], r21 = r[1][2], r22 = r[2][2], r10 = r[0][1], r00 = r[0][0];
    float sb = -r20;
    if (sb > 1.0f) sb = 1.0f;
    if (sb < -1.0f) sb = -1.0f;
    float b = asinf(sb), a, g;
    if (fabsf(sb) < 0.9999999f) {
        a = atan2f(r21, r22);
        g = atan2f(r10, r00);
    } else {
        a = 0.0f;
        g = atan2f(-r[1][0], r[1][1]);
    }
    const float r2d = (float)(180.0 / M_PI);
    outDegrees[0] = a * r2d;
    outDegrees[1] = b * r2d;
    outDegrees[2] = g * r2d;
}

void VKRT_decomposeMeshTransform(float worldTransform[4][4], float outPosition[3], float outRotation[3],
                                 float outScale[3]) {
    if (!worldTransform || !outPosition || !outRotation || !outScale) return;
    outPosition[0] = worldTransform[3][0];
    outPosition[1] = worldTransform[3][1];
    outPosition[2] = worldTransform[3][2];
    float rot[4][4];
    mat4Identity(rot);
    float absScale[3] = {1.0f, 1.0f, 1.0f};
    for (int axis = 0; axis < 3; axis++) {
        float col[3] = {worldTransform[axis][0], worldTransform[axis][1], worldTransform[axis][2]};
        float len = sqrtf(dot3(col, col));
        if (!(len >= 1e-6f) || !isfinite(len)) continue;
        absScale[axis] = len;
        for (int k = 0; k < 3; k++) rot[axis][k] = col[k] / len;
    }
    float c01[3], c0[3] = {rot[0][0], rot[0][1], rot[0][2]}, c1[3] = {rot[1][0], rot[1][1], rot[1][2]},
                  c2[3] = {rot[2][0], rot[2][1], rot[2][2]};
    cross3(c0, c1, c01);
    float det = dot3(c01, c2);
    int candidates = det < 0.0f ? 3 : 1;
    float bestErr = INFINITY;
    for (int cand = 0; cand < candidates; cand++) {
        float r[4][4];
        memcpy(r, rot, sizeof(r));
        float sc[3] = {absScale[0], absScale[1], absScale[2]};
        if (det < 0.0f) {
            for (int k = 0; k < 3; k++) r[cand][k] = -r[cand][k];
            sc[cand] = -sc[cand];
        }
        float deg[3];
        eulerFromRotation(r, deg);
        float recomposed[4][4];
        VKRT_buildMeshTransformMatrix(outPosition, deg, sc, recomposed);
        float err = 0.0f;
        for (int c = 0; c < 4; c++)
            for (int row = 0; row < 3; row++) {
                float d = fabsf(recomposed[c][row] - worldTransform[c][row]);
                if (d > err) err = d;
            }
        if (err < bestErr) {
            bestErr = err;
            memcpy(outRotation, deg, sizeof(deg));
            memcpy(outScale, sc, sizeof(sc));
        }
    }
}

/* src/core/scene/transform.c:84-111: glTF nodes live in a +Y-up frame; the
 * engine is +Z-up. The import basis B is a +90 degree rotation about X (built
 * like any mesh transform: position 0, rotation (90, 0, 0), scale 1), and a
 * node transform M becomes B * M * B^-1, so that composing converted node
 * transforms equals converting the composed transform. */
void VKRT_buildImportedNodeTransform(float worldTransform[4][4], float outEngineTransform[4][4]) {
    if (!worldTransform || !outEngineTransform) return;
    static const float zero[3] = {0.0f, 0.0f, 0.0f};
    static const float basisRotation[3] = {90.0f, 0.0f, 0.0f};
    static const float one[3] = {1.0f, 1.0f, 1.0f};
    float basis[4][4], inverse[4][4], tmp[4][4];
    VKRT_buildMeshTransformMatrix(zero, basisRotation, one, basis);
    if (!mat4Inverse(basis, inverse)) mat4Identity(inverse);
    mat4Mul(basis, worldTransform, tmp);
    mat4Mul(tmp, inverse, outEngineTransform);
}

/* src/core/scene/transform.c:113-119 */
void VKRT_decomposeMeshNodeTransform(float worldTransform[4][4], float outPosition[3], float outRotation[3],
                                     float outScale[3]) {
    if (!outPosition || !outRotation || !outScale) return;
    float engine[4][4];
    mat4Identity(engine);
    VKRT_buildImportedNodeTransform(worldTransform, engine);
    VKRT_decomposeMeshTransform(engine, outPosition, outRotation, outScale);
}
