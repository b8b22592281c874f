/*
 * Generator for the sRGB -> spectrum coefficient table used by spectral mode.
 *
 * The reference embeds a table produced offline by rgb2spec_opt
 * (external/rgb2spec/README.md) whose data file is not available; this tool
 * regenerates one with the published method of Jakob & Hanika 2019 ("A
 * Low-Dimensional Function Space for Efficient Spectral Upsampling"):
 * reflectance s(l) = sigmoid(c0 l^2 + c1 l + c2), sigmoid(x) = 0.5 + x/(2 sqrt(1+x^2)),
 * coefficients fitted per table cell by damped Gauss-Newton, warm-started
 * along each brightness column.
 *
 * Self-consistency with the renderer: the forward model is exactly the
 * integrator's own colour pipeline (src/shaders/utility/spectral.slang:66-118):
 * the Wyman-Sloan-Shirley multi-lobe CIE 1931 fit over 360..830 nm under an
 * equal-energy illuminant, normalised by 106.9461715, Bradford-adapted from E
 * to D65 and mapped to linear sRGB.
 *
 * Output layout is what src/core/scene/rgb2spec.c:17-59 parses and
 * src/shaders/utility/rgb2spec.slang:33-80 indexes: "SPEC", u32 res, res
 * brightness scales (smoothstep(smoothstep(k/(res-1)))), then
 * 3*res*res*res*3 floats ordered [dominant][z][y][x][coefficient] with
 * coefficients for the polynomial in nanometres.
 *
 * usage: rgb2spec_opt <res> <output.coeff>
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define LAMBDA_MIN 360.0
#define LAMBDA_MAX 830.0
#define NSAMPLES 471 /* 1 nm spacing, trapezoidal weights */
#define Y_INTEGRAL 106.9461715

static double g_weight[3][NSAMPLES]; /* CMF * dl / Y_INTEGRAL, already mapped to linear sRGB rows */
static double g_lambdaNorm[NSAMPLES];

static double lobe(double x, double mu, double s1, double s2) {
    double t = (x - mu) * (x < mu ? s1 : s2);
    return exp(-0.5 * t * t);
}

static void cmfFit(double l, double out[3]) {
    double x = 0.362 * lobe(l, 442.0, 0.0624, 0.0374) + 1.056 * lobe(l, 599.8, 0.0264, 0.0323) -
               0.065 * lobe(l, 501.1, 0.0490, 0.0382);
    double y = 0.821 * lobe(l, 568.8, 0.0213, 0.0247) + 0.286 * lobe(l, 530.9, 0.0613, 0.0322);
    double z = 1.217 * lobe(l, 437.0, 0.0845, 0.0278) + 0.681 * lobe(l, 459.0, 0.0385, 0.0725);
    out[0] = x > 0.0 ? x : 0.0;
    out[1] = y > 0.0 ? y : 0.0;
    out[2] = z > 0.0 ? z : 0.0;
}

static void xyzToSrgbMatrix(double m[3][3]) {
    static const double srgb[3][3] = {
        {3.2404542, -1.5371385, -0.4985314}, {-0.9692660, 1.8760108, 0.0415560}, {0.0556434, -0.2040259, 1.0572252}};
    static const double brad[3][3] = {{0.8951, 0.2664, -0.1614}, {-0.7502, 1.7135, 0.0367}, {0.0389, -0.0685, 1.0296}};
    static const double bradInv[3][3] = {
        {0.9869929, -0.1470543, 0.1599627}, {0.4323053, 0.5183603, 0.0492912}, {-0.0085287, 0.0400428, 0.9684867}};
    static const double scale[3] = {0.9413344, 1.0404175, 1.0895327};
    double a[3][3], b[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a[i][j] = bradInv[i][0] * scale[0] * brad[0][j] + bradInv[i][1] * scale[1] * brad[1][j] +
                                              bradInv[i][2] * scale[2] * brad[2][j];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) b[i][j] = srgb[i][0] * a[0][j] + srgb[i][1] * a[1][j] + srgb[i][2] * a[2][j];
    memcpy(m, b, sizeof(b));
}

static void initTables(void) {
    double m[3][3];
    xyzToSrgbMatrix(m);
    double dl = (LAMBDA_MAX - LAMBDA_MIN) / (NSAMPLES - 1);
    for (int i = 0; i < NSAMPLES; i++) {
        double l = LAMBDA_MIN + dl * i;
        double w = (i == 0 || i == NSAMPLES - 1) ? 0.5 * dl : dl;
        double xyz[3];
        cmfFit(l, xyz);
        for (int r = 0; r < 3; r++) g_weight[r][i] = (m[r][0] * xyz[0] + m[r][1] * xyz[1] + m[r][2] * xyz[2]) * w / Y_INTEGRAL;
        g_lambdaNorm[i] = (l - LAMBDA_MIN) / (LAMBDA_MAX - LAMBDA_MIN);
    }
}

/* rgb of the sigmoid spectrum with normalised-wavelength coefficients c, plus d rgb / d c */
static void evalModel(const double c[3], double rgb[3], double jac[3][3]) {
    for (int r = 0; r < 3; r++) {
        rgb[r] = 0.0;
        for (int k = 0; k < 3; k++) jac[r][k] = 0.0;
    }
    for (int i = 0; i < NSAMPLES; i++) {
        double t = g_lambdaNorm[i];
        double x = (c[0] * t + c[1]) * t + c[2];
        double q = 1.0 / sqrt(1.0 + x * x);
        double s = 0.5 * x * q + 0.5;
        double ds = 0.5 * q * q * q;
        double dx[3] = {t * t, t, 1.0};
        for (int r = 0; r < 3; r++) {
            double w = g_weight[r][i];
            rgb[r] += s * w;
            for (int k = 0; k < 3; k++) jac[r][k] += ds * dx[k] * w;
        }
    }
}

static int solve3(double a[3][3], const double b[3], double x[3]) {
    double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                 a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    if (fabs(det) < 1e-300) return 0;
    for (int k = 0; k < 3; k++) {
        double m[3][3];
        memcpy(m, a, sizeof(m));
        for (int r = 0; r < 3; r++) m[r][k] = b[r];
        x[k] = (m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0])) /
               det;
    }
    return 1;
}

/* Damped Gauss-Newton on ||rgb(c) - target||^2, c updated in place. */
static double fitCell(const double target[3], double c[3]) {
    double lambda = 1e-3;
    double rgb[3], jac[3][3];
    evalModel(c, rgb, jac);
    double err = 0.0;
    for (int r = 0; r < 3; r++) err += (rgb[r] - target[r]) * (rgb[r] - target[r]);
    for (int it = 0; it < 200 && err > 1e-14; it++) {
        double jtj[3][3], jtr[3], delta[3];
        for (int i = 0; i < 3; i++) {
            jtr[i] = 0.0;
            for (int j = 0; j < 3; j++) {
                jtj[i][j] = 0.0;
                for (int r = 0; r < 3; r++) jtj[i][j] += jac[r][i] * jac[r][j];
            }
            for (int r = 0; r < 3; r++) jtr[i] += jac[r][i] * (target[r] - rgb[r]);
        }
        int improved = 0;
        for (int tries = 0; tries < 12; tries++) {
            double a[3][3];
            memcpy(a, jtj, sizeof(a));
            for (int i = 0; i < 3; i++) a[i][i] += lambda * (jtj[i][i] + 1e-12);
            if (!solve3(a, jtr, delta)) {
                lambda *= 10.0;
                continue;
            }
            double cn[3] = {c[0] + delta[0], c[1] + delta[1], c[2] + delta[2]};
            double rgbn[3], jacn[3][3];
            evalModel(cn, rgbn, jacn);
            double errn = 0.0;
            for (int r = 0; r < 3; r++) errn += (rgbn[r] - target[r]) * (rgbn[r] - target[r]);
            if (errn < err) {
                memcpy(c, cn, sizeof(cn));
                memcpy(rgb, rgbn, sizeof(rgbn));
                memcpy(jac, jacn, sizeof(jacn));
                err = errn;
                lambda = lambda * 0.3 > 1e-9 ? lambda * 0.3 : 1e-9;
                improved = 1;
                break;
            }
            lambda *= 10.0;
        }
        if (!improved) break;
        double step = fabs(delta[0]) + fabs(delta[1]) + fabs(delta[2]);
        if (step < 1e-12) break;
    }
    return err;
}

static double smoothstep(double x) { return x * x * (3.0 - 2.0 * x); }

int main(int argc, char** argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s <res> <output.coeff>\n", argv[0]);
        return 2;
    }
    int res = atoi(argv[1]);
    if (res < 2 || res > 256) {
        fprintf(stderr, "res must be in [2, 256]\n");
        return 2;
    }
    initTables();
    size_t nScale = (size_t)res;
    size_t nData = (size_t)3 * res * res * res * 3;
    float* scales = (float*)malloc(sizeof(float) * nScale);
    float* data = (float*)calloc(nData, sizeof(float));
    if (!scales || !data) return 1;
    for (int k = 0; k < res; k++) scales[k] = (float)smoothstep(smoothstep((double)k / (double)(res - 1)));

    int start = res / 5;
    double worst = 0.0;
#pragma omp parallel for collapse(3) schedule(dynamic) reduction(max : worst)
    for (int l = 0; l < 3; l++) {
        for (int yi = 0; yi < res; yi++) {
            for (int xi = 0; xi < res; xi++) {
                double x = (double)xi / (double)(res - 1);
                double y = (double)yi / (double)(res - 1);
                double c[3] = {0.0, 0.0, 0.0};
                double cStart[3];
                /* sweep up from the start brightness, then down with a fresh warm start */
                for (int pass = 0; pass < 2; pass++) {
                    int zBegin = pass == 0 ? start : start - 1;
                    int zEnd = pass == 0 ? res : -1;
                    int zStep = pass == 0 ? 1 : -1;
                    if (pass == 1) memcpy(c, cStart, sizeof(c));
                    for (int zi = zBegin; zi != zEnd; zi += zStep) {
                        double z = scales[zi];
                        double target[3];
                        target[l] = z;
                        target[(l + 1) % 3] = x * z;
                        target[(l + 2) % 3] = y * z;
                        double err = fitCell(target, c);
                        if (zi == start) memcpy(cStart, c, sizeof(c));
                        if (err > worst) worst = err;
                        /* normalised-wavelength coefficients -> nanometre coefficients */
                        double a = LAMBDA_MIN, d = LAMBDA_MAX - LAMBDA_MIN;
                        double A = c[0] / (d * d);
                        double B = c[1] / d - 2.0 * c[0] * a / (d * d);
                        double C = c[0] * a * a / (d * d) - c[1] * a / d + c[2];
                        size_t off = ((((size_t)l * res + zi) * res + yi) * res + xi) * 3;
                        data[off + 0] = (float)A;
                        data[off + 1] = (float)B;
                        data[off + 2] = (float)C;
                    }
                }
            }
        }
    }
    FILE* f = fopen(argv[2], "wb");
    if (!f) {
        fprintf(stderr, "cannot open %s\n", argv[2]);
        return 1;
    }
    uint32_t r32 = (uint32_t)res;
    int ok = fwrite("SPEC", 1, 4, f) == 4 && fwrite(&r32, 4, 1, f) == 1 && fwrite(scales, sizeof(float), nScale, f) == nScale &&
             fwrite(data, sizeof(float), nData, f) == nData;
    fclose(f);
    free(scales);
    free(data);
    if (!ok) return 1;
    fprintf(stderr, "rgb2spec_opt: res %d written to %s (worst squared rgb residual %.3e)\n", res, argv[2], worst);
    return 0;
}
