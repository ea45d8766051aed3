/*
 * mlp_oracle.c — FP64 reference of the far-field correction network.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP forward/backward kernels.
 *
 * Network (no reference counterpart; defined by this build): 12 inputs ->
 * 64 (SiLU) -> 64 (SiLU) -> 6, all layers affine. Parameters are one flat
 * array in this order: W1[64][12], b1[64], W2[64][64], b2[64], W3[6][64],
 * b3[6] (5,382 values). The 6 outputs o are a residual quadrupole tensor
 * (xx, xy, xz, yy, yz, zz); a training row's target is (u, y, w) with u a
 * direction (3), y (3) and a weight w, the prediction B(u) o = -3 o u -
 * 1.5 tr(o) u + 7.5 (u.o.u) u (the softened quadrupole's shape for s = 1),
 * and the loss the mean over rows of w |B(u) o - y|^2. Pinned by finite differences of the loss
 * and by B(u) against a numerical gradient of the quadrupole potential
 * (tests/test_mlp_oracle.py).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define IN 12
#define H 64
#define OUT 6
#define TGT 7
#define NPARAM (H * IN + H + H * H + H + OUT * H + OUT)

static double silu(double x) { return x / (1.0 + exp(-x)); }
static double dsilu(double x) {
    double s = 1.0 / (1.0 + exp(-x));
    return s * (1.0 + x * (1.0 - s));
}

int mlp_num_params(void) { return NPARAM; }

/* the symmetric tensor o (xx xy xz yy yz zz) as a 3x3 matrix */
static void sym3(const double *o, double T[3][3]) {
    T[0][0] = o[0]; T[0][1] = T[1][0] = o[1]; T[0][2] = T[2][0] = o[2];
    T[1][1] = o[3]; T[1][2] = T[2][1] = o[4]; T[2][2] = o[5];
}

/* B(u) o */
void mlp_quad_basis_f64(const double *u, const double *o, double *p) {
    double T[3][3], Tu[3] = {0, 0, 0}, uTu = 0.0;
    sym3(o, T);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Tu[i] += T[i][j] * u[j];
    for (int i = 0; i < 3; i++) uTu += u[i] * Tu[i];
    double tr = T[0][0] + T[1][1] + T[2][2];
    for (int i = 0; i < 3; i++) p[i] = -3.0 * Tu[i] - 1.5 * tr * u[i] + 7.5 * uTu * u[i];
}

/* y[b][6] for x[b][12]; parameters in float (the trained master copy) */
void mlp_forward_f64(const float *p, const float *x, int batch, double *y) {
    const float *W1 = p, *b1 = W1 + H * IN, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H, *b3 = W3 + OUT * H;
    for (int b = 0; b < batch; b++) {
        double h1[H], h2[H];
        for (int j = 0; j < H; j++) {
            double s = b1[j];
            for (int k = 0; k < IN; k++) s += (double)W1[j * IN + k] * x[b * IN + k];
            h1[j] = silu(s);
        }
        for (int j = 0; j < H; j++) {
            double s = b2[j];
            for (int k = 0; k < H; k++) s += (double)W2[j * H + k] * h1[k];
            h2[j] = silu(s);
        }
        for (int j = 0; j < OUT; j++) {
            double s = b3[j];
            for (int k = 0; k < H; k++) s += (double)W3[j * H + k] * h2[k];
            y[b * OUT + j] = s;
        }
    }
}

/* loss = mean_b w_b |B(u_b) o_b - y_b|^2 for t[b] = (u_b, y_b, w_b), and its gradient
 * with respect to every parameter (the derivative of B(u) o in o by finite
 * basis vectors: B is linear in o) */
double mlp_loss_grad_f64(const float *p, const float *x, const float *t, int batch, double *grad) {
    const float *W1 = p, *b1 = W1 + H * IN, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H, *b3 = W3 + OUT * H;
    double *gW1 = grad, *gb1 = gW1 + H * IN, *gW2 = gb1 + H, *gb2 = gW2 + H * H, *gW3 = gb2 + H, *gb3 = gW3 + OUT * H;
    memset(grad, 0, sizeof(double) * NPARAM);
    double loss = 0.0;
    for (int b = 0; b < batch; b++) {
        double z1[H], h1[H], z2[H], h2[H], o[OUT], dy[OUT], dh2[H], dz2[H], dh1[H], dz1[H];
        for (int j = 0; j < H; j++) {
            double s = b1[j];
            for (int k = 0; k < IN; k++) s += (double)W1[j * IN + k] * x[b * IN + k];
            z1[j] = s; h1[j] = silu(s);
        }
        for (int j = 0; j < H; j++) {
            double s = b2[j];
            for (int k = 0; k < H; k++) s += (double)W2[j * H + k] * h1[k];
            z2[j] = s; h2[j] = silu(s);
        }
        for (int j = 0; j < OUT; j++) {
            double s = b3[j];
            for (int k = 0; k < H; k++) s += (double)W3[j * H + k] * h2[k];
            o[j] = s;
        }
        double u[3] = {t[b * TGT + 0], t[b * TGT + 1], t[b * TGT + 2]}, pr[3], e[3], w = t[b * TGT + 6];
        mlp_quad_basis_f64(u, o, pr);
        for (int i = 0; i < 3; i++) {
            e[i] = pr[i] - t[b * TGT + 3 + i];
            loss += w * e[i] * e[i];
        }
        for (int c = 0; c < OUT; c++) {
            double unit[OUT] = {0, 0, 0, 0, 0, 0}, bc[3];
            unit[c] = 1.0;
            mlp_quad_basis_f64(u, unit, bc);
            dy[c] = 2.0 * w * (e[0] * bc[0] + e[1] * bc[1] + e[2] * bc[2]) / batch;
        }
        for (int k = 0; k < H; k++) dh2[k] = 0.0;
        for (int j = 0; j < OUT; j++) {
            gb3[j] += dy[j];
            for (int k = 0; k < H; k++) { gW3[j * H + k] += dy[j] * h2[k]; dh2[k] += dy[j] * W3[j * H + k]; }
        }
        for (int k = 0; k < H; k++) { dz2[k] = dh2[k] * dsilu(z2[k]); dh1[k] = 0.0; }
        for (int j = 0; j < H; j++) {
            gb2[j] += dz2[j];
            for (int k = 0; k < H; k++) { gW2[j * H + k] += dz2[j] * h1[k]; dh1[k] += dz2[j] * W2[j * H + k]; }
        }
        for (int k = 0; k < H; k++) dz1[k] = dh1[k] * dsilu(z1[k]);
        for (int j = 0; j < H; j++) {
            gb1[j] += dz1[j];
            for (int k = 0; k < IN; k++) gW1[j * IN + k] += dz1[j] * x[b * IN + k];
        }
    }
    return loss / batch;
}
