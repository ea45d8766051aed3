/*
 * mlp_oracle.c — FP64 reference of the far-field correction network.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP forward/backward kernels.
 *
 * Network (no reference counterpart; defined by this build): 12 inputs ->
 * 64 (SiLU) -> 64 (SiLU) -> 3, all layers affine. Parameters are one flat
 * array in this order: W1[64][12], b1[64], W2[64][64], b2[64], W3[3][64],
 * b3[3] (5,187 values). Loss: mean over the batch of the squared error summed
 * over the 3 outputs. Pinned by finite differences of mlp_loss_f64 (tests).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define IN 12
#define H 64
#define OUT 3
#define NPARAM (H * IN + H + H * H + H + OUT * H + OUT)

static double silu(double x) { return x / (1.0 + exp(-x)); }
static double dsilu(double x) {
    double s = 1.0 / (1.0 + exp(-x));
    return s * (1.0 + x * (1.0 - s));
}

int mlp_num_params(void) { return NPARAM; }

/* y[b][3] for x[b][12]; parameters in float (the trained master copy) */
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

/* loss = mean_b sum_j (y - t)^2 and its gradient with respect to every parameter */
double mlp_loss_grad_f64(const float *p, const float *x, const float *t, int batch, double *grad) {
    const float *W1 = p, *b1 = W1 + H * IN, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H, *b3 = W3 + OUT * H;
    double *gW1 = grad, *gb1 = gW1 + H * IN, *gW2 = gb1 + H, *gb2 = gW2 + H * H, *gW3 = gb2 + H, *gb3 = gW3 + OUT * H;
    memset(grad, 0, sizeof(double) * NPARAM);
    double loss = 0.0;
    for (int b = 0; b < batch; b++) {
        double z1[H], h1[H], z2[H], h2[H], dy[OUT], dh2[H], dz2[H], dh1[H], dz1[H];
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
            double e = s - t[b * OUT + j];
            loss += e * e;
            dy[j] = 2.0 * e / batch;
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
