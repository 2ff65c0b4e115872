// Shared definitions of the learned-proposal network (gf_mlp.hip, gf_capi.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

// 2 -> H -> H -> 1, SiLU.  Flat FP32 parameter vector:
//   W1[k][0..1] (k < H), b1[k], W2[j][k] (row j = output neuron), b2[j], w3[j], b3
#define GF_MLP_H 128
#define GF_MLP_W1 0
#define GF_MLP_B1 (2 * GF_MLP_H)
#define GF_MLP_W2 (3 * GF_MLP_H)
#define GF_MLP_B2 (GF_MLP_W2 + GF_MLP_H * GF_MLP_H)
#define GF_MLP_W3 (GF_MLP_B2 + GF_MLP_H)
#define GF_MLP_B3 (GF_MLP_W3 + GF_MLP_H)
#define GF_MLP_NPARAM (GF_MLP_B3 + 1)

// per-block partial gradient layouts
#define GF_MLP_PART_IN (3 * GF_MLP_H)       // dW1 (interleaved) | db1
#define GF_MLP_PART_OUT (2 * GF_MLP_H + 1)  // dw3 | db2 | db3
#define GF_MLP_W2_BLOCKS 64                  // workgroups of the dW2 product

// Row of the 16 x 16 result of v_mfma_f32_16x16x32_bf16 in register r of lane l
// (the column is l % 16): four consecutive rows per group of 16 lanes.
#define GF_BF16_D_ROW(l, r) (4 * ((l) >> 4) + (r))

struct GfBox {
    double l12, u12, l13, u13;
};

// Per-step training state kept on the device, so that a captured training step
// can be replayed: k_step_begin advances it (batch key, learning rate, Adam bias
// corrections) before the step's kernels read it.
struct GfTrainState {
    uint64_t seed;
    int32_t k, steps;      // step within this training call, number of steps
    int32_t adam_t, pad;   // Adam step count over the network's life
    float lr0, b1, b2, eps;
    float lr, c1, c2, pad2;
    uint64_t key;
};

// Device buffers of one network (sized for batches of up to cap samples).
struct GfMlpDev {
    float *P, *m, *v, *grad;   // parameters, Adam moments, gradient
    uint16_t *w2b, *w2t;       // bf16 W2 and W2^T
    float2 *xy;                // network inputs (x = 2u - 1, y = 2v - 1)
    float *z1, *z2;            // pre-activations [cap][H]
    uint16_t *h1, *delta2;     // bf16 activations / layer-2 deltas [cap][H]
    float *out, *coef, *loss;  // f per sample, dL/df per sample, loss
    float *part_in, *part_w2, *part_out;
    GfTrainState *state;
    int64_t cap;
};

namespace gfm {
// one training step, all parameters from the device state
int launch_train_step(const GfMlpDev &D, const double *m12, const double *m13, int64_t ndata, GfBox box, int batch,
                      hipStream_t s);
int launch_batch_grid(int G, float2 *xy, hipStream_t s);
// n a multiple of 64, n <= cap
int launch_forward(const GfMlpDev &D, const float2 *xy, int n, float *out, hipStream_t s);
int launch_loss(const float *out, int nd, int nu, float *coef, float *loss, hipStream_t s);
// gradient of sum_s coef[s] f(x_s) into grad[GF_MLP_NPARAM] (after launch_forward on the same xy)
int launch_backward(const GfMlpDev &D, const float2 *xy, const float *coef, int n, float *grad, hipStream_t s);

int launch_w2_bf16(const GfMlpDev &D, hipStream_t s);
int launch_sample(const double *cdf, const double *pcell, int G, GfBox box, int64_t n, uint64_t key, double *m12,
                  double *m13, double *w, hipStream_t s);
int launch_sample_uniform(GfBox box, int64_t n, uint64_t key, double *m12, double *m13, double *w, hipStream_t s);
int launch_bf16_probe(const uint16_t *a8, const uint16_t *b8, float *d, hipStream_t s);
}  // namespace gfm
