// Learned importance-sampling proposal for the Dalitz normalisation integrals.
//
// A 2 -> 128 -> 128 -> 1 SiLU network f(u, v) over the unit square (the Dalitz
// box [l12, u12] x [l13, u13] mapped to [0, 1]^2) is trained by maximum
// likelihood on events the model generated: loss = -mean_data f + log mean_unif
// exp(f), the second term a Monte-Carlo estimate of the normaliser over the box.
// The 128 x 128 layer and both of its backward products run on BF16 MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation); the rest is FP32 VALU.
//
// Sampling: exp(f) tabulated on a G x G grid of cell centres, mixed with the
// uniform density (weight alpha, so every cell has probability >= alpha/G^2),
// sampled by inverse CDF (cell) and uniformly inside the cell.  The proposal's
// density is exactly that piecewise-constant table, so the weights
// w = 1 / (N q(x)) make sum_p w_p F(x_p) an unbiased estimate of the integral of
// F over the box whatever the network learned.
//
// No reference counterpart: GooFit normalises Amp3Body on a fixed midpoint grid
// (src/PDFs/physics/detail/SpecialResonanceIntegrator.cu:52-121).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "gf_device.h"
#include "gf_mlp.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// Row of the 16 x 16 MFMA result held by register r of lane l (the column is
// l % 16).  gfx950 result layout of v_mfma_f32_16x16x32_bf16, checked by
// gf_mlp_layout_check against a host product.
__device__ __forceinline__ int d_row(int l, int r) { return GF_BF16_D_ROW(l, r); }

// fp32 -> bf16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float x) {
    uint32_t b = __float_as_uint(x);
    b += 0x7FFFu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}

using gfd::mix64;
__device__ __forceinline__ double unif(uint64_t key, uint64_t ctr) { return gfd::unif01(key, ctr); }

__device__ __forceinline__ bf16x8 load8(const uint16_t *p) {
    const uint4 q = *reinterpret_cast<const uint4 *>(p);  // 16-byte aligned
    bf16x8 v;
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for(int t = 0; t < 4; ++t) {
        v[2 * t]     = __builtin_bit_cast(__bf16, (uint16_t)(w[t] & 0xFFFFu));
        v[2 * t + 1] = __builtin_bit_cast(__bf16, (uint16_t)(w[t] >> 16));
    }
    return v;
}

// ----------------------------------------------------------------- batches

// Training batch: rows [0, nd) are data events drawn with replacement, rows
// [nd, nd + nu) uniform points of the unit square.  Inputs to the network are
// x = 2u - 1, y = 2v - 1.
__global__ __launch_bounds__(256) void k_batch_train(const double *__restrict__ m12, const double *__restrict__ m13,
                                                     int64_t ndata, GfBox box, int nd, int nu,
                                                     const GfTrainState *__restrict__ st, float2 *__restrict__ xy) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if(s >= nd + nu)
        return;
    const uint64_t key = st->key;
    double u, v;
    if(s < nd) {
        const int64_t e = (int64_t)(mix64(key ^ mix64(2 * (uint64_t)s)) % (uint64_t)ndata);
        u = (m12[e] - box.l12) / (box.u12 - box.l12);
        v = (m13[e] - box.l13) / (box.u13 - box.l13);
    } else {
        u = unif(key + 1, 2 * (uint64_t)s);
        v = unif(key + 1, 2 * (uint64_t)s + 1);
    }
    xy[s] = make_float2((float)(2.0 * u - 1.0), (float)(2.0 * v - 1.0));
}

// Cell centres of a G x G grid over the unit square, cell c = cx + G * cy.
__global__ __launch_bounds__(256) void k_batch_grid(int G, int64_t ncell, float2 *__restrict__ xy) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(c >= ncell)
        return;
    const double u = ((double)(c % G) + 0.5) / G, v = ((double)(c / G) + 0.5) / G;
    xy[c] = make_float2((float)(2.0 * u - 1.0), (float)(2.0 * v - 1.0));
}

// ----------------------------------------------------------------- forward

// Layer 1 (2 -> 128, SiLU): one lane per (sample, neuron).  z1 kept in FP32 for
// the backward pass, h1 stored as bf16 for the MFMA layer.
__global__ __launch_bounds__(256) void k_layer1(const float2 *__restrict__ xy, int n, const float *__restrict__ P,
                                                float *__restrict__ z1, uint16_t *__restrict__ h1) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(id >= (int64_t)n * GF_MLP_H)
        return;
    const int s = (int)(id / GF_MLP_H), k = (int)(id % GF_MLP_H);
    const float2 x = xy[s];
    const float z  = P[GF_MLP_W1 + 2 * k] * x.x + P[GF_MLP_W1 + 2 * k + 1] * x.y + P[GF_MLP_B1 + k];
    z1[id]         = z;
    h1[id]         = f2bf(silu(z));
}

// Layer 2 (128 -> 128, SiLU, BF16 MFMA) and the output layer (128 -> 1).
// A workgroup of 4 waves takes 64 samples, wave w the 16 rows [16w, 16w + 16):
// for each of the 8 neuron tiles, 4 MFMAs over K = 128.  n is a multiple of 64.
__global__ __launch_bounds__(256) void k_layer2(const uint16_t *__restrict__ h1, const uint16_t *__restrict__ w2b,
                                                const float *__restrict__ P, int n, float *__restrict__ z2,
                                                float *__restrict__ out) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int row0 = blockIdx.x * 64 + 16 * w;
    if(row0 >= n)
        return;
    const int li = l & 15, lg = l >> 4;
    bf16x8 a[4];
#pragma unroll
    for(int c = 0; c < 4; ++c)
        a[c] = load8(h1 + (int64_t)(row0 + li) * GF_MLP_H + 32 * c + 8 * lg);
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for(int jt = 0; jt < GF_MLP_H / 16; ++jt) {
        const int j = 16 * jt + li;  // this lane's column (neuron)
        floatx4 d  = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for(int c = 0; c < 4; ++c) {
            const bf16x8 b = load8(w2b + (int64_t)j * GF_MLP_H + 32 * c + 8 * lg);  // B[k][j] = W2[j][k]
            d              = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[c], b, d, 0, 0, 0);
        }
        const float bj = P[GF_MLP_B2 + j], wj = P[GF_MLP_W3 + j];
#pragma unroll
        for(int r = 0; r < 4; ++r) {
            const int s   = row0 + d_row(l, r);
            const float z = d[r] + bj;
            z2[(int64_t)s * GF_MLP_H + j] = z;
            acc[r] += wj * silu(z);
        }
    }
    // sum over the 16 columns held by lanes with the same l / 16
#pragma unroll
    for(int r = 0; r < 4; ++r) {
#pragma unroll
        for(int off = 1; off < 16; off <<= 1)
            acc[r] += __shfl_xor(acc[r], off, 64);
    }
    if(li == 0) {
        const float b3 = P[GF_MLP_B3];
#pragma unroll
        for(int r = 0; r < 4; ++r)
            out[row0 + d_row(l, r)] = acc[r] + b3;
    }
}

// ----------------------------------------------------------------- loss

// One workgroup: per-sample output gradients of the loss
//   L = -(1/nd) sum_data f + log((1/nu) sum_unif exp f)
// (data: -1/nd; uniform: softmax weight), and L itself.
__global__ __launch_bounds__(1024) void k_loss(const float *__restrict__ out, int nd, int nu, float *__restrict__ coef,
                                               float *__restrict__ loss) {
    __shared__ float red[32];
    __shared__ float bcast;
    const int t = threadIdx.x, nw = blockDim.x >> 6, wv = t >> 6, ln = t & 63;
    // max over the uniform part
    float m = -INFINITY;
    for(int s = nd + t; s < nd + nu; s += blockDim.x)
        m = fmaxf(m, out[s]);
    for(int off = 32; off > 0; off >>= 1)
        m = fmaxf(m, __shfl_xor(m, off, 64));
    if(ln == 0)
        red[wv] = m;
    __syncthreads();
    if(t == 0) {
        float mm = red[0];
        for(int i = 1; i < nw; ++i)
            mm = fmaxf(mm, red[i]);
        bcast = mm;
    }
    __syncthreads();
    m = bcast;
    __syncthreads();
    // sum of exp over the uniform part, and of f over the data part
    double se = 0.0, sd = 0.0;
    for(int s = t; s < nd + nu; s += blockDim.x) {
        if(s < nd)
            sd += out[s];
        else
            se += exp((double)(out[s] - m));
    }
    for(int off = 32; off > 0; off >>= 1) {
        se += __shfl_xor(se, off, 64);
        sd += __shfl_xor(sd, off, 64);
    }
    __shared__ double rse[16], rsd[16];
    __shared__ double zb;
    if(ln == 0) {
        rse[wv] = se;
        rsd[wv] = sd;
    }
    __syncthreads();
    if(t == 0) {
        double a = 0, b = 0;
        for(int i = 0; i < nw; ++i) {
            a += rse[i];
            b += rsd[i];
        }
        zb    = a;
        *loss = (float)(-b / nd + (double)m + log(a / nu));
    }
    __syncthreads();
    for(int s = t; s < nd + nu; s += blockDim.x)
        coef[s] = s < nd ? -1.0f / nd : (float)(exp((double)(out[s] - m)) / zb);
}

// ----------------------------------------------------------------- backward

// delta2 = coef * w3 * silu'(z2) (stored as bf16) and this block's partial sums
// over its 64 samples of dL/dw3, dL/db2 and dL/db3.  Lane j of each of the
// GF_BO_SPLIT lane groups takes neuron j on a contiguous quarter of the samples;
// the groups' sums are added in group order.
#define GF_BO_SPLIT 4
__global__ __launch_bounds__(GF_MLP_H *GF_BO_SPLIT) void k_back_out(const float *__restrict__ z2,
                                                                   const float *__restrict__ coef,
                                                                   const float *__restrict__ P, int n,
                                                                   uint16_t *__restrict__ delta2,
                                                                   float *__restrict__ part) {
    __shared__ float red[GF_BO_SPLIT][2 * GF_MLP_H + 1];
    const int j  = threadIdx.x % GF_MLP_H, g = threadIdx.x / GF_MLP_H;
    const int s0 = blockIdx.x * 64 + g * (64 / GF_BO_SPLIT);
    const float wj = P[GF_MLP_W3 + j];
    float gw3 = 0.f, gb2 = 0.f, gb3 = 0.f;
#pragma unroll 4
    for(int s = s0; s < s0 + 64 / GF_BO_SPLIT; ++s) {
        if(s < n) {
            const float z  = z2[(int64_t)s * GF_MLP_H + j];
            const float cs = coef[s];
            const float d  = cs * wj * dsilu(z);
            delta2[(int64_t)s * GF_MLP_H + j] = f2bf(d);
            gw3 += cs * silu(z);
            gb2 += d;
            gb3 += cs;
        }
    }
    red[g][j]            = gw3;
    red[g][GF_MLP_H + j] = gb2;
    if(j == 0)
        red[g][2 * GF_MLP_H] = gb3;
    __syncthreads();
    if(g == 0) {
        float a = red[0][j], b = red[0][GF_MLP_H + j], c = red[0][2 * GF_MLP_H];
        for(int q = 1; q < GF_BO_SPLIT; ++q) {
            a += red[q][j];
            b += red[q][GF_MLP_H + j];
            c += red[q][2 * GF_MLP_H];
        }
        float *p = part + (int64_t)blockIdx.x * GF_MLP_PART_OUT;
        p[j]            = a;
        p[GF_MLP_H + j] = b;
        if(j == 0)
            p[2 * GF_MLP_H] = c;
    }
}

// dW2[j][k] = sum_s delta2[s][j] h1[s][k] over this block's samples (BF16 MFMA,
// K = samples).  Chunks of 32 samples are staged in LDS; wave w owns the 16
// output tiles (j-tile, k-tile) with j-tile in {2w, 2w + 1}.  Writes the
// block's 128 x 128 partial.
__global__ __launch_bounds__(256) void k_back_w2(const uint16_t *__restrict__ delta2, const uint16_t *__restrict__ h1,
                                                 int n, int chunks_per_block, float *__restrict__ part) {
    __shared__ uint16_t sd[32][GF_MLP_H + 8];
    __shared__ uint16_t sh[32][GF_MLP_H + 8];
    const int t = threadIdx.x, l = t & 63, w = t >> 6, li = l & 15, lg = l >> 4;
    floatx4 acc[2][8];
#pragma unroll
    for(int a = 0; a < 2; ++a)
#pragma unroll
        for(int b = 0; b < 8; ++b)
            acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
    for(int ch = 0; ch < chunks_per_block; ++ch) {
        const int s0 = (blockIdx.x * chunks_per_block + ch) * 32;
        if(s0 >= n)
            break;
        __syncthreads();
        // stage 32 x 128 of each operand: 256 lanes x 16 elements
        for(int e = t; e < 32 * GF_MLP_H / 8; e += 256) {
            const int r = e / (GF_MLP_H / 8), c8 = (e % (GF_MLP_H / 8)) * 8;
            const uint4 qd = *reinterpret_cast<const uint4 *>(delta2 + (int64_t)(s0 + r) * GF_MLP_H + c8);
            const uint4 qh = *reinterpret_cast<const uint4 *>(h1 + (int64_t)(s0 + r) * GF_MLP_H + c8);
            const uint32_t vd[4] = {qd.x, qd.y, qd.z, qd.w}, vh[4] = {qh.x, qh.y, qh.z, qh.w};
#pragma unroll
            for(int q = 0; q < 4; ++q) {
                sd[r][c8 + 2 * q]     = (uint16_t)(vd[q] & 0xFFFFu);
                sd[r][c8 + 2 * q + 1] = (uint16_t)(vd[q] >> 16);
                sh[r][c8 + 2 * q]     = (uint16_t)(vh[q] & 0xFFFFu);
                sh[r][c8 + 2 * q + 1] = (uint16_t)(vh[q] >> 16);
            }
        }
        __syncthreads();
        // A[i = j][kk = s] = delta2[s][j]; B[kk = s][jj = k] = h1[s][k]
#pragma unroll
        for(int a = 0; a < 2; ++a) {
            const int j = 16 * (2 * w + a) + li;
            bf16x8 av;
#pragma unroll
            for(int q = 0; q < 8; ++q)
                av[q] = __builtin_bit_cast(__bf16, sd[8 * lg + q][j]);
#pragma unroll
            for(int b = 0; b < 8; ++b) {
                const int k = 16 * b + li;
                bf16x8 bv;
#pragma unroll
                for(int q = 0; q < 8; ++q)
                    bv[q] = __builtin_bit_cast(__bf16, sh[8 * lg + q][k]);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[a][b], 0, 0, 0);
            }
        }
    }
    float *p = part + (int64_t)blockIdx.x * GF_MLP_H * GF_MLP_H;
#pragma unroll
    for(int a = 0; a < 2; ++a)
#pragma unroll
        for(int b = 0; b < 8; ++b)
#pragma unroll
            for(int r = 0; r < 4; ++r) {
                const int j = 16 * (2 * w + a) + d_row(l, r);
                const int k = 16 * b + li;
                p[j * GF_MLP_H + k] = acc[a][b][r];
            }
}

// dh1[s][k] = sum_j delta2[s][j] W2[j][k] (BF16 MFMA with the transposed W2
// copy), dz1 = dh1 * silu'(z1), and this block's partial sums over its 64
// samples of dL/dW1[k][0..1] and dL/db1[k].  Wave w takes rows [16w, 16w + 16).
__global__ __launch_bounds__(256) void k_back_h1(const uint16_t *__restrict__ delta2, const uint16_t *__restrict__ w2t,
                                                 const float *__restrict__ z1, const float2 *__restrict__ xy, int n,
                                                 float *__restrict__ part) {
    __shared__ float red[4][3][GF_MLP_H];
    const int t = threadIdx.x, l = t & 63, w = t >> 6, li = l & 15, lg = l >> 4;
    const int row0 = blockIdx.x * 64 + 16 * w;
    bf16x8 a[4];
#pragma unroll
    for(int c = 0; c < 4; ++c)
        a[c] = load8(delta2 + (int64_t)(row0 + li) * GF_MLP_H + 32 * c + 8 * lg);
    float2 x[4];
#pragma unroll
    for(int r = 0; r < 4; ++r)
        x[r] = xy[row0 + d_row(l, r)];
#pragma unroll
    for(int kt = 0; kt < GF_MLP_H / 16; ++kt) {
        const int k = 16 * kt + li;
        floatx4 d  = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for(int c = 0; c < 4; ++c) {
            const bf16x8 b = load8(w2t + (int64_t)k * GF_MLP_H + 32 * c + 8 * lg);  // B[j][k] = W2t[k][j]
            d              = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[c], b, d, 0, 0, 0);
        }
        float g0 = 0.f, g1 = 0.f, gb = 0.f;
#pragma unroll
        for(int r = 0; r < 4; ++r) {
            const int s    = row0 + d_row(l, r);
            const float dz = d[r] * dsilu(z1[(int64_t)s * GF_MLP_H + k]);
            g0 += dz * x[r].x;
            g1 += dz * x[r].y;
            gb += dz;
        }
        // lanes l, l+16, l+32, l+48 share column k
        g0 += __shfl_xor(g0, 16, 64);
        g1 += __shfl_xor(g1, 16, 64);
        gb += __shfl_xor(gb, 16, 64);
        g0 += __shfl_xor(g0, 32, 64);
        g1 += __shfl_xor(g1, 32, 64);
        gb += __shfl_xor(gb, 32, 64);
        if(lg == 0) {
            red[w][0][k] = g0;
            red[w][1][k] = g1;
            red[w][2][k] = gb;
        }
    }
    __syncthreads();
    if(t < GF_MLP_H) {
        float *p = part + (int64_t)blockIdx.x * GF_MLP_PART_IN;
        float s0 = 0.f, s1 = 0.f, s2 = 0.f;
        for(int q = 0; q < 4; ++q) {
            s0 += red[q][0][t];
            s1 += red[q][1][t];
            s2 += red[q][2][t];
        }
        p[2 * t]            = s0;
        p[2 * t + 1]        = s1;
        p[2 * GF_MLP_H + t] = s2;
    }
}

// Gradient vector: sum of the per-block partials.  A workgroup takes 64
// parameters; its 4 waves sum the partials b = w, w + 4, w + 8, ... (two
// interleaved accumulators each, so consecutive loads are independent), then
// the four sums are added in wave order: a fixed order for a given launch.
__global__ __launch_bounds__(256) void k_grad_sum(const float *__restrict__ p_in, int nb_in,
                                                  const float *__restrict__ p_w2, int nb_w2,
                                                  const float *__restrict__ p_out, int nb_out, float *__restrict__ g) {
    __shared__ float red[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i    = blockIdx.x * 64 + lane;
    const float *base = p_in;
    int64_t stride    = GF_MLP_PART_IN;
    int nb = nb_in, o = 0;
    if(i < GF_MLP_W2) {  // W1 (interleaved k, input) then b1
        o = i;
    } else if(i < GF_MLP_B2) {
        base   = p_w2;
        stride = (int64_t)GF_MLP_H * GF_MLP_H;
        nb     = nb_w2;
        o      = i - GF_MLP_W2;
    } else {
        // b2, w3, b3 live in the output partials as [w3 | b2 | b3]
        base   = p_out;
        stride = GF_MLP_PART_OUT;
        nb     = nb_out;
        if(i < GF_MLP_W3)
            o = GF_MLP_H + (i - GF_MLP_B2);
        else if(i < GF_MLP_B3)
            o = i - GF_MLP_W3;
        else
            o = 2 * GF_MLP_H;
    }
    float s0 = 0.f, s1 = 0.f;
    if(i < GF_MLP_NPARAM) {
        int b = w;
        for(; b + 4 < nb; b += 8) {
            s0 += base[(int64_t)b * stride + o];
            s1 += base[(int64_t)(b + 4) * stride + o];
        }
        if(b < nb)
            s0 += base[(int64_t)b * stride + o];
    }
    red[w][lane] = s0 + s1;
    __syncthreads();
    if(w == 0 && i < GF_MLP_NPARAM)
        g[i] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

// Advance the training state by one step: batch key, cosine-decayed learning
// rate lr0 (0.05 + 0.95 (1 + cos(pi k / steps)) / 2), Adam bias corrections.
__global__ void k_step_begin(GfTrainState *st) {
    const int k = st->k;
    st->adam_t += 1;
    st->key = st->seed * 0x9E3779B97F4A7C15ull + 2ull * (uint64_t)k + 1ull;
    const double cs = 0.5 * (1.0 + cos(3.141592653589793 * (double)k / (double)st->steps));
    st->lr = (float)(st->lr0 * (0.05 + 0.95 * cs));
    st->c1 = 1.f - powf(st->b1, (float)st->adam_t);
    st->c2 = 1.f - powf(st->b2, (float)st->adam_t);
    st->k  = k + 1;
}

// Adam step on the FP32 master parameters, then the bf16 copies of W2 and W2^T.
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ P, float *__restrict__ m, float *__restrict__ v,
                                              const float *__restrict__ g, const GfTrainState *__restrict__ st,
                                              uint16_t *__restrict__ w2b, uint16_t *__restrict__ w2t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= GF_MLP_NPARAM)
        return;
    const float lr = st->lr, b1 = st->b1, b2 = st->b2, eps = st->eps, c1 = st->c1, c2 = st->c2;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i]           = mi;
    v[i]           = vi;
    const float p  = P[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    P[i]           = p;
    if(i >= GF_MLP_W2 && i < GF_MLP_B2) {
        const int jk = i - GF_MLP_W2, j = jk / GF_MLP_H, k = jk % GF_MLP_H;
        const uint16_t h = f2bf(p);
        w2b[jk]                = h;
        w2t[k * GF_MLP_H + j]  = h;
    }
}

// bf16 copies of W2 and W2^T from the master parameters.
__global__ __launch_bounds__(256) void k_w2_bf16(const float *__restrict__ P, uint16_t *__restrict__ w2b,
                                                 uint16_t *__restrict__ w2t) {
    const int jk = blockIdx.x * blockDim.x + threadIdx.x;
    if(jk >= GF_MLP_H * GF_MLP_H)
        return;
    const int j = jk / GF_MLP_H, k = jk % GF_MLP_H;
    const uint16_t h = f2bf(P[GF_MLP_W2 + jk]);
    w2b[jk]               = h;
    w2t[k * GF_MLP_H + j] = h;
}

// ----------------------------------------------------------------- sampling

// n points of the tabulated proposal: cell by binary search of the FP64 CDF
// (cdf[c] = P(cell <= c), cdf[G*G - 1] = 1), position uniform in the cell;
// w = box area / (n * p_cell * G^2) = 1 / (n q(x)).
__global__ __launch_bounds__(256) void k_sample(const double *__restrict__ cdf, const double *__restrict__ pcell,
                                                int G, GfBox box, int64_t n, uint64_t key, double *__restrict__ m12,
                                                double *__restrict__ m13, double *__restrict__ wt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n)
        return;
    const int64_t ncell = (int64_t)G * G;
    const double r = unif(key, 3 * (uint64_t)i);
    int64_t lo = 0, hi = ncell - 1;  // first c with cdf[c] > r
    while(lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if(cdf[mid] > r)
            hi = mid;
        else
            lo = mid + 1;
    }
    const int64_t c = lo;
    const double u = ((double)(c % G) + unif(key, 3 * (uint64_t)i + 1)) / G;
    const double v = ((double)(c / G) + unif(key, 3 * (uint64_t)i + 2)) / G;
    const double area = (box.u12 - box.l12) * (box.u13 - box.l13);
    m12[i] = box.l12 + u * (box.u12 - box.l12);
    m13[i] = box.l13 + v * (box.u13 - box.l13);
    wt[i]  = area / ((double)n * pcell[c] * (double)ncell);
}

// Uniform points of the box with w = area / n.
__global__ __launch_bounds__(256) void k_sample_uniform(GfBox box, int64_t n, uint64_t key, double *__restrict__ m12,
                                                        double *__restrict__ m13, double *__restrict__ wt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n)
        return;
    const double area = (box.u12 - box.l12) * (box.u13 - box.l13);
    m12[i] = box.l12 + unif(key, 2 * (uint64_t)i) * (box.u12 - box.l12);
    m13[i] = box.l13 + unif(key, 2 * (uint64_t)i + 1) * (box.u13 - box.l13);
    wt[i]  = area / (double)n;
}

// Probe of the MFMA operand/result layout: lane l feeds a8[l][0..7], b8[l][0..7]
// and writes its 4 result registers.
__global__ void k_bf16_probe(const uint16_t *a8, const uint16_t *b8, float *d) {
    const int l = threadIdx.x;
    bf16x8 av, bv;
#pragma unroll
    for(int t = 0; t < 8; ++t) {
        av[t] = __builtin_bit_cast(__bf16, a8[8 * l + t]);
        bv[t] = __builtin_bit_cast(__bf16, b8[8 * l + t]);
    }
    floatx4 r = {0.f, 0.f, 0.f, 0.f};
    r         = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, r, 0, 0, 0);
#pragma unroll
    for(int t = 0; t < 4; ++t)
        d[4 * l + t] = r[t];
}

inline int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

namespace gfm {

int check(hipError_t e) { return e == hipSuccess ? 0 : (int)e; }

int launch_train_step(const GfMlpDev &D, const double *m12, const double *m13, int64_t ndata, GfBox box, int batch,
                      hipStream_t s) {
    const int half = batch / 2;
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(1), 0, s, D.state);
    int rc = check(hipGetLastError());
    if(rc)
        return rc;
    hipLaunchKernelGGL(k_batch_train, dim3(nblk(batch, 256)), dim3(256), 0, s, m12, m13, ndata, box, half, half,
                       (const GfTrainState *)D.state, D.xy);
    if((rc = check(hipGetLastError())) || (rc = launch_forward(D, D.xy, batch, D.out, s))
       || (rc = launch_loss(D.out, half, half, D.coef, D.loss, s))
       || (rc = launch_backward(D, D.xy, D.coef, batch, D.grad, s)))
        return rc;
    hipLaunchKernelGGL(k_adam, dim3(nblk(GF_MLP_NPARAM, 256)), dim3(256), 0, s, D.P, D.m, D.v, D.grad,
                       (const GfTrainState *)D.state, D.w2b, D.w2t);
    return check(hipGetLastError());
}

int launch_batch_grid(int G, float2 *xy, hipStream_t s) {
    const int64_t nc = (int64_t)G * G;
    hipLaunchKernelGGL(k_batch_grid, dim3(nblk(nc, 256)), dim3(256), 0, s, G, nc, xy);
    return check(hipGetLastError());
}

int launch_forward(const GfMlpDev &D, const float2 *xy, int n, float *out, hipStream_t s) {
    hipLaunchKernelGGL(k_layer1, dim3(nblk((int64_t)n * GF_MLP_H, 256)), dim3(256), 0, s, xy, n, D.P, D.z1, D.h1);
    int rc = check(hipGetLastError());
    if(rc)
        return rc;
    hipLaunchKernelGGL(k_layer2, dim3(n / 64), dim3(256), 0, s, D.h1, D.w2b, D.P, n, D.z2, out);
    return check(hipGetLastError());
}

int launch_loss(const float *out, int nd, int nu, float *coef, float *loss, hipStream_t s) {
    hipLaunchKernelGGL(k_loss, dim3(1), dim3(1024), 0, s, out, nd, nu, coef, loss);
    return check(hipGetLastError());
}

int launch_backward(const GfMlpDev &D, const float2 *xy, const float *coef, int n, float *grad, hipStream_t s) {
    const int nb = n / 64;
    hipLaunchKernelGGL(k_back_out, dim3(nb), dim3(GF_MLP_H * GF_BO_SPLIT), 0, s, D.z2, coef, D.P, n, D.delta2,
                       D.part_out);
    int rc = check(hipGetLastError());
    if(rc)
        return rc;
    const int chunks = n / 32;
    const int nbw    = chunks < GF_MLP_W2_BLOCKS ? chunks : GF_MLP_W2_BLOCKS;
    const int cpb    = (chunks + nbw - 1) / nbw;
    hipLaunchKernelGGL(k_back_w2, dim3(nbw), dim3(256), 0, s, D.delta2, D.h1, n, cpb, D.part_w2);
    if((rc = check(hipGetLastError())))
        return rc;
    hipLaunchKernelGGL(k_back_h1, dim3(nb), dim3(256), 0, s, D.delta2, D.w2t, D.z1, xy, n, D.part_in);
    if((rc = check(hipGetLastError())))
        return rc;
    hipLaunchKernelGGL(k_grad_sum, dim3(nblk(GF_MLP_NPARAM, 64)), dim3(256), 0, s, D.part_in, nb, D.part_w2, nbw,
                       D.part_out, nb, grad);
    return check(hipGetLastError());
}


int launch_w2_bf16(const GfMlpDev &D, hipStream_t s) {
    hipLaunchKernelGGL(k_w2_bf16, dim3(nblk(GF_MLP_H * GF_MLP_H, 256)), dim3(256), 0, s, D.P, D.w2b, D.w2t);
    return check(hipGetLastError());
}

int launch_sample(const double *cdf, const double *pcell, int G, GfBox box, int64_t n, uint64_t key, double *m12,
                  double *m13, double *w, hipStream_t s) {
    hipLaunchKernelGGL(k_sample, dim3(nblk(n, 256)), dim3(256), 0, s, cdf, pcell, G, box, n, key, m12, m13, w);
    return check(hipGetLastError());
}

int launch_sample_uniform(GfBox box, int64_t n, uint64_t key, double *m12, double *m13, double *w, hipStream_t s) {
    hipLaunchKernelGGL(k_sample_uniform, dim3(nblk(n, 256)), dim3(256), 0, s, box, n, key, m12, m13, w);
    return check(hipGetLastError());
}

int launch_bf16_probe(const uint16_t *a8, const uint16_t *b8, float *d, hipStream_t s) {
    hipLaunchKernelGGL(k_bf16_probe, dim3(1), dim3(64), 0, s, a8, b8, d);
    return check(hipGetLastError());
}

}  // namespace gfm
