// nn.h -- learned signed-distance cull: one small MLP per grain template,
// positional encoding (6 octaves) -> 64 -> 64 -> 64 -> 1 with ReLU, evaluated with
// gfx950's BF16 MFMA (v_mfma_f32_16x16x32_bf16; one 16x16x16 step for the
// 48-wide input layer's last 16 columns) and FP32 accumulation.
//
// No reference counterpart (the reference has only the exact BVH query and its
// conservative hull culls).  The network predicts the signed distance of a point
// in a template's frame; a node whose prediction exceeds the template's margin
// (three times the largest validation error) skips the exact query.
//
// MFMA operand layouts per lane l (0..63), D[m][n] with n = l % 16 and
// m = 4 (l / 16) + {0..3} for both shapes:
//   16x16x16: A[m][k] with m = l % 16, k = 4 (l / 16) + {0..3}; B[k][n] with n = l % 16
//             and the same k;
//   16x16x32: A[m][k] with m = l % 16, k = 8 (l / 16) + {0..7}; B[k][n] likewise.
#pragma once
#include "dem_math.h"

namespace nn {
constexpr int IN = 39;  // xyz + sin/cos of 2^k pi x, k = 0..5, per axis
constexpr int K1 = 48;  // input width padded to three MFMA k-steps
constexpr int H = 64;
constexpr int OCTAVES = 6;
// FP32 master parameters per template (row-major [out][in] weights).
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + H * IN;
constexpr int OFF_W2 = OFF_B1 + H;
constexpr int OFF_B2 = OFF_W2 + H * H;
constexpr int OFF_W3 = OFF_B2 + H;
constexpr int OFF_B3 = OFF_W3 + H * H;
constexpr int OFF_W4 = OFF_B3 + H;
constexpr int OFF_B4 = OFF_W4 + H;
constexpr int PARAMS = OFF_B4 + 1;  // 10,945
constexpr int STRIDE = 11008;        // per-template stride of master / gradient / Adam buffers
// BF16 weight block per template (MFMA B operands; W1 zero-padded to K1 columns).
constexpr int BW1 = 0, BW2 = H * K1, BW3 = BW2 + H * H, BSTRIDE = BW3 + H * H;
// FP32 inference block per template: biases, output layer, input normalisation, margin
// (label units), the signed distance in template units at or below which the network
// cannot cull a node within its validated error ((margin - error) x label scale), and
// the label scale (template units per label unit).
constexpr int FB1 = 0, FB2 = 64, FB3 = 128, FW4 = 192, FB4 = 256, FCENTER = 257, FSCALE = 260, FMARGIN = 261,
              FNEAR = 262, FLABEL = 263;
constexpr int FSTRIDE = 264;
}  // namespace nn

typedef short nn_bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 nn_bf16x8 __attribute__((ext_vector_type(8)));
typedef float nn_f32x4 __attribute__((ext_vector_type(4)));

// acc += A B over one 32-wide k block: a and b are lane l's 8 consecutive k values.
DEM_D nn_f32x4 nn_mfma32(const uint16_t* a, const uint16_t* b, nn_f32x4 acc) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<const nn_bf16x8*>(a),
                                                   *reinterpret_cast<const nn_bf16x8*>(b), acc, 0, 0, 0);
}
// The same from 8 values gathered per lane (strided operands).
DEM_D nn_f32x4 nn_mfma32g(const uint16_t (&a)[8], const uint16_t (&b)[8], nn_f32x4 acc) {
    nn_bf16x8 av, bv;
    for (int j = 0; j < 8; j++) {
        av[j] = __builtin_bit_cast(__bf16, a[j]);
        bv[j] = __builtin_bit_cast(__bf16, b[j]);
    }
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
}

DEM_D uint16_t nn_to_bf16(float x) {  // round to nearest even
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
DEM_D float nn_from_bf16(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// Input features of one point into a K1-wide BF16 row.
DEM_D void nn_features(v3 p, const float* f, uint16_t* row) {
    float s = f[nn::FSCALE];
    float u[3] = {(p.x - f[nn::FCENTER]) * s, (p.y - f[nn::FCENTER + 1]) * s, (p.z - f[nn::FCENTER + 2]) * s};
    row[0] = nn_to_bf16(u[0]);
    row[1] = nn_to_bf16(u[1]);
    row[2] = nn_to_bf16(u[2]);
    float scale = 3.14159265358979f;
#pragma unroll 1
    for (int k = 0; k < nn::OCTAVES; k++) {
        for (int d = 0; d < 3; d++) {
            float a = scale * u[d];
            row[3 + 6 * k + 2 * d] = nn_to_bf16(__sinf(a));
            row[3 + 6 * k + 2 * d + 1] = nn_to_bf16(__cosf(a));
        }
        scale *= 2.0f;
    }
    for (int c = nn::IN; c < nn::K1; c++) row[c] = 0;
}

// out[r][:] = relu(in[r][:] W^T + b) for RT 16-row tiles; in has KT*16 columns.
// W is [64][KT*16] BF16 row-major in global memory; in/out are BF16 in LDS.
// K32 selects the 16x16x32 MFMA for 32-wide k blocks (the rest in 16-wide steps).
// The collection kernel's inference uses the 16x16x16 form only: under that
// kernel's register pressure hipcc (ROCm 7.2) assigns an MFMA's destination to
// registers of its A or B operand.  The 16x16x16 form gives correct results so;
// the 16x16x32 form, whose operands span twice the registers, gave wrong
// predictions there, while in the training and batch-inference kernels (operands
// kept apart) it equals the 16x16x16 form bit for bit.
template <int RT, int KT, bool K32 = true>
DEM_D void nn_layer(const uint16_t* in, uint16_t* out, const uint16_t* __restrict__ W, const float* __restrict__ bias,
                    int lane) {
    constexpr int LD = KT * 16;
    constexpr int BLOCKS32 = K32 ? KT / 2 : 0;
    const int r = lane & 15, q = (lane >> 4) * 4, q8 = (lane >> 4) * 8;
#pragma unroll 1
    for (int nt = 0; nt < 4; nt++) {  // not unrolled: few registers live inside the collection kernel
        nn_f32x4 acc[RT];
        for (int mt = 0; mt < RT; mt++) acc[mt] = nn_f32x4{0, 0, 0, 0};
        for (int kb = 0; kb < BLOCKS32; kb++) {
            const uint16_t* b = W + (nt * 16 + r) * LD + kb * 32 + q8;
            for (int mt = 0; mt < RT; mt++) acc[mt] = nn_mfma32(in + (mt * 16 + r) * LD + kb * 32 + q8, b, acc[mt]);
        }
        for (int kt = 2 * BLOCKS32; kt < KT; kt++) {
            nn_bf16x4 b = *reinterpret_cast<const nn_bf16x4*>(W + (nt * 16 + r) * LD + kt * 16 + q);
            for (int mt = 0; mt < RT; mt++) {
                nn_bf16x4 a = *reinterpret_cast<const nn_bf16x4*>(in + (mt * 16 + r) * LD + kt * 16 + q);
                acc[mt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[mt], 0, 0, 0);
            }
        }
        const int col = nt * 16 + r;
        const float bv = bias[col];
        for (int mt = 0; mt < RT; mt++)
            for (int i = 0; i < 4; i++) out[(mt * 16 + q + i) * nn::H + col] = nn_to_bf16(fmaxf(acc[mt][i] + bv, 0.0f));
    }
}

// Predicted signed distance of RT*16 points whose feature rows are in `bufA`
// (row r at bufA + r*K1); `bufB` is scratch of RT*16*64 BF16.  Lane l < RT*16
// returns the prediction of row l.
template <int RT, bool K32 = true>
DEM_D float nn_forward(uint16_t* bufA, uint16_t* bufB, const uint16_t* __restrict__ wb, const float* __restrict__ f,
                       int lane) {
    nn_layer<RT, 3, K32>(bufA, bufB, wb + nn::BW1, f + nn::FB1, lane);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    nn_layer<RT, 4, K32>(bufB, bufA, wb + nn::BW2, f + nn::FB2, lane);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    nn_layer<RT, 4, K32>(bufA, bufB, wb + nn::BW3, f + nn::FB3, lane);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    float y = f[nn::FB4];
    if (lane < RT * 16) {
        const uint16_t* h = bufB + lane * nn::H;
        for (int c = 0; c < nn::H; c++) y += nn_from_bf16(h[c]) * f[nn::FW4 + c];
    }
    return y;
}

// ---------------------------------------------------------------- training
// dW[o][i] += sum_r G[r][o] A[r][i]  (G: 64 rows x 64 outputs, A: 64 rows x LDA
// inputs, both BF16 in LDS); accumulated into an FP32 [64][NIN] block with LDS
// atomics.  M = outputs, N = inputs, K = rows.
template <int LDA, int NIN>
DEM_D void nn_weight_grad(const uint16_t* G, const uint16_t* A, float* accW, int lane) {
    const int r = lane & 15, q = (lane >> 4) * 4, q8 = (lane >> 4) * 8;
    constexpr int NT = (NIN + 15) / 16;
    for (int mt = 0; mt < 4; mt++)
        for (int nt = 0; nt < NT; nt++) {
            nn_f32x4 acc = {0, 0, 0, 0};
            for (int kt = 0; kt < 2; kt++) {  // 64 rows = two 32-wide k blocks
                uint16_t a[8], b[8];
                for (int j = 0; j < 8; j++) {
                    int row = kt * 32 + q8 + j;
                    a[j] = G[row * nn::H + mt * 16 + r];
                    b[j] = A[row * LDA + nt * 16 + r];
                }
                acc = nn_mfma32g(a, b, acc);
            }
            int in = nt * 16 + r;
            if (in < NIN)
                for (int i = 0; i < 4; i++) atomicAdd(accW + (mt * 16 + q + i) * NIN + in, acc[i]);
        }
}
// Gout[r][i] = (sum_o G[r][o] W[o][i]) * (Hmask[r][i] > 0), W BF16 [64][64] global;
// the bias gradient of the lower layer (column sums of the FP32 values, before
// BF16 rounding) is added to accB.
DEM_D void nn_backprop(const uint16_t* G, const uint16_t* __restrict__ W, const uint16_t* Hmask, uint16_t* Gout,
                       float* accB, int lane) {
    const int r = lane & 15, q = (lane >> 4) * 4, q8 = (lane >> 4) * 8;
    for (int nt = 0; nt < 4; nt++) {
        const int col = nt * 16 + r;
        float colsum = 0.0f;
        for (int mt = 0; mt < 4; mt++) {
            nn_f32x4 acc = {0, 0, 0, 0};
            for (int kt = 0; kt < 2; kt++) {
                uint16_t b[8];
                for (int j = 0; j < 8; j++) b[j] = W[(kt * 32 + q8 + j) * nn::H + nt * 16 + r];
                acc = nn_mfma32(G + (mt * 16 + r) * nn::H + kt * 32 + q8, b, acc);
            }
            for (int i = 0; i < 4; i++) {
                int row = mt * 16 + q + i;
                float v = nn_from_bf16(Hmask[row * nn::H + col]) > 0.0f ? acc[i] : 0.0f;
                Gout[row * nn::H + col] = nn_to_bf16(v);
                colsum += v;
            }
        }
        colsum += __shfl_xor(colsum, 16);
        colsum += __shfl_xor(colsum, 32);
        if (lane < 16) atomicAdd(accB + col, colsum);
    }
}

struct NNTrainArgs {
    const float4* pts;       // template-frame points
    const float* labels;     // target signed distance
    const uint32_t* tstart;  // first point of each template
    const uint32_t* tcount;  // points of each template (multiple of 64)
    const uint16_t* wb;      // BF16 weight blocks
    const float* fb;         // FP32 inference blocks
    float* partial;          // [template][slice][STRIDE] gradient partials
    float* loss;             // [template] sum of squared errors
    uint32_t slices;
    float lossScale;         // 1 / points per template: gradient of the mean squared error
};

// One (slice, template) per block of two waves; each wave processes 64-point tiles:
// MFMA forward, loss, MFMA backward; gradients summed in LDS, then written once.
__global__ __launch_bounds__(128) void k_nn_grad(NNTrainArgs a) {
    __shared__ uint16_t X[2][64 * nn::K1];
    __shared__ uint16_t H1[2][64 * nn::H], H2[2][64 * nn::H], H3[2][64 * nn::H];
    __shared__ uint16_t GA[2][64 * nn::H], GB[2][64 * nn::H];
    __shared__ float DY[2][64];
    __shared__ float acc[nn::STRIDE];
    const uint32_t t = blockIdx.y, slice = blockIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int k = threadIdx.x; k < nn::STRIDE; k += blockDim.x) acc[k] = 0.0f;
    __syncthreads();
    const uint16_t* wb = a.wb + size_t(t) * nn::BSTRIDE;
    const float* fb = a.fb + size_t(t) * nn::FSTRIDE;
    const uint32_t tiles = a.tcount[t] / 64;
    float sq = 0.0f;
    for (uint32_t tile = slice * 2 + wv; tile < tiles; tile += a.slices * 2) {
        uint32_t idx = a.tstart[t] + tile * 64 + lane;
        float4 pt = a.pts[idx];
        float label = a.labels[idx];
        nn_features({pt.x, pt.y, pt.z}, fb, X[wv] + lane * nn::K1);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        nn_layer<4, 3>(X[wv], H1[wv], wb + nn::BW1, fb + nn::FB1, lane);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        nn_layer<4, 4>(H1[wv], H2[wv], wb + nn::BW2, fb + nn::FB2, lane);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        nn_layer<4, 4>(H2[wv], H3[wv], wb + nn::BW3, fb + nn::FB3, lane);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const uint16_t* h3 = H3[wv] + lane * nn::H;
        float y = fb[nn::FB4];
        for (int c = 0; c < nn::H; c++) y += nn_from_bf16(h3[c]) * fb[nn::FW4 + c];
        float err = y - label;
        sq += err * err;
        const float dy = 2.0f * err * a.lossScale;
        DY[wv][lane] = dy;
        for (int c = 0; c < nn::H; c++) {
            float g = nn_from_bf16(h3[c]) > 0.0f ? dy * fb[nn::FW4 + c] : 0.0f;
            GA[wv][lane * nn::H + c] = nn_to_bf16(g);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        {   // output layer: lane c sums column c over the tile's rows
            float gw4 = 0.0f, gb3 = 0.0f, gb4 = dy;
            for (int row = 0; row < 64; row++) {
                float h = nn_from_bf16(H3[wv][row * nn::H + lane]), d = DY[wv][row];
                gw4 += h * d;
                gb3 += h > 0.0f ? d : 0.0f;
            }
            for (int d = 32; d > 0; d >>= 1) gb4 += __shfl_xor(gb4, d);
            atomicAdd(acc + nn::OFF_W4 + lane, gw4);
            atomicAdd(acc + nn::OFF_B3 + lane, gb3 * fb[nn::FW4 + lane]);
            if (lane == 0) atomicAdd(acc + nn::OFF_B4, gb4);
        }
        nn_weight_grad<nn::H, nn::H>(GA[wv], H2[wv], acc + nn::OFF_W3, lane);
        nn_backprop(GA[wv], wb + nn::BW3, H2[wv], GB[wv], acc + nn::OFF_B2, lane);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        nn_weight_grad<nn::H, nn::H>(GB[wv], H1[wv], acc + nn::OFF_W2, lane);
        nn_backprop(GB[wv], wb + nn::BW2, H1[wv], GA[wv], acc + nn::OFF_B1, lane);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        nn_weight_grad<nn::K1, nn::IN>(GA[wv], X[wv], acc + nn::OFF_W1, lane);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    for (int d = 32; d > 0; d >>= 1) sq += __shfl_xor(sq, d);
    if (lane == 0 && sq != 0.0f) atomicAdd(a.loss + t, sq);
    __syncthreads();
    float* out = a.partial + (size_t(t) * a.slices + slice) * nn::STRIDE;
    for (int k = threadIdx.x; k < nn::PARAMS; k += blockDim.x) out[k] = acc[k];
}

// Sum of slice partials -> gradient; optional Adam step on the FP32 master.
__global__ void k_nn_adam(float* master, float* m1, float* m2, const float* partial, float* grad, uint32_t slices,
                          uint32_t templates, float lr, float beta1, float beta2, float eps, float bc1, float bc2) {
    const size_t total = size_t(templates) * nn::PARAMS;
    for (size_t idx = blockIdx.x * size_t(blockDim.x) + threadIdx.x; idx < total; idx += size_t(gridDim.x) * blockDim.x) {
        size_t t = idx / nn::PARAMS, k = idx % nn::PARAMS, at = t * nn::STRIDE + k;
        float g = 0.0f;
        for (uint32_t s = 0; s < slices; s++) g += partial[(t * slices + s) * nn::STRIDE + k];
        grad[at] = g;
        if (lr > 0.0f) {
            float a = beta1 * m1[at] + (1.0f - beta1) * g;
            float b = beta2 * m2[at] + (1.0f - beta2) * g * g;
            m1[at] = a;
            m2[at] = b;
            master[at] -= lr * (a / bc1) / (sqrtf(b / bc2) + eps);
        }
    }
}
// FP32 master -> BF16 weight blocks and FP32 inference blocks (layer parameters only).
__global__ void k_nn_pack(const float* master, uint16_t* wb, float* fb, uint32_t templates) {
    const uint32_t t = blockIdx.y;
    if (t >= templates) return;
    const float* p = master + size_t(t) * nn::STRIDE;
    uint16_t* w = wb + size_t(t) * nn::BSTRIDE;
    float* f = fb + size_t(t) * nn::FSTRIDE;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nn::BSTRIDE; k += gridDim.x * blockDim.x) {
        float v;
        if (k < nn::BW2) {
            uint32_t o = k / nn::K1, c = k % nn::K1;
            v = c < nn::IN ? p[nn::OFF_W1 + o * nn::IN + c] : 0.0f;
        } else if (k < nn::BW3) {
            v = p[nn::OFF_W2 + (k - nn::BW2)];
        } else {
            v = p[nn::OFF_W3 + (k - nn::BW3)];
        }
        w[k] = nn_to_bf16(v);
        if (k < 64) {
            f[nn::FB1 + k] = p[nn::OFF_B1 + k];
            f[nn::FB2 + k] = p[nn::OFF_B2 + k];
            f[nn::FB3 + k] = p[nn::OFF_B3 + k];
            f[nn::FW4 + k] = p[nn::OFF_W4 + k];
        }
        if (k == 0) f[nn::FB4] = p[nn::OFF_B4];
    }
}
// Predictions for template-major 64-point tiles.
__global__ __launch_bounds__(64) void k_nn_infer(const float4* pts, const uint32_t* tileTemplate, const uint16_t* wb,
                                                 const float* fb, float* out) {
    __shared__ uint16_t A[64 * nn::H], B[64 * nn::H];
    const uint32_t tile = blockIdx.x;
    const int lane = threadIdx.x;
    const uint32_t t = tileTemplate[tile];
    const float* f = fb + size_t(t) * nn::FSTRIDE;
    float4 pt = pts[size_t(tile) * 64 + lane];
    nn_features({pt.x, pt.y, pt.z}, f, A + lane * nn::K1);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    float y = nn_forward<4>(A, B, wb + size_t(t) * nn::BSTRIDE, f, lane);
    out[size_t(tile) * 64 + lane] = y;
}
// MFMA operand-layout probe: D = A B for one BF16 tile, 16x16xK with K = 16 or 32
// (A is 16 x K and B is K x 16, both row-major).
__global__ __launch_bounds__(64) void k_mfma_probe(const uint16_t* A, const uint16_t* B, float* D, int k32) {
    const int lane = threadIdx.x, r = lane & 15, q = (lane >> 4) * 4, q8 = (lane >> 4) * 8;
    nn_f32x4 d = {0, 0, 0, 0};
    if (k32) {
        uint16_t a[8], b[8];
        for (int j = 0; j < 8; j++) {
            a[j] = A[r * 32 + q8 + j];
            b[j] = B[(q8 + j) * 16 + r];
        }
        d = nn_mfma32g(a, b, d);
    } else {
        nn_bf16x4 a, b;
        for (int j = 0; j < 4; j++) {
            a[j] = short(A[r * 16 + q + j]);
            b[j] = short(B[(q + j) * 16 + r]);
        }
        d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, d, 0, 0, 0);
    }
    for (int i = 0; i < 4; i++) D[(q + i) * 16 + r] = d[i];
}
