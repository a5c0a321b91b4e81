// Radiance-cache MLP on gfx950 MFMA: BF16 operands, FP32 accumulation.
//
// The network replaces the reference's never-compiled RadianceField
// (editor/mlp.cuh:71-154: dense layers out = W x + b with W row-major by
// output, mlp.cuh:48-67; xavier init :13-31) with the online-trained cache:
// 62 encoded inputs -> 5 x 64 ReLU -> 3 linear outputs.
//
// Layout.  Every dense layer is computed transposed, Z^T = W . X^T, with
// v_mfma_f32_16x16x32_bf16: A = weights (16 output rows x 32 inputs per
// fragment), B = activations (32 inputs x 16 queries), D = 16 outputs x 16
// queries.  Lane l holds A row l%16 / B column l%16 at k = 8(l/16) + j, and D
// rows 4(l/16) + v of column l%16.  A layer's D registers are the next layer's
// B operand without any data movement: chunk c of the next layer takes row
// blocks 2c and 2c+1, so its k index j maps to hidden unit
// 16(2c + j/4) + 4(l/16) + j%4; the weight fragments are packed in that
// order.  Weights live in fragment order (64 lanes x 8 bf16 per fragment),
// regenerated from the FP32 master copy after every optimiser step.
//
// Training.  One workgroup owns 128 samples (4 waves x 2 tiles of 16).  The
// forward pass keeps each layer's bf16 input in LDS as [unit][sample]; the
// backward pass runs the transposed products W^T . delta^T for the deltas and
// delta^T . H over the workgroup's samples for the weight gradient (K = 32
// samples per MFMA).  Each workgroup writes its partial gradient; a second
// kernel sums the partials in a fixed order (deterministic), unscales by the
// loss scale and flags non-finite values; Adam then updates the FP32 master
// weights unless the flag is set, in which case the step is skipped and the
// loss scale halves.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "krt_device.h"
#include "krt_internal.h"

namespace krt {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kFragF = 42;  // forward fragments: 5 layers x (4 row blocks x 2 chunks) + 2
constexpr int kFragB = 36;  // backward fragments: 4 layers x 8 + 4
constexpr int kFragElems = 64 * 8;
constexpr int kStride = 136;  // LDS row stride in bf16 (128 samples + 8), 16-byte aligned rows
constexpr int kTrainBlock = 256;
constexpr int kTrainSamples = 128;
constexpr int kInferBlock = 256;

__host__ __device__ constexpr int in_dim(int L) { return L == 0 ? kMlpEncoded : kMlpHidden; }
__host__ __device__ constexpr int out_dim(int L) { return L == kMlpLayers - 1 ? kMlpOut : kMlpHidden; }
__host__ __device__ constexpr int w_off(int L) {
    int o = 0;
    for (int i = 0; i < L; i++) o += in_dim(i) * out_dim(i) + out_dim(i);
    return o;
}
__host__ __device__ constexpr int b_off(int L) { return w_off(L) + in_dim(L) * out_dim(L); }
static_assert(w_off(kMlpLayers) == kMlpParams, "parameter count");

KRT_DEV f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

// Layer 0's inputs in fragment order.  Fragment input k = 32c + 8g + j
// (chunk c, lane group g, element j) holds encoded feature enc_feature(k):
// every lane group computes six of the 24 sin/cos pairs (features 2p, 2p + 1
// of pair p) and four raw inputs, so the sine work is spread evenly:
//   chunk 0: pairs 6g .. 6g + 3      (features 12g .. 12g + 7)
//   chunk 1: pairs 6g + 4, 6g + 5    (features 12g + 8 .. 12g + 11),
//            raw features 48 + 4g .. 48 + 4g + 3
// Features 0..47 are the sin/cos pairs, 48..60 the raw inputs (pos mapped to
// [0, 1]), 61 the constant 1, 62..63 zero.  The weights are packed, and the
// layer-0 gradient written, through the same map.
__host__ __device__ constexpr int enc_feature(int k) {
    const int c = k >> 5, g = (k >> 3) & 3, j = k & 7;
    return c == 0 ? 12 * g + j : (j < 4 ? 12 * g + 8 + j : 48 + 4 * g + (j - 4));
}

// Turns of frequency pair p (0..23): pos x 4 octaves, dir x 2, normal x 2,
// from the query's 13 mapped inputs q (pos already in [0, 1]).
KRT_DEV float pair_turns(int p, const float *q) {
    const int vec = p < 12 ? 0 : (p < 18 ? 1 : 2);
    const int base = vec == 0 ? 0 : (vec == 1 ? 12 : 18);
    const int comp = (p - base) % 3, k = (p - base) / 3;
    return q[3 * vec + comp] * (0.5f * (float)(1 << k));
}

// The 13 raw floats of 16 consecutive queries (one MFMA tile) are 208
// contiguous floats: the wave's 64 lanes load them coalesced (lane l takes
// floats l, l + 64, l + 128, l + 192), then stage them into the wave's LDS
// slice with pos mapped to [0, 1].
KRT_DEV void load_tile(const float *__restrict__ raw, uint32_t q0, uint32_t n, int lane, float (&v)[4]) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int e = lane + 64 * i;
        const uint32_t q = q0 + (uint32_t)(e / kMlpRaw);
        v[i] = (e < 16 * kMlpRaw && q < n) ? raw[(size_t)kMlpRaw * q0 + e] : 0.0f;
    }
}

KRT_DEV void stage_tile(const float (&v)[4], const MlpEncoding &enc, int lane, float *tile) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int e = lane + 64 * i;
        if (e < 16 * kMlpRaw) {
            const int a = e % kMlpRaw;
            tile[e] = a < 3 ? (v[i] - enc.pmin[a]) * enc.pscale[a] : v[i];
        }
    }
    __builtin_amdgcn_wave_barrier();
}

// The 16 features lane group g owns for one query (q: its 13 mapped inputs).
KRT_DEV void encode(const float *q, int g, bf16x8 &b0, bf16x8 &b1) {
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
        float c, s;
        sincos_turn(pair_turns(6 * g + jj, q), c, s);
        b0[2 * jj] = (__bf16)s;
        b0[2 * jj + 1] = (__bf16)c;
    }
#pragma unroll
    for (int jj = 0; jj < 2; jj++) {
        float c, s;
        sincos_turn(pair_turns(6 * g + 4 + jj, q), c, s);
        b1[2 * jj] = (__bf16)s;
        b1[2 * jj + 1] = (__bf16)c;
    }
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
        const int r = 4 * g + jj;  // feature 48 + r
        b1[4 + jj] = (__bf16)(r < kMlpRaw ? q[r] : (r == kMlpRaw ? 1.0f : 0.0f));
    }
}

KRT_DEV bf16x8 zero8() {
    bf16x8 z;
#pragma unroll
    for (int j = 0; j < 8; j++) z[j] = (__bf16)0.0f;
    return z;
}

// ---------------------------------------------------------------------------
// Fragment packing: FP32 master -> bf16 fragments (forward W, backward W^T).
// ---------------------------------------------------------------------------
__global__ void pack_kernel(const float *__restrict__ params, __bf16 *__restrict__ fragF,
                            __bf16 *__restrict__ fragB) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (kFragF + kFragB) * kFragElems) return;
    const int frag = e / kFragElems, lane = (e / 8) % 64, j = e % 8, g = lane >> 4, r = lane & 15;
    float val = 0.0f;
    if (frag < kFragF) {
        const int L = frag < 40 ? frag / 8 : 5, loc = frag - 8 * L, m = loc / 2, c = loc % 2;
        const int o = 16 * m + r;
        const int k = (L == 0) ? enc_feature(32 * c + 8 * g + j) : 16 * (2 * c + (j >> 2)) + 4 * g + (j & 3);
        if (o < out_dim(L) && k < in_dim(L)) val = params[w_off(L) + o * in_dim(L) + k];
        fragF[e] = (__bf16)val;
    } else {
        const int fb = frag - kFragF;
        const int L = fb < 32 ? fb / 8 + 1 : 5, loc = fb - 8 * (L - 1);
        const int m = (L == 5) ? loc : loc / 2, c = (L == 5) ? 0 : loc % 2;
        const int i = 16 * m + r;  // input unit of layer L
        int k = 16 * (2 * c + (j >> 2)) + 4 * g + (j & 3);  // output unit of layer L
        bool ok = true;
        if (L == 5) {
            k = 4 * g + j;
            ok = j < 4 && k < kMlpOut;
        }
        if (ok) val = params[w_off(L) + k * in_dim(L) + i];
        fragB[e - kFragF * kFragElems] = (__bf16)val;
    }
}

// ---------------------------------------------------------------------------
// Inference: persistent grid, forward fragments resident in LDS, 64 queries
// (4 tiles) per wave iteration.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kInferBlock) __attribute__((amdgpu_waves_per_eu(3))) void infer_kernel(const bf16x8 *__restrict__ fragF,
                                                             const float *__restrict__ params,
                                                             const float *__restrict__ raw, uint32_t n,
                                                             float *__restrict__ out, MlpEncoding enc) {
    __shared__ bf16x8 sF[kFragF * 64];
    __shared__ float sBias[5 * kMlpHidden + 4];  // hidden layers' biases, then the output's
    __shared__ float sQ[kInferBlock / 64][16 * kMlpRaw];  // each wave's staged query tile
    for (int i = threadIdx.x; i < kFragF * 64; i += kInferBlock) sF[i] = fragF[i];
    for (int i = threadIdx.x; i < 5 * kMlpHidden + 4; i += kInferBlock)
        sBias[i] = i < 5 * kMlpHidden ? params[b_off(i / kMlpHidden) + i % kMlpHidden]
                                      : (i - 5 * kMlpHidden < kMlpOut ? params[b_off(5) + i - 5 * kMlpHidden] : 0.0f);
    __syncthreads();
    const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
    const uint32_t wave = blockIdx.x * (kInferBlock / 64) + (threadIdx.x >> 6);
    const uint32_t nwaves = gridDim.x * (kInferBlock / 64);
    // the next iteration's queries are loaded while this one computes
    float cur[4][4], nxt[4][4];
#pragma unroll
    for (int t = 0; t < 4; t++) load_tile(raw, wave * 64 + 16 * t, n, lane, cur[t]);
    for (uint32_t base = wave * 64; base < n; base += nwaves * 64) {
        const uint32_t next = base + nwaves * 64;
#pragma unroll
        for (int t = 0; t < 4; t++) load_tile(raw, next + 16 * t, n, lane, nxt[t]);
        bf16x8 B[2][4];
        float *tile = sQ[threadIdx.x >> 6];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            stage_tile(cur[t], enc, lane, tile);
            encode(tile + kMlpRaw * r, g, B[0][t], B[1][t]);
            __builtin_amdgcn_wave_barrier();  // the next tile overwrites the slice
        }
        f32x4 acc[4][4];
#pragma unroll
        for (int L = 0; L < 5; L++) {
#pragma unroll
            for (int m = 0; m < 4; m++) {
                f32x4 bias;
#pragma unroll
                for (int v = 0; v < 4; v++) bias[v] = sBias[kMlpHidden * L + 16 * m + 4 * g + v];
#pragma unroll
                for (int t = 0; t < 4; t++) acc[m][t] = bias;
            }
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    const bf16x8 A = sF[(8 * L + 2 * m + c) * 64 + lane];
#pragma unroll
                    for (int t = 0; t < 4; t++) acc[m][t] = mfma(A, B[c][t], acc[m][t]);
                }
#pragma unroll
            for (int t = 0; t < 4; t++)
#pragma unroll
                for (int c = 0; c < 2; c++)
#pragma unroll
                    for (int j = 0; j < 8; j++) B[c][t][j] = (__bf16)fmaxf(acc[2 * c + (j >> 2)][t][j & 3], 0.0f);
        }
        f32x4 bias5;
#pragma unroll
        for (int v = 0; v < 4; v++) bias5[v] = g == 0 ? sBias[5 * kMlpHidden + v] : 0.0f;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            f32x4 y = bias5;
#pragma unroll
            for (int c = 0; c < 2; c++) y = mfma(sF[(40 + c) * 64 + lane], B[c][t], y);
            const uint32_t q = base + 16 * t + r;
            if (g == 0 && q < n) {
#pragma unroll
                for (int v = 0; v < kMlpOut; v++) out[(size_t)kMlpOut * q + v] = y[v];
            }
        }
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int i = 0; i < 4; i++) cur[t][i] = nxt[t][i];
    }
}

// ---------------------------------------------------------------------------
// Training: forward + backward of 128 samples per workgroup.
// ---------------------------------------------------------------------------
struct TrainArgs {
    const bf16x8 *fragF, *fragB;
    const float *params;
    const float *raw, *target;
    uint32_t n;
    MlpEncoding enc;
    const MlpDevState *state;
    float *partial;         // per workgroup: kMlpParams gradient sums + 1 loss sum
    const uint32_t *n_dev;  // batch size on the device (overrides n when set)
};

__device__ __forceinline__ __bf16 &lds(__bf16 *base, int row, int col) { return base[row * kStride + col]; }

__global__ __launch_bounds__(kTrainBlock) void train_kernel(TrainArgs a) {
    __shared__ __attribute__((aligned(16))) __bf16 sH[6 * 64 * kStride];
    __shared__ __attribute__((aligned(16))) __bf16 sD[2 * 64 * kStride];
    __shared__ float sLoss[kTrainBlock / 64];
    __shared__ float sQ[kTrainBlock / 64][16 * kMlpRaw];  // each wave's staged sample tile
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, g = lane >> 4, r = lane & 15;
    const uint32_t n = a.n_dev ? *a.n_dev : a.n;
    if (blockIdx.x * kTrainSamples >= n) return;  // beyond the batch: no partial
    const uint32_t s0 = blockIdx.x * kTrainSamples + 32 * w;
    const float gscale = a.state->scale / (float)(3u * n);
    int col[2];
    bool valid[2];
    bf16x8 B[2][2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const uint32_t q = s0 + 16 * t + r;
        valid[t] = q < n;
        col[t] = 32 * w + 16 * t + r;
        float v[4];
        load_tile(a.raw, s0 + 16 * t, n, lane, v);
        stage_tile(v, a.enc, lane, sQ[w]);
        encode(sQ[w] + kMlpRaw * r, g, B[0][t], B[1][t]);
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int j = 0; j < 8; j++) lds(sH, 32 * c + 8 * g + j, col[t]) = B[c][t][j];
    }
    // forward
    f32x4 acc[4][2];
#pragma unroll
    for (int L = 0; L < 5; L++) {
#pragma unroll
        for (int m = 0; m < 4; m++) {
            f32x4 bias;
#pragma unroll
            for (int v = 0; v < 4; v++) bias[v] = a.params[b_off(L) + 16 * m + 4 * g + v];
            acc[m][0] = acc[m][1] = bias;
        }
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const bf16x8 A = a.fragF[(8 * L + 2 * m + c) * 64 + lane];
#pragma unroll
                for (int t = 0; t < 2; t++) acc[m][t] = mfma(A, B[c][t], acc[m][t]);
            }
        __bf16 *H = sH + (L + 1) * 64 * kStride;
#pragma unroll
        for (int t = 0; t < 2; t++)
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const __bf16 h = (__bf16)fmaxf(acc[m][t][v], 0.0f);
                    lds(H, 16 * m + 4 * g + v, col[t]) = h;
                    B[m >> 1][t][4 * (m & 1) + v] = h;
                }
    }
    // output layer and loss gradient
    bf16x8 Bd[2][2];
    float lsum = 0.0f;
    {
        f32x4 bias5;
#pragma unroll
        for (int v = 0; v < 4; v++) bias5[v] = (g == 0 && v < kMlpOut) ? a.params[b_off(5) + v] : 0.0f;
        __bf16 *D5 = sD + 64 * kStride;  // delta of layer 5 lives in buffer 5 & 1
#pragma unroll
        for (int t = 0; t < 2; t++) {
            f32x4 y = bias5;
#pragma unroll
            for (int c = 0; c < 2; c++) y = mfma(a.fragF[(40 + c) * 64 + lane], B[c][t], y);
            const uint32_t q = s0 + 16 * t + r;
            Bd[0][t] = zero8();
            Bd[1][t] = zero8();
#pragma unroll
            for (int v = 0; v < 4; v++) {
                float gv = 0.0f;
                if (g == 0 && v < kMlpOut && valid[t]) {
                    const float yv = y[v];
                    const float rr = yv - a.target[(size_t)kMlpOut * q + v];
                    const float den = yv * yv + 0.01f;
                    gv = (2.0f * rr / den) * gscale;
                    lsum += rr * rr / den;
                }
                const __bf16 d = (__bf16)gv;
                lds(D5, 4 * g + v, col[t]) = d;
                Bd[0][t][v] = d;
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_down(lsum, off, 64);
    if (lane == 0) sLoss[w] = lsum;

    float *part = a.partial + (size_t)blockIdx.x * (kMlpParams + 1);
    for (int L = 5; L >= 0; L--) {
        __syncthreads();
        const int in = in_dim(L), out = out_dim(L);
        const __bf16 *D = sD + (L & 1) * 64 * kStride;
        const __bf16 *H = sH + L * 64 * kStride;
        // weight gradient: wave w owns input units 16w..16w+15
        const int nm = (L == 5) ? 1 : 4;
        for (int m = 0; m < nm; m++) {
            f32x4 gw = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int qq = 0; qq < kTrainSamples / 32; qq++) {
                const bf16x8 A = *reinterpret_cast<const bf16x8 *>(D + (16 * m + r) * kStride + 32 * qq + 8 * g);
                const bf16x8 Bh = *reinterpret_cast<const bf16x8 *>(H + (16 * w + r) * kStride + 32 * qq + 8 * g);
                gw = mfma(A, Bh, gw);
            }
            const int i = L == 0 ? enc_feature(16 * w + r) : 16 * w + r;  // layer 0: fragment input -> feature
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int o = 16 * m + 4 * g + v;
                if (o < out && i < in) part[w_off(L) + o * in + i] = gw[v];
            }
        }
        // bias gradient: 4 threads per output unit, 32 samples each
        {
            const int row = tid >> 2, quarter = tid & 3;
            float s = 0.0f;
            for (int k = 0; k < 32; k++) s += (float)D[row * kStride + 32 * quarter + k];
            const float s1 = __shfl_down(s, 1, 64), s2 = __shfl_down(s, 2, 64), s3 = __shfl_down(s, 3, 64);
            if (quarter == 0 && row < out) part[b_off(L) + row] = ((s + s1) + s2) + s3;
        }
        if (L == 0) break;
        // delta of layer L-1 = [H_L > 0] * (W_L^T delta_L), for this wave's samples
        __bf16 *Dn = sD + ((L - 1) & 1) * 64 * kStride;
        f32x4 ad[4][2];
#pragma unroll
        for (int m = 0; m < 4; m++) ad[m][0] = ad[m][1] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        const int nc = (L == 5) ? 1 : 2;
        for (int c = 0; c < nc; c++)
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const int fi = (L == 5) ? 32 + m : 8 * (L - 1) + 2 * m + c;
                const bf16x8 A = a.fragB[fi * 64 + lane];
#pragma unroll
                for (int t = 0; t < 2; t++) ad[m][t] = mfma(A, Bd[c][t], ad[m][t]);
            }
#pragma unroll
        for (int t = 0; t < 2; t++)
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int unit = 16 * m + 4 * g + v;
                    const __bf16 hb = H[unit * kStride + col[t]];
                    unsigned short bits;
                    __builtin_memcpy(&bits, &hb, 2);
                    const __bf16 d = (bits & 0x7fffu) ? (__bf16)ad[m][t][v] : (__bf16)0.0f;
                    lds(Dn, unit, col[t]) = d;
                    Bd[m >> 1][t][4 * (m & 1) + v] = d;
                }
    }
    __syncthreads();
    if (tid == 0) part[kMlpParams] = ((sLoss[0] + sLoss[1]) + sLoss[2]) + sLoss[3];
}

// Sum of the workgroup partials in a fixed order (deterministic): groups of
// kReduceGroup consecutive partials are summed in workgroup order
// (reduce_groups_kernel, one grid row per group), then the group sums in
// group order (reduce_kernel), unscaled; flags non-finite gradients (bad = 1)
// and an empty batch (bad = 2: no step, no scale change).
constexpr int kReduceGroup = 32;

__global__ void reduce_groups_kernel(const float *__restrict__ partial, uint32_t n_host, const uint32_t *n_dev,
                                     float *__restrict__ mid) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kMlpParams) return;
    const uint32_t n = n_dev ? *n_dev : n_host;
    const int nparts = (int)((n + kTrainSamples - 1) / kTrainSamples);
    const int k0 = (int)blockIdx.y * kReduceGroup, k1 = min(k0 + kReduceGroup, nparts);
    float s = 0.0f;
    for (int k = k0; k < k1; k++) s += partial[(size_t)k * (kMlpParams + 1) + p];
    mid[(size_t)blockIdx.y * (kMlpParams + 1) + p] = s;
}

__global__ void reduce_kernel(const float *__restrict__ mid, int ngroups, uint32_t n_host, const uint32_t *n_dev,
                              MlpDevState *st, float *__restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kMlpParams) return;
    const uint32_t n = n_dev ? *n_dev : n_host;
    if (n == 0) {
        if (p == kMlpParams) {
            st->loss = 0.0f;
            st->bad = 2;
        }
        return;
    }
    float s = 0.0f;
    for (int k = 0; k < ngroups; k++) s += mid[(size_t)k * (kMlpParams + 1) + p];
    if (p == kMlpParams) {
        st->loss = s / (float)(3u * n);
        return;
    }
    const float gv = s / st->scale;
    grad[p] = gv;
    if (!isfinite(gv)) atomicOr(&st->bad, 1);
}

__global__ void check_kernel(const float *__restrict__ grad, MlpDevState *st) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kMlpParams && !isfinite(grad[p])) atomicOr(&st->bad, 1);
}

__global__ void adam_kernel(float *__restrict__ params, float *__restrict__ m, float *__restrict__ v,
                            const float *__restrict__ grad, const MlpDevState *st, MlpConfig cfg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpParams || st->bad) return;
    const float p1 = st->b1t * cfg.b1, p2 = st->b2t * cfg.b2;
    const float c1 = 1.0f - p1, c2 = 1.0f - p2;
    const float g = grad[i];
    const float mi = cfg.b1 * m[i] + (1.0f - cfg.b1) * g;
    const float vi = cfg.b2 * v[i] + (1.0f - cfg.b2) * (g * g);
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / c1, vh = vi / c2;
    params[i] = params[i] - cfg.lr * mh / (sqrtf(vh) + cfg.eps);
}

__global__ void finalize_kernel(MlpDevState *st, MlpConfig cfg) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    MlpDevState s = *st;
    if (s.bad == 2) {
        // empty batch: nothing to apply
    } else if (s.bad) {
        s.scale = s.scale * 0.5f;
        s.good = 0;
        s.skipped += 1;
    } else {
        s.b1t = s.b1t * cfg.b1;
        s.b2t = s.b2t * cfg.b2;
        s.step += 1;
        s.good += 1;
        if (s.good >= cfg.growth_interval) {
            if (s.scale < 16777216.0f) s.scale = s.scale * 2.0f;
            s.good = 0;
        }
    }
    s.bad = 0;
    *st = s;
}

__global__ void clear_flag_kernel(MlpDevState *st) {
    if (threadIdx.x == 0 && blockIdx.x == 0) st->bad = 0;
}

uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void pack(Context &ctx, Mlp &net) {
    const int total = (kFragF + kFragB) * kFragElems;
    pack_kernel<<<(total + 255) / 256, 256, 0, ctx.stream>>>(net.params.as<float>(), net.frag_f.as<__bf16>(),
                                                            net.frag_b.as<__bf16>());
    KRT_HIP(hipGetLastError());
}

void apply(Context &ctx, Mlp &net) {
    const int nb = (kMlpParams + 255) / 256;
    adam_kernel<<<nb, 256, 0, ctx.stream>>>(net.params.as<float>(), net.m.as<float>(), net.v.as<float>(),
                                           net.grad.as<float>(), net.state.as<MlpDevState>(), net.cfg);
    finalize_kernel<<<1, 64, 0, ctx.stream>>>(net.state.as<MlpDevState>(), net.cfg);
    KRT_HIP(hipGetLastError());
    pack(ctx, net);
}

}  // namespace

void mlp_init(Context &ctx, Mlp &net, const MlpConfig &cfg, uint64_t seed) {
    net.cfg = cfg;
    net.params.alloc(sizeof(float) * kMlpParams);
    net.m.alloc(sizeof(float) * kMlpParams);
    net.v.alloc(sizeof(float) * kMlpParams);
    net.grad.alloc(sizeof(float) * kMlpParams);
    net.frag_f.alloc(sizeof(__bf16) * kFragF * kFragElems);
    net.frag_b.alloc(sizeof(__bf16) * kFragB * kFragElems);
    net.state.alloc(sizeof(MlpDevState));
    // xavier(n, m) * (u - 0.5) for weights and biases (editor/mlp.cuh:13-31),
    // u from a seeded SplitMix64 stream instead of rand()
    std::vector<float> h(kMlpParams);
    uint64_t s = seed;
    for (int L = 0; L < kMlpLayers; L++) {
        const float a = sqrtf(6.0f / (float)(in_dim(L) + out_dim(L)));
        for (int i = w_off(L); i < w_off(L + 1); i++) {
            const float u = (float)(splitmix64(s) >> 40) * (1.0f / 16777216.0f);
            h[i] = a * (u - 0.5f);
        }
    }
    mlp_set_params(ctx, net, h.data());
    KRT_HIP(hipMemsetAsync(net.m.p, 0, net.m.bytes, ctx.stream));
    KRT_HIP(hipMemsetAsync(net.v.p, 0, net.v.bytes, ctx.stream));
    KRT_HIP(hipMemsetAsync(net.frag_f.p, 0, net.frag_f.bytes, ctx.stream));
    KRT_HIP(hipMemsetAsync(net.frag_b.p, 0, net.frag_b.bytes, ctx.stream));
    MlpDevState st{};
    st.scale = cfg.loss_scale;
    st.b1t = 1.0f;
    st.b2t = 1.0f;
    KRT_HIP(hipMemcpyAsync(net.state.p, &st, sizeof(st), hipMemcpyHostToDevice, ctx.stream));
    pack(ctx, net);
    KRT_HIP(hipStreamSynchronize(ctx.stream));
}

void mlp_set_params(Context &ctx, Mlp &net, const float *host) {
    KRT_HIP(hipMemcpyAsync(net.params.p, host, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, ctx.stream));
    pack(ctx, net);
    KRT_HIP(hipStreamSynchronize(ctx.stream));
}

void mlp_get_params(Context &ctx, Mlp &net, float *host) {
    KRT_HIP(hipMemcpyAsync(host, net.params.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, ctx.stream));
    KRT_HIP(hipStreamSynchronize(ctx.stream));
}

uint32_t mlp_infer_blocks(uint32_t n) {
    uint32_t b = (n + kInferBlock - 1) / kInferBlock;
    const uint32_t cap = 256 * 3;  // three 43 KB-LDS workgroups per CU
    return b < 1 ? 1 : (b < cap ? b : cap);
}

void mlp_infer_dev(Mlp &net, const float *raw, uint32_t n, float *out, hipStream_t stream) {
    if (n == 0) return;
    infer_kernel<<<mlp_infer_blocks(n), kInferBlock, 0, stream>>>(net.frag_f.as<bf16x8>(), net.params.as<float>(),
                                                                   raw, n, out, net.enc);
    KRT_HIP(hipGetLastError());
}

void mlp_infer_dev(Context &ctx, Mlp &net, const float *raw, uint32_t n, float *out) {
    mlp_infer_dev(net, raw, n, out, ctx.stream);
}

void mlp_gradient_dev(Context &ctx, Mlp &net, const float *raw, const float *target, uint32_t n,
                      const uint32_t *n_dev) {
    KRT_REQUIRE(n > 0, KRT_ERR_INVALID, "empty training batch");
    const uint32_t nb = (n + kTrainSamples - 1) / kTrainSamples;
    const size_t need = sizeof(float) * (size_t)nb * (kMlpParams + 1);
    if (net.partial.bytes < need) net.partial.alloc(need);
    TrainArgs a{net.frag_f.as<bf16x8>(), net.frag_b.as<bf16x8>(), net.params.as<float>(), raw, target, n, net.enc,
                net.state.as<MlpDevState>(), net.partial.as<float>(), n_dev};
    train_kernel<<<nb, kTrainBlock, 0, ctx.stream>>>(a);
    const int ngroups = (int)((nb + kReduceGroup - 1) / kReduceGroup);
    const size_t mid_bytes = sizeof(float) * (size_t)ngroups * (kMlpParams + 1);
    if (net.mid.bytes < mid_bytes) net.mid.alloc(mid_bytes);
    reduce_groups_kernel<<<dim3((kMlpParams + 1 + 255) / 256, ngroups), 256, 0, ctx.stream>>>(
        net.partial.as<float>(), n, n_dev, net.mid.as<float>());
    reduce_kernel<<<(kMlpParams + 1 + 255) / 256, 256, 0, ctx.stream>>>(net.mid.as<float>(), ngroups, n, n_dev,
                                                                        net.state.as<MlpDevState>(),
                                                                        net.grad.as<float>());
    KRT_HIP(hipGetLastError());
}

void mlp_apply_dev(Context &ctx, Mlp &net) { apply(ctx, net); }

void mlp_apply_host_gradient(Context &ctx, Mlp &net, const float *grad) {
    KRT_HIP(hipMemcpyAsync(net.grad.p, grad, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, ctx.stream));
    check_kernel<<<(kMlpParams + 255) / 256, 256, 0, ctx.stream>>>(net.grad.as<float>(),
                                                                   net.state.as<MlpDevState>());
    KRT_HIP(hipGetLastError());
    apply(ctx, net);
}

void mlp_clear_flag(Context &ctx, Mlp &net) {
    clear_flag_kernel<<<1, 64, 0, ctx.stream>>>(net.state.as<MlpDevState>());
    KRT_HIP(hipGetLastError());
}

}  // namespace krt
