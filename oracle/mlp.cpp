// Radiance-cache MLP oracle — TEST INFRASTRUCTURE ONLY.
//
// The reference's network (editor/mlp.cuh:6-154: Matrix<N, M>, xavier,
// dev_matrix, the forward kernel, RadianceField) is a dense FP32 forward pass
// only, never compiled and never trained.  The network restated here is the
// one the renderer trains online:
//   62 encoded inputs -> 5 x 64 ReLU -> 3 linear outputs, 20,867 parameters,
// each dense layer out[o] = sum_k W[o * in + k] * x[k] + b[o] as in
// mlp.cuh:48-67 (weights row-major by output, bias added after the sum).
//
// Two precisions:
//   FP64   every value in double, no rounding: the model's mathematics.
//   BF16   the rounding points of the HIP kernels: weights and layer inputs
//          rounded to bfloat16 (round to nearest even), each dense layer
//          summed exactly and rounded to float, ReLU, then bfloat16; the
//          backward pass rounds every propagated delta to bfloat16 the same
//          way and scales the output delta by the loss scale.
// Loss: relative L2, (y - t)^2 / (sg(y)^2 + 0.01), averaged over 3 N outputs;
// the denominator is held constant under differentiation.
// Adam: FP32 with the same expression order as the device update.
//
// Pinning (tests/test_oracle_mlp.py): the FP64 gradient against central finite
// differences of the FP64 loss (frozen denominators); the encoding against
// numpy sin/cos; Adam's first step against its closed form
// (p - lr * sign(g)); the BF16 mode against the FP64 mode within bf16 error.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle.h"
#include "rt.hpp"

namespace {

constexpr int kRaw = 13, kFeat = 64, kEnc = 62, kHid = 64, kOut = 3, kLayers = 6;
constexpr int kIn[kLayers] = {kEnc, kHid, kHid, kHid, kHid, kHid};
constexpr int kOutL[kLayers] = {kHid, kHid, kHid, kHid, kHid, kOut};

int w_off(int L) {
    int o = 0;
    for (int i = 0; i < L; i++) o += kIn[i] * kOutL[i] + kOutL[i];
    return o;
}
int b_off(int L) { return w_off(L) + kIn[L] * kOutL[L]; }

float bf16_round(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) {  // inf / nan: truncate, keep nan quiet
        if (u & 0x007fffffu) u |= 0x00400000u;
        u &= 0xffff0000u;
    } else {
        u += 0x7fffu + ((u >> 16) & 1u);
        u &= 0xffff0000u;
    }
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

// One query's 64 features (62 used, 2 zero): float, same expressions as the
// HIP encoder.
void encode(const float *raw, const float *pmin, const float *pscale, float *feat) {
    float pn[9];
    for (int a = 0; a < 3; a++) pn[a] = (raw[a] - pmin[a]) * pscale[a];
    for (int a = 3; a < 9; a++) pn[a] = raw[a];
    for (int p = 0; p < 24; p++) {
        int vec = p < 12 ? 0 : (p < 18 ? 1 : 2);
        int base = vec == 0 ? 0 : (vec == 1 ? 12 : 18);
        int comp = (p - base) % 3, k = (p - base) / 3;
        float u = pn[3 * vec + comp] * (0.5f * (float)(1 << k));
        float c, s;
        orc::sincos_turn(u, c, s);
        feat[2 * p] = s;
        feat[2 * p + 1] = c;
    }
    for (int r = 0; r < 3; r++) feat[48 + r] = pn[r];
    for (int r = 3; r < 13; r++) feat[48 + r] = raw[r];
    feat[61] = 1.0f;
    feat[62] = 0.0f;
    feat[63] = 0.0f;
}

struct Net {
    // weights / biases as the kernels see them
    std::vector<double> W[kLayers], B[kLayers];
    Net(const double *params, bool bf16) {
        for (int L = 0; L < kLayers; L++) {
            W[L].resize((size_t)kIn[L] * kOutL[L]);
            B[L].resize(kOutL[L]);
            for (size_t i = 0; i < W[L].size(); i++) {
                double w = params[w_off(L) + i];
                W[L][i] = bf16 ? (double)bf16_round((float)w) : w;
            }
            for (int o = 0; o < kOutL[L]; o++) {
                double b = params[b_off(L) + o];
                B[L][o] = bf16 ? (double)(float)b : b;
            }
        }
    }
};

// Forward of one query; acts[L] = input of layer L (acts[0] = features).
void forward(const Net &net, const float *feat, bool bf16, double acts[kLayers][kFeat], double y[kOut]) {
    for (int k = 0; k < kFeat; k++) acts[0][k] = bf16 ? (double)bf16_round(feat[k]) : (double)feat[k];
    for (int L = 0; L < kLayers; L++) {
        const int in = kIn[L], out = kOutL[L];
        for (int o = 0; o < out; o++) {
            double z = net.B[L][o];
            for (int k = 0; k < in; k++) z += net.W[L][(size_t)o * in + k] * acts[L][k];
            if (bf16) z = (double)(float)z;
            if (L == kLayers - 1) {
                y[o] = z;
            } else {
                double h = z > 0.0 ? z : 0.0;
                acts[L + 1][o] = bf16 ? (double)bf16_round((float)h) : h;
            }
        }
        if (L < kLayers - 1)
            for (int o = out; o < kFeat; o++) acts[L + 1][o] = 0.0;
    }
}

}  // namespace

extern "C" {

int orc_mlp_param_count(void) { return w_off(kLayers); }

void orc_mlp_encode(const float *raw, int n, const float *pmin, const float *pscale, float *feat) {
    for (int q = 0; q < n; q++) encode(raw + (size_t)kRaw * q, pmin, pscale, feat + (size_t)kFeat * q);
}

void orc_mlp_forward(const double *params, const float *raw, int n, const float *pmin, const float *pscale,
                     int bf16, double *out) {
    Net net(params, bf16 != 0);
    float feat[kFeat];
    double acts[kLayers][kFeat], y[kOut];
    for (int q = 0; q < n; q++) {
        encode(raw + (size_t)kRaw * q, pmin, pscale, feat);
        forward(net, feat, bf16 != 0, acts, y);
        for (int c = 0; c < kOut; c++) out[(size_t)kOut * q + c] = y[c];
    }
}

// Loss with frozen denominators den[3n] (den == NULL: y^2 + 0.01 of this
// forward pass).  FP64 only.
double orc_mlp_loss(const double *params, const float *raw, const float *target, int n, const float *pmin,
                    const float *pscale, const double *den) {
    Net net(params, false);
    float feat[kFeat];
    double acts[kLayers][kFeat], y[kOut], sum = 0.0;
    for (int q = 0; q < n; q++) {
        encode(raw + (size_t)kRaw * q, pmin, pscale, feat);
        forward(net, feat, false, acts, y);
        for (int c = 0; c < kOut; c++) {
            double r = y[c] - (double)target[(size_t)kOut * q + c];
            double d = den ? den[(size_t)kOut * q + c] : y[c] * y[c] + 0.01;
            sum += r * r / d;
        }
    }
    return sum / (3.0 * n);
}

// Gradient of the mean loss over the n queries with respect to every
// parameter (canonical order), divided back by the loss scale in BF16 mode.
// Returns the loss.
double orc_mlp_grad(const double *params, const float *raw, const float *target, int n, const float *pmin,
                    const float *pscale, int bf16, float loss_scale, double *grad) {
    const bool h = bf16 != 0;
    Net net(params, h);
    const int P = w_off(kLayers);
    std::fill(grad, grad + P, 0.0);
    float feat[kFeat];
    double acts[kLayers][kFeat], y[kOut], delta[kFeat], prev[kFeat], loss = 0.0;
    const float gscale = loss_scale / (float)(3 * n);
    for (int q = 0; q < n; q++) {
        encode(raw + (size_t)kRaw * q, pmin, pscale, feat);
        forward(net, feat, h, acts, y);
        for (int c = 0; c < kOut; c++) {
            const float t = target[(size_t)kOut * q + c];
            if (h) {
                float yf = (float)y[c];
                float r = yf - t;
                float den = yf * yf + 0.01f;
                delta[c] = (double)bf16_round((2.0f * r / den) * gscale);
                loss += (double)(r * r / den);
            } else {
                double r = y[c] - (double)t;
                double den = y[c] * y[c] + 0.01;
                delta[c] = 2.0 * r / den / (3.0 * n);
                loss += r * r / den;
            }
        }
        for (int L = kLayers - 1; L >= 0; L--) {
            const int in = kIn[L], out = kOutL[L];
            for (int o = 0; o < out; o++) {
                grad[b_off(L) + o] += delta[o];
                for (int k = 0; k < in; k++) grad[w_off(L) + (size_t)o * in + k] += delta[o] * acts[L][k];
            }
            if (L == 0) break;
            for (int k = 0; k < in; k++) {
                double s = 0.0;
                for (int o = 0; o < out; o++) s += net.W[L][(size_t)o * in + k] * delta[o];
                bool on = acts[L][k] > 0.0;
                if (h) s = (double)bf16_round((float)s);
                prev[k] = on ? s : 0.0;
            }
            std::memcpy(delta, prev, sizeof(double) * in);
        }
    }
    if (h)
        for (int p = 0; p < P; p++) grad[p] /= (double)loss_scale;
    return loss / (3.0 * n);
}

// One Adam step in FP32 (device expression order).  b1t / b2t hold beta^t of
// the previous step and are advanced here.
void orc_mlp_adam(float *params, float *m, float *v, const float *grad, int n, float lr, float b1, float b2,
                  float eps, float *b1t, float *b2t) {
    const float p1 = *b1t * b1, p2 = *b2t * b2;
    const float c1 = 1.0f - p1, c2 = 1.0f - p2;
    for (int i = 0; i < n; i++) {
        const float g = grad[i];
        const float mi = b1 * m[i] + (1.0f - b1) * g;
        const float vi = b2 * v[i] + (1.0f - b2) * (g * g);
        m[i] = mi;
        v[i] = vi;
        const float mh = mi / c1, vh = vi / c2;
        params[i] = params[i] - lr * mh / (sqrtf(vh) + eps);
    }
    *b1t = p1;
    *b2t = p2;
}

}  // extern "C"
