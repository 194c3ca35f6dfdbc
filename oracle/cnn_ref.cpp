/*
 * cnn_ref.cpp — TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of the learned denoiser (no counterpart in the
 * reference): a kernel-predicting 3-level U-Net with 32/64/128 channels and
 * 3x3 convolutions, a 1x1 head predicting 25 logits per pixel, a softmax and
 * the 5x5 per-pixel kernel applied to the noisy radiance, trained on relMSE.
 * Forward and backward are written out loop by loop; finite differences of
 * the loss pin the backward (tests/test_cnn_oracle.py).
 *
 * Conventions shared with the device network:
 *   input: 10 planar channels [radiance 3, albedo 3, normal 3, depth 1] per
 *     image, transformed to [log1p(max(rad,0)), albedo, normal, log1p(depth)];
 *   conv: out[y][x][co] = b[co] + sum_{dy,dx,ci} in[y+dy][x+dx][ci] W[t][ci][co],
 *     t = (dy+1)*3 + (dx+1), zero padding; ReLU after every conv but the head;
 *   pool: 2x2 average; up: nearest; concat: [upsampled, skip]; sizes that
 *     are not a multiple of 4 are padded up to one by repeating the last row
 *     and column (forward only; the backward takes multiples of 4);
 *   kernel: logits -> softmax over k = (dy+2)*5 + (dx+2), applied to the
 *     radiance at the clamped neighbour;
 *   loss: mean over pixels and channels of (x - r)^2 / (r^2 + 1e-2);
 *   parameters: per layer W[t][ci][co] then b[co], layers in the order
 *     e0a e0b e1a e1b ba bb d1a d1b d0a d0b head.
 */
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct Layer { int cin, cout, taps; };
const Layer kLayers[11] = {{10, 32, 9}, {32, 32, 9}, {32, 64, 9}, {64, 64, 9}, {64, 128, 9}, {128, 128, 9},
                           {192, 64, 9}, {64, 64, 9}, {96, 32, 9}, {32, 32, 9}, {32, 25, 1}};

size_t layer_params(const Layer& l) { return (size_t)l.taps * l.cin * l.cout + l.cout; }

struct T {  // H x W x C, channel fastest
    int h = 0, w = 0, c = 0;
    std::vector<double> v;
    T() = default;
    T(int h_, int w_, int c_) : h(h_), w(w_), c(c_), v((size_t)h_ * w_ * c_, 0.0) {}
    double& at(int y, int x, int k) { return v[((size_t)y * w + x) * c + k]; }
    double at(int y, int x, int k) const { return v[((size_t)y * w + x) * c + k]; }
};

// Rows of an output on the host's threads (each row written by one thread,
// so the result does not depend on the thread count).
template <class F>
void for_rows(int n, F f) {
    const int threads = std::max(1, std::min<int>(n, (int)std::thread::hardware_concurrency()));
    if (threads <= 1 || n < 16) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> ts;
    for (int t = 0; t < threads; ++t)
        ts.emplace_back([&]() { for (int i; (i = next.fetch_add(1)) < n;) f(i); });
    for (auto& t : ts) t.join();
}

void conv(const T& in, const double* W, const double* b, const Layer& L, bool relu, T& out) {
    out = T(in.h, in.w, L.cout);
    const int r = L.taps == 9 ? 1 : 0;
    for_rows(in.h, [&](int y) {
        for (int x = 0; x < in.w; ++x)
            for (int co = 0; co < L.cout; ++co) {
                double s = b[co];
                for (int dy = -r; dy <= r; ++dy)
                    for (int dx = -r; dx <= r; ++dx) {
                        const int yy = y + dy, xx = x + dx;
                        if (yy < 0 || yy >= in.h || xx < 0 || xx >= in.w) continue;
                        const int t = L.taps == 9 ? (dy + 1) * 3 + (dx + 1) : 0;
                        const double* wt = W + (size_t)t * L.cin * L.cout;
                        for (int ci = 0; ci < L.cin; ++ci) s += in.at(yy, xx, ci) * wt[(size_t)ci * L.cout + co];
                    }
                out.at(y, x, co) = relu ? (s > 0 ? s : 0) : s;
            }
    });
}

// gradients of conv: given dPre (gradient of the pre-activation), accumulate dW, db, and dIn
void conv_bwd(const T& in, const double* W, const Layer& L, const T& dpre, double* dW, double* db, T* din) {
    const int r = L.taps == 9 ? 1 : 0;
    if (din) *din = T(in.h, in.w, L.cin);
    for (int y = 0; y < in.h; ++y)
        for (int x = 0; x < in.w; ++x)
            for (int co = 0; co < L.cout; ++co) {
                const double g = dpre.at(y, x, co);
                if (g == 0.0) continue;
                db[co] += g;
                for (int dy = -r; dy <= r; ++dy)
                    for (int dx = -r; dx <= r; ++dx) {
                        const int yy = y + dy, xx = x + dx;
                        if (yy < 0 || yy >= in.h || xx < 0 || xx >= in.w) continue;
                        const int t = L.taps == 9 ? (dy + 1) * 3 + (dx + 1) : 0;
                        const size_t wo = (size_t)t * L.cin * L.cout;
                        for (int ci = 0; ci < L.cin; ++ci) {
                            dW[wo + (size_t)ci * L.cout + co] += in.at(yy, xx, ci) * g;
                            if (din) din->at(yy, xx, ci) += W[wo + (size_t)ci * L.cout + co] * g;
                        }
                    }
            }
}

T pool(const T& a) {
    T o(a.h / 2, a.w / 2, a.c);
    for (int y = 0; y < o.h; ++y)
        for (int x = 0; x < o.w; ++x)
            for (int k = 0; k < a.c; ++k)
                o.at(y, x, k) = 0.25 * (a.at(2 * y, 2 * x, k) + a.at(2 * y, 2 * x + 1, k) + a.at(2 * y + 1, 2 * x, k) +
                                        a.at(2 * y + 1, 2 * x + 1, k));
    return o;
}
T pool_bwd(const T& d, int h, int w) {
    T o(h, w, d.c);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int k = 0; k < d.c; ++k) o.at(y, x, k) = 0.25 * d.at(y / 2, x / 2, k);
    return o;
}
T up_concat(const T& lo, const T& skip) {
    T o(skip.h, skip.w, lo.c + skip.c);
    for (int y = 0; y < skip.h; ++y)
        for (int x = 0; x < skip.w; ++x) {
            for (int k = 0; k < lo.c; ++k) o.at(y, x, k) = lo.at(y / 2, x / 2, k);
            for (int k = 0; k < skip.c; ++k) o.at(y, x, lo.c + k) = skip.at(y, x, k);
        }
    return o;
}
void relu_mask(T& g, const T& act) {
    for (size_t i = 0; i < g.v.size(); ++i)
        if (!(act.v[i] > 0)) g.v[i] = 0;
}

struct Net {
    const double* p[11][2];
    explicit Net(const double* params) {
        size_t o = 0;
        for (int l = 0; l < 11; ++l) {
            p[l][0] = params + o;
            o += (size_t)kLayers[l].taps * kLayers[l].cin * kLayers[l].cout;
            p[l][1] = params + o;
            o += kLayers[l].cout;
        }
    }
};

struct Acts {
    T x, e0a, e0b, p0, e1a, e1b, p1, ba, bb, c1, d1a, d1b, c0, d0a, d0b, head;
    std::vector<double> wk;  // softmax weights, 25 per pixel
};

// Round to bfloat16 (round to nearest even, through float), as the device
// stores its weights and activations.
double round_bf16(double v) {
    float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return (double)f;
}
void round_all(T& t) {
    for (double& v : t.v) v = round_bf16(v);
}

// `bf16`: round the input, every stored activation (conv outputs after the
// ReLU, pools) to bfloat16 and the head's logits to float, as the device
// network stores them; the caller passes bf16 weights and float biases.
void forward(const Net& n, const float* in10, int h, int w, Acts& A, std::vector<double>& out_rgb, bool bf16 = false) {
    // the network runs on the image padded to a multiple of 4 by repeating its
    // last row and column; the kernel is applied on the image itself
    const int hp = (h + 3) / 4 * 4, wp = (w + 3) / 4 * 4;
    A.x = T(hp, wp, 10);
    for (int y = 0; y < hp; ++y)
        for (int x = 0; x < wp; ++x)
            for (int k = 0; k < 10; ++k) {
                double v = in10[(size_t)k * h * w + (size_t)std::min(y, h - 1) * w + std::min(x, w - 1)];
                if (k < 3) v = std::log1p(v > 0 ? v : 0.0);
                if (k == 9) v = std::log1p(v);
                A.x.at(y, x, k) = bf16 ? round_bf16((double)(float)v) : v;
            }
    auto R = [&](T& t) { if (bf16) round_all(t); };
    conv(A.x, n.p[0][0], n.p[0][1], kLayers[0], true, A.e0a); R(A.e0a);
    conv(A.e0a, n.p[1][0], n.p[1][1], kLayers[1], true, A.e0b); R(A.e0b);
    A.p0 = pool(A.e0b); R(A.p0);
    conv(A.p0, n.p[2][0], n.p[2][1], kLayers[2], true, A.e1a); R(A.e1a);
    conv(A.e1a, n.p[3][0], n.p[3][1], kLayers[3], true, A.e1b); R(A.e1b);
    A.p1 = pool(A.e1b); R(A.p1);
    conv(A.p1, n.p[4][0], n.p[4][1], kLayers[4], true, A.ba); R(A.ba);
    conv(A.ba, n.p[5][0], n.p[5][1], kLayers[5], true, A.bb); R(A.bb);
    A.c1 = up_concat(A.bb, A.e1b);
    conv(A.c1, n.p[6][0], n.p[6][1], kLayers[6], true, A.d1a); R(A.d1a);
    conv(A.d1a, n.p[7][0], n.p[7][1], kLayers[7], true, A.d1b); R(A.d1b);
    A.c0 = up_concat(A.d1b, A.e0b);
    conv(A.c0, n.p[8][0], n.p[8][1], kLayers[8], true, A.d0a); R(A.d0a);
    conv(A.d0a, n.p[9][0], n.p[9][1], kLayers[9], true, A.d0b); R(A.d0b);
    conv(A.d0b, n.p[10][0], n.p[10][1], kLayers[10], false, A.head);
    if (bf16)
        for (double& v : A.head.v) v = (double)(float)v;
    A.wk.assign((size_t)h * w * 25, 0.0);
    out_rgb.assign((size_t)h * w * 3, 0.0);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            double m = -1e300;
            for (int k = 0; k < 25; ++k) m = std::max(m, A.head.at(y, x, k));
            double s = 0;
            double* wk = &A.wk[((size_t)y * w + x) * 25];
            for (int k = 0; k < 25; ++k) { wk[k] = std::exp(A.head.at(y, x, k) - m); s += wk[k]; }
            for (int k = 0; k < 25; ++k) wk[k] /= s;
            for (int k = 0; k < 25; ++k) {
                const int yy = std::min(std::max(y + k / 5 - 2, 0), h - 1), xx = std::min(std::max(x + k % 5 - 2, 0), w - 1);
                for (int c = 0; c < 3; ++c) out_rgb[((size_t)y * w + x) * 3 + c] += wk[k] * in10[(size_t)c * h * w + (size_t)yy * w + xx];
            }
        }
}

}  // namespace

extern "C" {

uint32_t cnnref_num_params(void) {
    size_t n = 0;
    for (auto& l : kLayers) n += layer_params(l);
    return (uint32_t)n;
}

/* Forward of one image: out_rgb is h*w*3 (packed per pixel). */
void cnnref_forward(const double* params, const float* in10, int h, int w, double* out_rgb) {
    Net n(params);
    Acts A;
    std::vector<double> o;
    forward(n, in10, h, w, A, o);
    std::memcpy(out_rgb, o.data(), o.size() * sizeof(double));
}

/* The same forward at the device's storage precision: weights rounded to
 * bfloat16, biases to float, the input and every stored activation to
 * bfloat16, the logits to float; sums in FP64.  What remains between this
 * and the device is the order and precision of the device's FP32 sums. */
void cnnref_forward_bf16(const double* params, const float* in10, int h, int w, double* out_rgb) {
    std::vector<double> p(cnnref_num_params());
    size_t o = 0;
    for (auto& l : kLayers) {
        const size_t nw = (size_t)l.taps * l.cin * l.cout;
        for (size_t i = 0; i < nw; ++i, ++o) p[o] = round_bf16(params[o]);
        for (int i = 0; i < l.cout; ++i, ++o) p[o] = (double)(float)params[o];
    }
    Net n(p.data());
    Acts A;
    std::vector<double> out;
    forward(n, in10, h, w, A, out, true);
    std::memcpy(out_rgb, out.data(), out.size() * sizeof(double));
}

/* Loss (mean relMSE over the batch's pixels and channels) and its gradient
 * with respect to every parameter (added into grads, FP64). targets packed
 * h*w*3 per image. */
}  // extern "C"

namespace {

// Weights rounded to bfloat16 and biases to float, as the device keeps them.
std::vector<double> device_params(const double* params) {
    std::vector<double> p(cnnref_num_params());
    size_t o = 0;
    for (auto& l : kLayers) {
        const size_t nw = (size_t)l.taps * l.cin * l.cout;
        for (size_t i = 0; i < nw; ++i, ++o) p[o] = round_bf16(params[o]);
        for (int i = 0; i < l.cout; ++i, ++o) p[o] = (double)(float)params[o];
    }
    return p;
}

// `bf16`: the forward at the device's storage precision, and every gradient
// tensor the device stores in bfloat16 rounded where it stores it (logit
// gradients, each layer's input gradient after its ReLU mask, the
// upsampling's sums, the pooling gradients and the skip sums).
double loss_grad(const double* params, const float* in10, const float* targets, int batch, int h, int w,
                 double* grads, bool bf16) {
    if (h % 4 || w % 4) return NAN;  // the backward is written for sizes the pooling divides
    Net n(params);
    auto R = [&](T& t) { if (bf16) round_all(t); };
    const double count = (double)batch * h * w * 3;
    double loss = 0.0;
    std::vector<double> zero(cnnref_num_params(), 0.0);
    if (grads) std::memset(grads, 0, sizeof(double) * zero.size());
    for (int b = 0; b < batch; ++b) {
        const float* in = in10 + (size_t)b * 10 * h * w;
        const float* tg = targets + (size_t)b * h * w * 3;
        Acts A;
        std::vector<double> out;
        forward(n, in, h, w, A, out, bf16);
        T dhead(h, w, 25);
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                const size_t p = (size_t)y * w + x;
                double dout[3];
                for (int c = 0; c < 3; ++c) {
                    const double r = tg[p * 3 + c], d = out[p * 3 + c] - r;
                    loss += d * d / (r * r + 1e-2);
                    dout[c] = 2.0 * d / (r * r + 1e-2) / count;
                }
                // kernel apply + softmax backward
                const double* wk = &A.wk[p * 25];
                double dw[25], dot = 0;
                for (int k = 0; k < 25; ++k) {
                    const int yy = std::min(std::max(y + k / 5 - 2, 0), h - 1), xx = std::min(std::max(x + k % 5 - 2, 0), w - 1);
                    dw[k] = 0;
                    for (int c = 0; c < 3; ++c) dw[k] += dout[c] * in[(size_t)c * h * w + (size_t)yy * w + xx];
                    dot += wk[k] * dw[k];
                }
                for (int k = 0; k < 25; ++k) dhead.at(y, x, k) = wk[k] * (dw[k] - dot);
            }
        R(dhead);
        if (!grads) continue;
        auto G = [&](int l, int k) {
            size_t o = 0;
            for (int i = 0; i < l; ++i) o += layer_params(kLayers[i]);
            return grads + o + (k ? (size_t)kLayers[l].taps * kLayers[l].cin * kLayers[l].cout : 0);
        };
        T d;
        conv_bwd(A.d0b, n.p[10][0], kLayers[10], dhead, G(10, 0), G(10, 1), &d);
        relu_mask(d, A.d0b);
        R(d);
        T d2;
        conv_bwd(A.d0a, n.p[9][0], kLayers[9], d, G(9, 0), G(9, 1), &d2);
        relu_mask(d2, A.d0a);
        R(d2);
        T dc0;
        conv_bwd(A.c0, n.p[8][0], kLayers[8], d2, G(8, 0), G(8, 1), &dc0);
        relu_mask(dc0, A.c0);
        R(dc0);
        // split concat0 = [up(d1b) 64, e0b 32]
        T dd1b(h / 2, w / 2, 64), de0b(h, w, 32);
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                for (int k = 0; k < 64; ++k) dd1b.at(y / 2, x / 2, k) += dc0.at(y, x, k);
                for (int k = 0; k < 32; ++k) de0b.at(y, x, k) = dc0.at(y, x, 64 + k);
            }
        R(dd1b);
        T d3;
        conv_bwd(A.d1a, n.p[7][0], kLayers[7], dd1b, G(7, 0), G(7, 1), &d3);
        relu_mask(d3, A.d1a);
        R(d3);
        T dc1;
        conv_bwd(A.c1, n.p[6][0], kLayers[6], d3, G(6, 0), G(6, 1), &dc1);
        relu_mask(dc1, A.c1);
        R(dc1);
        T dbb(h / 4, w / 4, 128), de1b(h / 2, w / 2, 64);
        for (int y = 0; y < h / 2; ++y)
            for (int x = 0; x < w / 2; ++x) {
                for (int k = 0; k < 128; ++k) dbb.at(y / 2, x / 2, k) += dc1.at(y, x, k);
                for (int k = 0; k < 64; ++k) de1b.at(y, x, k) = dc1.at(y, x, 128 + k);
            }
        R(dbb);
        T d4;
        conv_bwd(A.ba, n.p[5][0], kLayers[5], dbb, G(5, 0), G(5, 1), &d4);
        relu_mask(d4, A.ba);
        R(d4);
        T dp1;
        conv_bwd(A.p1, n.p[4][0], kLayers[4], d4, G(4, 0), G(4, 1), &dp1);
        R(dp1);
        T pb1 = pool_bwd(dp1, h / 2, w / 2);
        for (size_t i = 0; i < de1b.v.size(); ++i) de1b.v[i] += pb1.v[i];
        relu_mask(de1b, A.e1b);
        R(de1b);
        T d5;
        conv_bwd(A.e1a, n.p[3][0], kLayers[3], de1b, G(3, 0), G(3, 1), &d5);
        relu_mask(d5, A.e1a);
        R(d5);
        T dp0;
        conv_bwd(A.p0, n.p[2][0], kLayers[2], d5, G(2, 0), G(2, 1), &dp0);
        R(dp0);
        T pb0 = pool_bwd(dp0, h, w);
        for (size_t i = 0; i < de0b.v.size(); ++i) de0b.v[i] += pb0.v[i];
        relu_mask(de0b, A.e0b);
        R(de0b);
        T d6;
        conv_bwd(A.e0a, n.p[1][0], kLayers[1], de0b, G(1, 0), G(1, 1), &d6);
        relu_mask(d6, A.e0a);
        R(d6);
        conv_bwd(A.x, n.p[0][0], kLayers[0], d6, G(0, 0), G(0, 1), nullptr);
    }
    return loss / count;
}

}  // namespace

extern "C" {

double cnnref_loss_grad(const double* params, const float* in10, const float* targets, int batch, int h, int w,
                        double* grads) {
    return loss_grad(params, in10, targets, batch, h, w, grads, false);
}

/* The same at the device's storage precision (see cnnref_forward_bf16), with
 * the backward's stored gradients rounded to bfloat16 where the device
 * stores them. */
double cnnref_loss_grad_bf16(const double* params, const float* in10, const float* targets, int batch, int h, int w,
                             double* grads) {
    const std::vector<double> p = device_params(params);
    return loss_grad(p.data(), in10, targets, batch, h, w, grads, true);
}

}  // extern "C"
