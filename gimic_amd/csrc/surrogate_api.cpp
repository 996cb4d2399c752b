// C-ABI for the learned surrogate (include/gimic_hip.h, gimic_mlp_*): owns the
// FP32 master parameters, their BF16 copies, Adam state and work buffers,
// and sequences the kernels of surrogate.hip for training and inference.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gimic_hip.h"

namespace gimic {
hipError_t mlp_gemm(const __bf16 *A, long long sam, long long sak, const __bf16 *B, long long sbk,
                    long long sbn, float *C, long long ldc, int M, int N, int K, int ksplit,
                    hipStream_t s);
hipError_t mlp_features(const double *r, long long npts, const float *Bfreq, int F,
                        const float *cs, __bf16 *X, int K0, hipStream_t s);
hipError_t mlp_bias_silu(float *Z, const float *b, __bf16 *H, long long total, int N,
                         hipStream_t s);
hipError_t mlp_loss(const float *Z, const float *b, const double *target, const float *out_scale,
                    float eps2, long long npts, int NO, float *dY, __bf16 *dYb, float *loss,
                    hipStream_t s);
hipError_t mlp_dsilu(const float *dH, const float *Z, float *dZ, __bf16 *dZb, long long total,
                     hipStream_t s);
hipError_t mlp_colsum(const float *dZ, long long npts, int N, float *db, hipStream_t s);
hipError_t mlp_adam(float *p, const float *grad, float *m, float *v, long long n, float lr,
                    float b1, float b2, float eps, float bc1, float bc2, __bf16 *pb,
                    hipStream_t s);
hipError_t mlp_cast(const float *p, __bf16 *pb, long long n, hipStream_t s);
hipError_t mlp_output(const float *Z, const float *b, const float *out_scale, long long npts,
                      int NO, double *J, hipStream_t s);
hipError_t mlp_infer(const double *r, long long npts, const float *Bfreq, int F, const float *cs,
                     const __bf16 *const *W, const float *const *bias, const float *out_scale,
                     double *J, hipStream_t s);
std::string &error_slot();
}  // namespace gimic

using namespace gimic;

namespace {

constexpr int kF = 64, kK0 = 160, kH = 128, kD = 4, kNO = 16;

struct MlpError : std::runtime_error {
    int code;
    MlpError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void chk(hipError_t e, const char *what) {
    if (e != hipSuccess)
        throw MlpError(e == hipErrorOutOfMemory ? GIMIC_ENOMEM : GIMIC_EDEVICE,
                       std::string(what) + ": " + hipGetErrorString(e));
}

struct Buf {
    void *p = nullptr;
    size_t n = 0;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    void need(size_t bytes) {
        if (bytes <= n && p) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        chk(hipMalloc(&p, bytes ? bytes : 256), "hipMalloc");
        n = bytes;
    }
    template <class T>
    T *as() const { return (T *)p; }
};

}  // namespace

struct gimic_mlp {
    hipStream_t stream = nullptr;     // own stream
    hipStream_t user = nullptr;       // caller's stream, when set
    long long np = 0;                 // parameter count
    long long woff[kD + 1], boff[kD + 1];
    int kin[kD + 1], nout[kD + 1];
    Buf params, params_b, grad, adam_m, adam_v, freq, cs, oscale, ptrs;
    Buf X[kD + 1], Z[kD + 1], dH, dZ, dZb, dY, dYb, loss;
    long long cap = 0;  // batch capacity of the work buffers
    long long step = 0;
    ~gimic_mlp() {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }
};

namespace {

std::string &errstr() { return error_slot(); }

hipStream_t work_stream(const gimic_mlp *n) { return n->user ? n->user : n->stream; }

template <class F>
int guard(F &&f) {
    try {
        f();
        errstr().clear();
        return GIMIC_OK;
    } catch (const MlpError &e) {
        errstr() = e.what();
        return e.code;
    } catch (const std::exception &e) {
        errstr() = e.what();
        return GIMIC_EINVAL;
    }
}

void reserve(gimic_mlp *n, long long P) {
    if (P <= n->cap) return;
    for (int l = 0; l <= kD; l++) {
        n->X[l].need((size_t)P * n->kin[l] * 2);
        n->Z[l].need((size_t)P * n->nout[l] * 4);
    }
    n->dH.need((size_t)P * kK0 * 4);
    n->dZ.need((size_t)P * kH * 4);
    n->dZb.need((size_t)P * kH * 2);
    n->dY.need((size_t)P * kNO * 4);
    n->dYb.need((size_t)P * kNO * 2);
    n->cap = P;
}

// forward pass of a batch (training form: activations kept)
void forward(gimic_mlp *n, long long P, const double *d_r, hipStream_t s) {
    float *prm = n->params.as<float>();
    __bf16 *pb = n->params_b.as<__bf16>();
    chk(mlp_features(d_r, P, n->freq.as<float>(), kF, n->cs.as<float>(), n->X[0].as<__bf16>(), kK0,
                     s), "features");
    for (int l = 0; l <= kD; l++) {
        const int K = n->kin[l], N = n->nout[l];
        chk(mlp_gemm(n->X[l].as<__bf16>(), K, 1, pb + n->woff[l], 1, K, n->Z[l].as<float>(), N,
                     (int)P, N, K, 1, s), "gemm fwd");
        if (l < kD)
            chk(mlp_bias_silu(n->Z[l].as<float>(), prm + n->boff[l], n->X[l + 1].as<__bf16>(),
                              P * N, N, s), "bias_silu");
    }
}

void backward(gimic_mlp *n, long long P, const double *d_target, float eps2, hipStream_t s) {
    float *prm = n->params.as<float>();
    __bf16 *pb = n->params_b.as<__bf16>();
    float *gr = n->grad.as<float>();
    chk(hipMemsetAsync(gr, 0, (size_t)n->np * 4, s), "memset grad");
    chk(hipMemsetAsync(n->loss.p, 0, 4, s), "memset loss");
    chk(mlp_loss(n->Z[kD].as<float>(), prm + n->boff[kD], d_target, n->oscale.as<float>(), eps2, P,
                 kNO, n->dY.as<float>(), n->dYb.as<__bf16>(), n->loss.as<float>(), s), "loss");
    const int ks = (int)std::max<long long>(1, std::min<long long>(64, P / 1024));
    // output layer
    {
        const int l = kD, K = n->kin[l], N = n->nout[l];
        chk(mlp_gemm(n->dYb.as<__bf16>(), 1, N, n->X[l].as<__bf16>(), K, 1, gr + n->woff[l], K, N,
                     K, (int)P, ks, s), "gemm dW out");
        chk(mlp_colsum(n->dY.as<float>(), P, N, gr + n->boff[l], s), "colsum");
        chk(mlp_gemm(n->dYb.as<__bf16>(), N, 1, pb + n->woff[l], K, 1, n->dH.as<float>(), K,
                     (int)P, K, N, 1, s), "gemm dX out");
    }
    for (int l = kD - 1; l >= 0; l--) {
        const int K = n->kin[l], N = n->nout[l];
        chk(mlp_dsilu(n->dH.as<float>(), n->Z[l].as<float>(), n->dZ.as<float>(),
                      n->dZb.as<__bf16>(), P * N, s), "dsilu");
        chk(mlp_gemm(n->dZb.as<__bf16>(), 1, N, n->X[l].as<__bf16>(), K, 1, gr + n->woff[l], K, N,
                     K, (int)P, ks, s), "gemm dW");
        chk(mlp_colsum(n->dZ.as<float>(), P, N, gr + n->boff[l], s), "colsum");
        if (l > 0)
            chk(mlp_gemm(n->dZb.as<__bf16>(), N, 1, pb + n->woff[l], K, 1, n->dH.as<float>(), K,
                         (int)P, K, N, 1, s), "gemm dX");
    }
}

}  // namespace

extern "C" {

int gimic_mlp_create(gimic_mlp **out, double sigma, const double *center_scale,
                     const double *out_scale, uint64_t seed) {
    return guard([&] {
        if (!out || !center_scale || !out_scale) throw MlpError(GIMIC_EINVAL, "null argument");
        auto *n = new gimic_mlp();
        try {
            chk(hipStreamCreate(&n->stream), "hipStreamCreate");
            long long off = 0;
            for (int l = 0; l <= kD; l++) {
                n->kin[l] = l == 0 ? kK0 : kH;
                n->nout[l] = l == kD ? kNO : kH;
                n->woff[l] = off;
                off += (long long)n->kin[l] * n->nout[l];
                n->boff[l] = off;
                off += n->nout[l];
            }
            n->np = off;
            // He-style initialisation, zero biases; frequencies ~ N(0, sigma^2)
            std::mt19937_64 rng(seed);
            std::normal_distribution<float> nd(0.f, 1.f);
            std::vector<float> p((size_t)off, 0.f);
            for (int l = 0; l <= kD; l++) {
                const float sc = std::sqrt((l == kD ? 1.f : 2.f) / (float)n->kin[l]);
                for (long long i = 0; i < (long long)n->kin[l] * n->nout[l]; i++)
                    p[n->woff[l] + i] = sc * nd(rng);
            }
            std::vector<float> fr(3 * kF);
            for (auto &x : fr) x = (float)sigma * nd(rng);
            float cs[4] = {(float)center_scale[0], (float)center_scale[1], (float)center_scale[2],
                           (float)center_scale[3]};
            float os[9];
            for (int i = 0; i < 9; i++) os[i] = (float)out_scale[i];
            hipStream_t s = work_stream(n);
            n->params.need(off * 4);
            n->params_b.need(off * 2);
            n->grad.need(off * 4);
            n->adam_m.need(off * 4);
            n->adam_v.need(off * 4);
            n->freq.need(fr.size() * 4);
            n->cs.need(16);
            n->oscale.need(36);
            n->loss.need(4);
            chk(hipMemcpyAsync(n->params.p, p.data(), off * 4, hipMemcpyHostToDevice, s), "copy");
            chk(hipMemsetAsync(n->adam_m.p, 0, off * 4, s), "memset");
            chk(hipMemsetAsync(n->adam_v.p, 0, off * 4, s), "memset");
            chk(hipMemcpyAsync(n->freq.p, fr.data(), fr.size() * 4, hipMemcpyHostToDevice, s), "copy");
            chk(hipMemcpyAsync(n->cs.p, cs, 16, hipMemcpyHostToDevice, s), "copy");
            chk(hipMemcpyAsync(n->oscale.p, os, 36, hipMemcpyHostToDevice, s), "copy");
            chk(mlp_cast(n->params.as<float>(), n->params_b.as<__bf16>(), off, s), "cast");
            // device table of layer pointers for the fused inference kernel
            std::vector<void *> tab(2 * (kD + 1));
            for (int l = 0; l <= kD; l++) {
                tab[l] = n->params_b.as<__bf16>() + n->woff[l];
                tab[kD + 1 + l] = n->params.as<float>() + n->boff[l];
            }
            n->ptrs.need(tab.size() * sizeof(void *));
            chk(hipMemcpyAsync(n->ptrs.p, tab.data(), tab.size() * sizeof(void *),
                               hipMemcpyHostToDevice, s), "copy");
            chk(hipStreamSynchronize(s), "sync");
        } catch (...) {
            delete n;
            throw;
        }
        *out = n;
    });
}

int gimic_mlp_destroy(gimic_mlp *n) {
    return guard([&] { delete n; });
}

int gimic_mlp_set_stream(gimic_mlp *n, void *stream) {
    return guard([&] {
        if (!n) throw MlpError(GIMIC_EINVAL, "null network");
        chk(hipStreamSynchronize(work_stream(n)), "sync");
        n->user = (hipStream_t)stream;
    });
}

int gimic_mlp_info(const gimic_mlp *n, long long *nparams, float *freqs) {
    return guard([&] {
        if (!n) throw MlpError(GIMIC_EINVAL, "null network");
        if (nparams) *nparams = n->np;
        if (freqs) chk(hipMemcpy(freqs, n->freq.p, 3 * kF * 4, hipMemcpyDeviceToHost), "copy");
    });
}

int gimic_mlp_params(gimic_mlp *n, float *host, int set) {
    return guard([&] {
        if (!n || !host) throw MlpError(GIMIC_EINVAL, "null argument");
        hipStream_t s = work_stream(n);
        if (set) {
            chk(hipMemcpyAsync(n->params.p, host, n->np * 4, hipMemcpyHostToDevice, s), "copy");
            chk(mlp_cast(n->params.as<float>(), n->params_b.as<__bf16>(), n->np, s), "cast");
        } else {
            chk(hipMemcpyAsync(host, n->params.p, n->np * 4, hipMemcpyDeviceToHost, s), "copy");
        }
        chk(hipStreamSynchronize(s), "sync");
    });
}

int gimic_mlp_loss_grad(gimic_mlp *n, long long npts, const double *d_r, const double *d_target,
                        double eps, double *loss, float *host_grad) {
    return guard([&] {
        if (!n || npts <= 0 || !d_r || !d_target) throw MlpError(GIMIC_EINVAL, "bad argument");
        hipStream_t s = work_stream(n);
        reserve(n, npts);
        forward(n, npts, d_r, s);
        backward(n, npts, d_target, (float)(eps * eps), s);
        float l = 0.f;
        chk(hipMemcpyAsync(&l, n->loss.p, 4, hipMemcpyDeviceToHost, s), "copy");
        if (host_grad)
            chk(hipMemcpyAsync(host_grad, n->grad.p, n->np * 4, hipMemcpyDeviceToHost, s), "copy");
        chk(hipStreamSynchronize(s), "sync");
        if (loss) *loss = l;
    });
}

int gimic_mlp_train_step(gimic_mlp *n, long long npts, const double *d_r, const double *d_target,
                         double eps, double lr, double *loss) {
    return guard([&] {
        if (!n || npts <= 0 || !d_r || !d_target) throw MlpError(GIMIC_EINVAL, "bad argument");
        hipStream_t s = work_stream(n);
        reserve(n, npts);
        forward(n, npts, d_r, s);
        backward(n, npts, d_target, (float)(eps * eps), s);
        n->step++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float bc1 = 1.f - std::pow(b1, (float)n->step), bc2 = 1.f - std::pow(b2, (float)n->step);
        chk(mlp_adam(n->params.as<float>(), n->grad.as<float>(), n->adam_m.as<float>(),
                     n->adam_v.as<float>(), n->np, (float)lr, b1, b2, 1e-8f, bc1, bc2,
                     n->params_b.as<__bf16>(), s), "adam");
        if (loss) {
            float l = 0.f;
            chk(hipMemcpyAsync(&l, n->loss.p, 4, hipMemcpyDeviceToHost, s), "copy");
            chk(hipStreamSynchronize(s), "sync");
            *loss = l;
        }
    });
}

int gimic_mlp_predict(gimic_mlp *n, long long npts, const double *d_r, double *d_J, void *stream) {
    return guard([&] {
        if (!n || npts < 0 || (npts > 0 && (!d_r || !d_J))) throw MlpError(GIMIC_EINVAL, "bad argument");
        hipStream_t s = stream ? (hipStream_t)stream : work_stream(n);
        const __bf16 *const *W = (const __bf16 *const *)n->ptrs.p;
        const float *const *B = (const float *const *)((void **)n->ptrs.p + kD + 1);
        chk(mlp_infer(d_r, npts, n->freq.as<float>(), kF, n->cs.as<float>(), W, B,
                      n->oscale.as<float>(), d_J, s), "infer");
    });
}

int gimic_mlp_predict_unfused(gimic_mlp *n, long long npts, const double *d_r, double *d_J) {
    return guard([&] {
        if (!n || npts <= 0 || !d_r || !d_J) throw MlpError(GIMIC_EINVAL, "bad argument");
        hipStream_t s = work_stream(n);
        reserve(n, npts);
        forward(n, npts, d_r, s);
        chk(mlp_output(n->Z[kD].as<float>(), n->params.as<float>() + n->boff[kD],
                       n->oscale.as<float>(), npts, kNO, d_J, s), "output");
        chk(hipStreamSynchronize(s), "sync");
    });
}

}  // extern "C"
