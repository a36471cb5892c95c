// surrogate.cpp -- C-ABI of the learned local-moment surrogate (include/leo2d.h).
//
// A multilayer perceptron maps the Fourier features of a local stacking
// configuration (interlayer shift, symmetric strain, twist) to the local
// Chebyshev moments of the four sublattices of a bilayer.  Training and
// inference run on the GPU through the BF16 MFMA kernels of mlp.hip; the host
// here only sequences launches, shuffles the training set and initialises the
// weights.  There is no counterpart in the reference: the moments it learns
// come from this library's KPM recursion (leo2d_batch_kpm with target starts).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "devbuf.h"
#include "mlp.h"

using leo2d::check;
using leo2d::dalloc;
using leo2d::DevBuf;
using leo2d::HipError;

struct leo2d_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<int> dims;                 // in, hidden x layers, out
    std::vector<int64_t> woff, boff;       // offsets of W_l and b_l in the flat parameter vector
    int64_t nparams = 0;
    DevBuf P32, P16, G, Mom, Var;          // master weights, bf16 copy, gradient, Adam moments
    int64_t adam_step = 0;
    // batch workspace
    int cap = 0;
    DevBuf X16, Y, T, dY16, dZa, dZb, rowloss, xin, idx;
    std::vector<DevBuf> Z, H;              // per hidden layer: pre-activation FP32, activation bf16

    int layers() const { return (int)dims.size() - 1; }
    int in() const { return dims.front(); }
    int out() const { return dims.back(); }
    ~leo2d_mlp()
    {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

struct ArgError : std::runtime_error {
    explicit ArgError(const std::string& m) : std::runtime_error(m) {}
};

template <typename F>
int guarded(leo2d_mlp* m, F&& f)
{
    try {
        if (m) check(hipSetDevice(m->device), "hipSetDevice");
        return f();
    } catch (const HipError& e) {
        if (m) m->err = e.what();
        return e.code;
    } catch (const ArgError& e) {
        if (m) m->err = e.what();
        return LEO2D_ERR_ARG;
    } catch (const std::bad_alloc&) {
        if (m) m->err = "host allocation failed";
        return LEO2D_ERR_ALLOC;
    } catch (const std::exception& e) {
        if (m) m->err = e.what();
        return LEO2D_ERR_ARG;
    }
}

const uint16_t* bf(const DevBuf& b, int64_t off = 0) { return static_cast<const uint16_t*>(b.p) + off; }
uint16_t* bfw(DevBuf& b, int64_t off = 0) { return static_cast<uint16_t*>(b.p) + off; }
float* f32(DevBuf& b, int64_t off = 0) { return static_cast<float*>(b.p) + off; }

void to_device_f32(DevBuf& b, const std::vector<float>& v, hipStream_t s)
{
    dalloc(b, sizeof(float) * v.size(), "adam corrections");
    check(hipMemcpyAsync(b.p, v.data(), sizeof(float) * v.size(), hipMemcpyHostToDevice, s), "adam corrections");
}

void ensure_capacity(leo2d_mlp* m, int rows)
{
    if (rows <= m->cap) return;
    const int cap = std::max(rows, 256);
    const int hmax = *std::max_element(m->dims.begin() + 1, m->dims.end() - 1);
    dalloc(m->X16, sizeof(uint16_t) * (size_t)cap * m->in(), "mlp inputs");
    dalloc(m->xin, sizeof(float) * (size_t)cap * m->in(), "mlp inputs");
    dalloc(m->Y, sizeof(float) * (size_t)cap * m->out(), "mlp outputs");
    dalloc(m->T, sizeof(float) * (size_t)cap * m->out(), "mlp targets");
    dalloc(m->dY16, sizeof(uint16_t) * (size_t)cap * m->out(), "mlp output gradient");
    dalloc(m->dZa, sizeof(uint16_t) * (size_t)cap * hmax, "mlp gradient");
    dalloc(m->dZb, sizeof(uint16_t) * (size_t)cap * hmax, "mlp gradient");
    dalloc(m->rowloss, sizeof(double) * (size_t)cap, "mlp loss");
    m->Z.clear();
    m->H.clear();
    m->Z.resize(m->layers() - 1);
    m->H.resize(m->layers() - 1);
    for (int l = 0; l + 1 < m->layers(); ++l) {
        dalloc(m->Z[l], sizeof(float) * (size_t)cap * m->dims[l + 1], "mlp pre-activations");
        dalloc(m->H[l], sizeof(uint16_t) * (size_t)cap * m->dims[l + 1], "mlp activations");
    }
    m->cap = cap;
}

// Y[rows][out] = MLP(X16[rows][in]) ; Y may point outside the workspace
void forward(leo2d_mlp* m, int rows, float* Y)
{
    hipStream_t s = m->stream;
    const uint16_t* hin = bf(m->X16);
    for (int l = 0; l < m->layers(); ++l) {
        const int din = m->dims[l], dout = m->dims[l + 1];
        leo2d::GemmArgs g{};
        g.A = hin;
        g.sam = din;
        g.sak = 1;
        g.B = bf(m->P16, m->woff[l]);   // B(k, n) = W[n][k]
        g.sbk = 1;
        g.sbn = din;
        g.M = rows;
        g.N = dout;
        g.K = din;
        g.bias = f32(m->P32, m->boff[l]);
        g.scm = dout;
        if (l + 1 < m->layers()) {
            g.Z = f32(m->Z[l]);
            g.C16 = bfw(m->H[l]);
            check(leo2d::launch_gemm_bf16(g, leo2d::EPI_BIAS_GELU, s), "mlp forward");
            hin = bf(m->H[l]);
        } else {
            g.C32 = Y;
            check(leo2d::launch_gemm_bf16(g, leo2d::EPI_BIAS_F32, s), "mlp forward");
        }
    }
}

// gradient of the mean squared error over rows x out into m->G; row losses in
// row_loss (m->rowloss when null)
void backward(leo2d_mlp* m, int rows, double* row_loss = nullptr)
{
    hipStream_t s = m->stream;
    const int L = m->layers();
    check(leo2d::launch_mse_grad(f32(m->Y), f32(m->T), rows, m->out(), 2.0f / ((float)rows * (float)m->out()),
                                 bfw(m->dY16), row_loss ? row_loss : static_cast<double*>(m->rowloss.p), s),
          "mlp loss");
    const uint16_t* gz = bf(m->dY16);
    DevBuf* next = &m->dZa;
    for (int l = L - 1; l >= 0; --l) {
        const int din = m->dims[l], dout = m->dims[l + 1];
        const uint16_t* hin = l == 0 ? bf(m->X16) : bf(m->H[l - 1]);
        // dW_l[o][i] = sum_b dZ[b][o] Hin[b][i]
        leo2d::GemmArgs gw{};
        gw.A = gz;
        gw.sam = 1;
        gw.sak = dout;
        gw.B = hin;
        gw.sbk = din;
        gw.sbn = 1;
        gw.M = dout;
        gw.N = din;
        gw.K = rows;
        gw.C32 = f32(m->G, m->woff[l]);
        gw.scm = din;
        check(leo2d::launch_gemm_bf16(gw, leo2d::EPI_STORE_F32, s), "mlp weight gradient");
        check(leo2d::launch_colsum(gz, rows, dout, f32(m->G, m->boff[l]), s), "mlp bias gradient");
        if (l == 0) break;
        // dZ_{l-1} = (dZ_l W_l) * gelu'(Z_{l-1})
        leo2d::GemmArgs gx{};
        gx.A = gz;
        gx.sam = dout;
        gx.sak = 1;
        gx.B = bf(m->P16, m->woff[l]);   // B(k=o, n=i) = W[o][i]
        gx.sbk = din;
        gx.sbn = 1;
        gx.M = rows;
        gx.N = din;
        gx.K = dout;
        gx.Z = f32(m->Z[l - 1]);
        gx.C16 = bfw(*next);
        gx.scm = din;
        check(leo2d::launch_gemm_bf16(gx, leo2d::EPI_GELU_BWD, s), "mlp input gradient");
        gz = bf(*next);
        next = next == &m->dZa ? &m->dZb : &m->dZa;
    }
}

void upload_inputs(leo2d_mlp* m, const float* x, int rows)
{
    check(hipMemcpyAsync(m->xin.p, x, sizeof(float) * (size_t)rows * m->in(), hipMemcpyHostToDevice, m->stream),
          "mlp inputs");
    check(leo2d::launch_cast_bf16(f32(m->xin), (int64_t)rows * m->in(), bfw(m->X16), m->stream), "mlp inputs");
}

// Fourier features of one configuration (b1, b2, exx, eyy, exy, theta in degrees):
//   cos, sin 2 pi (p b1 + q b2) over 16 reciprocal vectors (p, q)          32
//   strain / 0.003                                                           3
//   strain x (cos, sin) of the first three harmonics                        18
//   t = (theta - 1.9)/1.1, t^2, 1/theta                                      3
//   t x (cos, sin) of the first three harmonics                              6
//   1, t (exx + eyy)/0.003                                                   2
const int kHarm[16][2] = {{1, 0}, {0, 1}, {1, 1}, {1, -1}, {2, 0}, {0, 2}, {2, 1}, {1, 2},
                          {2, -1}, {1, -2}, {2, 2}, {2, -2}, {3, 0}, {0, 3}, {3, 1}, {1, 3}};

void features_one(const double* c, float* f)
{
    const double two_pi = 6.283185307179586;
    double cs[16], sn[16];
    for (int h = 0; h < 16; ++h) {
        const double ph = two_pi * (kHarm[h][0] * c[0] + kHarm[h][1] * c[1]);
        cs[h] = std::cos(ph);
        sn[h] = std::sin(ph);
    }
    int k = 0;
    for (int h = 0; h < 16; ++h) {
        f[k++] = (float)cs[h];
        f[k++] = (float)sn[h];
    }
    const double e[3] = {c[2] / 0.003, c[3] / 0.003, c[4] / 0.003};
    for (int a = 0; a < 3; ++a) f[k++] = (float)e[a];
    for (int a = 0; a < 3; ++a)
        for (int h = 0; h < 3; ++h) {
            f[k++] = (float)(e[a] * cs[h]);
            f[k++] = (float)(e[a] * sn[h]);
        }
    const double t = (c[5] - 1.9) / 1.1;
    f[k++] = (float)t;
    f[k++] = (float)(t * t);
    f[k++] = (float)(1.0 / c[5]);
    for (int h = 0; h < 3; ++h) {
        f[k++] = (float)(t * cs[h]);
        f[k++] = (float)(t * sn[h]);
    }
    f[k++] = 1.0f;
    f[k++] = (float)(t * (e[0] + e[1]));
}

}  // namespace

extern "C" {

int leo2d_mlp_features(const double* cfg, int64_t n, float* feat)
{
    if (!cfg || !feat || n < 0) return LEO2D_ERR_ARG;
    for (int64_t i = 0; i < n; ++i) {
        if (!(cfg[6 * i + 5] > 0.0)) return LEO2D_ERR_ARG;   // twist in degrees, positive
        features_one(cfg + 6 * i, feat + (size_t)LEO2D_MLP_FEATURES * i);
    }
    return LEO2D_OK;
}

int leo2d_mlp_create(int32_t device, const leo2d_mlp_desc* d, leo2d_mlp** out)
{
    if (!d || !out || d->in <= 0 || d->hidden <= 0 || d->layers < 1 || d->out <= 0) return LEO2D_ERR_ARG;
    *out = nullptr;
    auto* m = new leo2d_mlp;
    m->device = device;
    const int rc = guarded(m, [&]() {
        check(hipStreamCreate(&m->stream), "hipStreamCreate");
        m->dims.push_back(d->in);
        for (int l = 0; l < d->layers; ++l) m->dims.push_back(d->hidden);
        m->dims.push_back(d->out);
        int64_t off = 0;
        for (int l = 0; l < m->layers(); ++l) {
            m->woff.push_back(off);
            off += (int64_t)m->dims[l + 1] * m->dims[l];
            m->boff.push_back(off);
            off += m->dims[l + 1];
        }
        m->nparams = off;
        // Glorot-uniform weights, zero biases
        std::vector<float> p((size_t)off, 0.0f);
        std::mt19937_64 rng(d->seed);
        for (int l = 0; l < m->layers(); ++l) {
            const double a = std::sqrt(6.0 / (m->dims[l] + m->dims[l + 1]));
            std::uniform_real_distribution<double> U(-a, a);
            for (int64_t q = 0; q < (int64_t)m->dims[l + 1] * m->dims[l]; ++q) p[m->woff[l] + q] = (float)U(rng);
        }
        dalloc(m->P32, sizeof(float) * off, "mlp weights");
        dalloc(m->P16, sizeof(uint16_t) * off, "mlp weights");
        dalloc(m->G, sizeof(float) * off, "mlp gradient");
        dalloc(m->Mom, sizeof(float) * off, "mlp optimiser");
        dalloc(m->Var, sizeof(float) * off, "mlp optimiser");
        check(hipMemcpyAsync(m->P32.p, p.data(), sizeof(float) * off, hipMemcpyHostToDevice, m->stream), "mlp init");
        check(leo2d::launch_cast_bf16(f32(m->P32), off, bfw(m->P16), m->stream), "mlp init");
        check(hipMemsetAsync(m->Mom.p, 0, sizeof(float) * off, m->stream), "mlp init");
        check(hipMemsetAsync(m->Var.p, 0, sizeof(float) * off, m->stream), "mlp init");
        check(hipStreamSynchronize(m->stream), "mlp init");
        return LEO2D_OK;
    });
    if (rc != LEO2D_OK) {
        delete m;
        return rc;
    }
    *out = m;
    return LEO2D_OK;
}

void leo2d_mlp_destroy(leo2d_mlp* m)
{
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    delete m;
}

const char* leo2d_mlp_last_error(const leo2d_mlp* m) { return m ? m->err.c_str() : "null network"; }

int64_t leo2d_mlp_num_params(const leo2d_mlp* m) { return m ? m->nparams : 0; }

int leo2d_mlp_get_params(leo2d_mlp* m, float* p)
{
    if (!m || !p) return LEO2D_ERR_ARG;
    return guarded(m, [&]() {
        check(hipMemcpy(p, m->P32.p, sizeof(float) * m->nparams, hipMemcpyDeviceToHost), "mlp params");
        return LEO2D_OK;
    });
}

int leo2d_mlp_set_params(leo2d_mlp* m, const float* p)
{
    if (!m || !p) return LEO2D_ERR_ARG;
    return guarded(m, [&]() {
        check(hipMemcpyAsync(m->P32.p, p, sizeof(float) * m->nparams, hipMemcpyHostToDevice, m->stream), "mlp params");
        check(leo2d::launch_cast_bf16(f32(m->P32), m->nparams, bfw(m->P16), m->stream), "mlp params");
        check(hipMemsetAsync(m->Mom.p, 0, sizeof(float) * m->nparams, m->stream), "mlp params");
        check(hipMemsetAsync(m->Var.p, 0, sizeof(float) * m->nparams, m->stream), "mlp params");
        m->adam_step = 0;
        check(hipStreamSynchronize(m->stream), "mlp params");
        return LEO2D_OK;
    });
}

int leo2d_mlp_forward(leo2d_mlp* m, const float* x, int64_t n, float* y)
{
    if (!m || !x || !y || n < 0) return LEO2D_ERR_ARG;
    return guarded(m, [&]() {
        const int chunk = 4096;
        ensure_capacity(m, (int)std::min<int64_t>(n, chunk));
        for (int64_t r0 = 0; r0 < n; r0 += chunk) {
            const int rows = (int)std::min<int64_t>(chunk, n - r0);
            upload_inputs(m, x + (size_t)r0 * m->in(), rows);
            forward(m, rows, f32(m->Y));
            check(hipMemcpyAsync(y + (size_t)r0 * m->out(), m->Y.p, sizeof(float) * (size_t)rows * m->out(),
                                 hipMemcpyDeviceToHost, m->stream),
                  "mlp outputs");
            check(hipStreamSynchronize(m->stream), "mlp forward");
        }
        return LEO2D_OK;
    });
}

int leo2d_mlp_gradients(leo2d_mlp* m, const float* x, const float* t, int32_t n, double* loss, float* grad)
{
    if (!m || !x || !t || n <= 0) return LEO2D_ERR_ARG;
    return guarded(m, [&]() {
        ensure_capacity(m, n);
        upload_inputs(m, x, n);
        check(hipMemcpyAsync(m->T.p, t, sizeof(float) * (size_t)n * m->out(), hipMemcpyHostToDevice, m->stream),
              "mlp targets");
        forward(m, n, f32(m->Y));
        backward(m, n);
        std::vector<double> rl(n);
        check(hipMemcpyAsync(rl.data(), m->rowloss.p, sizeof(double) * n, hipMemcpyDeviceToHost, m->stream), "loss");
        if (grad)
            check(hipMemcpyAsync(grad, m->G.p, sizeof(float) * m->nparams, hipMemcpyDeviceToHost, m->stream),
                  "mlp gradient");
        check(hipStreamSynchronize(m->stream), "mlp gradients");
        if (loss) *loss = std::accumulate(rl.begin(), rl.end(), 0.0) / ((double)n * m->out());
        return LEO2D_OK;
    });
}

int leo2d_mlp_train(leo2d_mlp* m, const float* x, const float* t, int64_t n, const leo2d_mlp_train_params* p,
                    double* epoch_loss)
{
    if (!m || !x || !t || !p || n <= 0 || p->epochs < 0 || p->batch <= 0) return LEO2D_ERR_ARG;
    return guarded(m, [&]() {
        hipStream_t s = m->stream;
        const int B = (int)std::min<int64_t>(p->batch, n);
        ensure_capacity(m, B);
        DevBuf X, T, idx, losses;
        dalloc(X, sizeof(float) * (size_t)n * m->in(), "training inputs");
        dalloc(T, sizeof(float) * (size_t)n * m->out(), "training targets");
        dalloc(idx, sizeof(int64_t) * (size_t)n, "training order");
        dalloc(losses, sizeof(double) * (size_t)n, "training losses");
        check(hipMemcpyAsync(X.p, x, sizeof(float) * (size_t)n * m->in(), hipMemcpyHostToDevice, s), "training inputs");
        check(hipMemcpyAsync(T.p, t, sizeof(float) * (size_t)n * m->out(), hipMemcpyHostToDevice, s), "training targets");
        std::vector<int64_t> order((size_t)n);
        std::iota(order.begin(), order.end(), (int64_t)0);
        std::mt19937_64 rng(p->seed);
        std::vector<double> rl((size_t)n);
        for (int e = 0; e < p->epochs; ++e) {
            std::shuffle(order.begin(), order.end(), rng);
            check(hipMemcpyAsync(idx.p, order.data(), sizeof(int64_t) * (size_t)n, hipMemcpyHostToDevice, s), "order");
            for (int64_t b0 = 0; b0 < n; b0 += B) {
                const int rows = (int)std::min<int64_t>(B, n - b0);
                check(leo2d::launch_gather(static_cast<const float*>(X.p), m->in(), static_cast<const float*>(T.p),
                                           m->out(), static_cast<const int64_t*>(idx.p) + b0, rows, bfw(m->X16),
                                           f32(m->T), s),
                      "batch gather");
                forward(m, rows, f32(m->Y));
                backward(m, rows);
                check(hipMemcpyAsync(static_cast<double*>(losses.p) + b0, m->rowloss.p, sizeof(double) * rows,
                                     hipMemcpyDeviceToDevice, s),
                      "loss");
                ++m->adam_step;
                check(leo2d::launch_adam(f32(m->P32), f32(m->G), f32(m->Mom), f32(m->Var), m->nparams, p->lr, p->beta1,
                                         p->beta2, p->eps, (int)m->adam_step, bfw(m->P16), s),
                      "adam");
            }
            check(hipMemcpyAsync(rl.data(), losses.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, s), "loss");
            check(hipStreamSynchronize(s), "epoch");
            if (epoch_loss) epoch_loss[e] = std::accumulate(rl.begin(), rl.end(), 0.0) / ((double)n * m->out());
        }
        return LEO2D_OK;
    });
}

// Several networks trained together: each epoch of each network is captured once
// as a graph (gather, forward, backward, Adam for every mini-batch, the Adam step
// read from a device counter) and replayed per epoch on the network's own stream,
// so the networks' short launches run side by side and the host issues one launch
// per network and epoch.  Each network's arithmetic, shuffles and updates are
// those of leo2d_mlp_train with its own seed, bit for bit.
int leo2d_mlp_train_ensemble(leo2d_mlp* const* nets, int32_t count, const float* x, const float* t, int64_t n,
                             const leo2d_mlp_train_params* p, const uint64_t* seeds, double* losses)
{
    if (!nets || count < 1 || !x || !t || !p || !seeds || n <= 0 || p->epochs < 0 || p->batch <= 0)
        return LEO2D_ERR_ARG;
    leo2d_mlp* m0 = nets[0];
    return guarded(m0, [&]() {
        for (int k = 0; k < count; ++k) {
            if (!nets[k]) throw ArgError("null network");
            if (nets[k]->dims != m0->dims || nets[k]->device != m0->device)
                throw ArgError("ensemble members differ in shape or device");
        }
        const int B = (int)std::min<int64_t>(p->batch, n);
        const int64_t steps = (n + B - 1) / B;
        const int64_t total = steps * (int64_t)p->epochs;
        if (total + 1 > 0x7fffffffLL) throw ArgError("too many training steps");
        if (p->epochs == 0) return LEO2D_OK;
        hipStream_t s0 = m0->stream;
        DevBuf X, T;
        dalloc(X, sizeof(float) * (size_t)n * m0->in(), "training inputs");
        dalloc(T, sizeof(float) * (size_t)n * m0->out(), "training targets");
        check(hipMemcpyAsync(X.p, x, sizeof(float) * (size_t)n * m0->in(), hipMemcpyHostToDevice, s0), "training inputs");
        check(hipMemcpyAsync(T.p, t, sizeof(float) * (size_t)n * m0->out(), hipMemcpyHostToDevice, s0),
              "training targets");
        check(hipStreamSynchronize(s0), "training data");
        struct Member {
            DevBuf idx, loss, corr, ctr;
            std::vector<int64_t> order;
            std::vector<double> rl;
            std::mt19937_64 rng;
            hipGraphExec_t exec = nullptr;
        };
        std::vector<Member> mem((size_t)count);
        auto release = [&]() {
            for (auto& q : mem)
                if (q.exec) (void)hipGraphExecDestroy(q.exec);
        };
        try {
            for (int k = 0; k < count; ++k) {
                leo2d_mlp* m = nets[k];
                Member& q = mem[(size_t)k];
                hipStream_t s = m->stream;
                ensure_capacity(m, B);
                dalloc(q.idx, sizeof(int64_t) * (size_t)n, "training order");
                dalloc(q.loss, sizeof(double) * (size_t)n, "training losses");
                dalloc(q.ctr, sizeof(int), "adam step");
                // bias corrections of every step of this call, from the network's own step count
                std::vector<float> corr((size_t)(2 * (total + 1)));
                for (int64_t i = 0; i <= total; ++i)
                    leo2d::adam_corrections(p->beta1, p->beta2, (int)(m->adam_step + 1 + i), &corr[2 * (size_t)i],
                                            &corr[2 * (size_t)i + 1]);
                to_device_f32(q.corr, corr, s);
                const int zero = 0;
                check(hipMemcpyAsync(q.ctr.p, &zero, sizeof(int), hipMemcpyHostToDevice, s), "adam step");
                check(hipStreamSynchronize(s), "training setup");
                q.order.resize((size_t)n);
                std::iota(q.order.begin(), q.order.end(), (int64_t)0);
                q.rng.seed(seeds[k]);
                q.rl.resize((size_t)n);
                // one epoch, captured
                hipGraph_t graph = nullptr;
                check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "epoch capture");
                try {
                    for (int64_t b0 = 0; b0 < n; b0 += B) {
                        const int rows = (int)std::min<int64_t>(B, n - b0);
                        check(leo2d::launch_gather(static_cast<const float*>(X.p), m->in(),
                                                   static_cast<const float*>(T.p), m->out(),
                                                   static_cast<const int64_t*>(q.idx.p) + b0, rows, bfw(m->X16),
                                                   f32(m->T), s),
                              "batch gather");
                        forward(m, rows, f32(m->Y));
                        backward(m, rows, static_cast<double*>(q.loss.p) + b0);
                        check(leo2d::launch_adam_table(f32(m->P32), f32(m->G), f32(m->Mom), f32(m->Var), m->nparams,
                                                       p->lr, p->beta1, p->beta2, p->eps, f32(q.corr),
                                                       static_cast<int*>(q.ctr.p), bfw(m->P16), s),
                              "adam");
                    }
                } catch (...) {
                    (void)hipStreamEndCapture(s, &graph);
                    if (graph) (void)hipGraphDestroy(graph);
                    throw;
                }
                check(hipStreamEndCapture(s, &graph), "epoch capture");
                const hipError_t ie = hipGraphInstantiate(&q.exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                if (ie != hipSuccess) {
                    q.exec = nullptr;
                    check(ie, "epoch graph");
                }
            }
            for (int e = 0; e < p->epochs; ++e) {
                // every member's epoch is queued before any loss is copied back: a copy into
                // pageable host memory returns only when its stream has drained
                for (int k = 0; k < count; ++k) {
                    Member& q = mem[(size_t)k];
                    hipStream_t s = nets[k]->stream;
                    std::shuffle(q.order.begin(), q.order.end(), q.rng);
                    check(hipMemcpyAsync(q.idx.p, q.order.data(), sizeof(int64_t) * (size_t)n, hipMemcpyHostToDevice, s),
                          "order");
                    check(hipGraphLaunch(q.exec, s), "epoch launch");
                }
                for (int k = 0; k < count; ++k) {
                    Member& q = mem[(size_t)k];
                    check(hipMemcpyAsync(q.rl.data(), q.loss.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost,
                                         nets[k]->stream),
                          "loss");
                }
                for (int k = 0; k < count; ++k) {
                    Member& q = mem[(size_t)k];
                    check(hipStreamSynchronize(nets[k]->stream), "epoch");
                    if (losses)
                        losses[(size_t)k * p->epochs + e] =
                            std::accumulate(q.rl.begin(), q.rl.end(), 0.0) / ((double)n * nets[k]->out());
                }
            }
        } catch (...) {
            release();
            throw;
        }
        release();
        for (int k = 0; k < count; ++k) nets[k]->adam_step += total;
        return LEO2D_OK;
    });
}

int leo2d_mlp_ensemble(leo2d_mlp* const* nets, int32_t count, const float* x, int64_t n, int32_t blocks, float* mean,
                       float* spread)
{
    if (!nets || count < 1 || !x || n < 0 || !spread || blocks < 1) return LEO2D_ERR_ARG;
    leo2d_mlp* m0 = nets[0];
    if (m0->out() % blocks != 0) return LEO2D_ERR_ARG;
    for (int e = 0; e < count; ++e)
        if (!nets[e] || nets[e]->dims != m0->dims || nets[e]->device != m0->device) return LEO2D_ERR_ARG;
    return guarded(m0, [&]() {
        const int chunk = 4096;
        const int out = m0->out();
        DevBuf Ys, dmean, dspread;
        dalloc(Ys, sizeof(float) * (size_t)count * chunk * out, "ensemble outputs");
        dalloc(dmean, sizeof(float) * (size_t)chunk * out, "ensemble mean");
        dalloc(dspread, sizeof(float) * (size_t)chunk * blocks, "ensemble spread");
        for (int e = 0; e < count; ++e) ensure_capacity(nets[e], std::min<int64_t>(n, chunk));
        for (int64_t r0 = 0; r0 < n; r0 += chunk) {
            const int rows = (int)std::min<int64_t>(chunk, n - r0);
            for (int e = 0; e < count; ++e) {
                leo2d_mlp* m = nets[e];
                upload_inputs(m, x + (size_t)r0 * m->in(), rows);
                forward(m, rows, static_cast<float*>(Ys.p) + (size_t)e * chunk * out);
                check(hipStreamSynchronize(m->stream), "ensemble forward");
            }
            check(leo2d::launch_ensemble_stats(static_cast<const float*>(Ys.p), count, (int64_t)chunk * out, rows, out,
                                               blocks, f32(dmean), f32(dspread), m0->stream),
                  "ensemble stats");
            if (mean)
                check(hipMemcpyAsync(mean + (size_t)r0 * out, dmean.p, sizeof(float) * (size_t)rows * out,
                                     hipMemcpyDeviceToHost, m0->stream),
                      "ensemble mean");
            check(hipMemcpyAsync(spread + r0 * blocks, dspread.p, sizeof(float) * rows * blocks, hipMemcpyDeviceToHost,
                                 m0->stream),
                  "ensemble spread");
            check(hipStreamSynchronize(m0->stream), "ensemble");
        }
        return LEO2D_OK;
    });
}

}  // extern "C"
