// mhg_mlp_host.cpp -- host side of the learned limiter: the network object
// (parameters, Adam state, training samples on the device), its training
// loop, and the per-step inference attached to a context.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "mhg_ctx.hpp"
#include "mhg_mlp_kernels.hpp"

using namespace mhg;

struct mhg_mlp {
  hipStream_t stream = nullptr;
  int device = 0;  // the device holding the network's arrays
  std::vector<void*> owned;
  float *prm = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr;  // FP32 master, Adam moments, gradient + loss
  unsigned short* wb = nullptr;                                          // BF16 copy of the parameters
  float* part = nullptr;                                                 // per-block partial gradients
  long part_blocks = 0;
  unsigned short* X = nullptr;  // training samples [rows][128] bf16
  float* T = nullptr;           // targets [rows][kMlpTgt]
  int* perm = nullptr;
  float* losses = nullptr;      // minibatch losses of one epoch
  long losses_cap = 0;
  long rows = 0;
  int t = 0;                    // Adam steps taken
  std::string err;
};

namespace {

int nfail(mhg_mlp* n, int code, const std::string& msg) {
  if (n) n->err = msg;
  return code;
}

#define NHIP(n, x)                                                                                      \
  do {                                                                                                  \
    hipError_t e_ = (x);                                                                                \
    if (e_ != hipSuccess) return nfail(n, MHG_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int nalloc(mhg_mlp* n, T** p, size_t count) {
  void* q = nullptr;
  if (hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return nfail(n, MHG_E_ALLOC, "hipMalloc failed");
  if (hipMemset(q, 0, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return nfail(n, MHG_E_HIP, "hipMemset failed");
  n->owned.push_back(q);
  *p = static_cast<T*>(q);
  return MHG_OK;
}

void nfree(mhg_mlp* n, void* p) {
  if (!p) return;
  auto it = std::find(n->owned.begin(), n->owned.end(), p);
  if (it != n->owned.end()) n->owned.erase(it);
  hipFree(p);
}

int ensure_part(mhg_mlp* n, long blocks) {
  if (blocks <= n->part_blocks) return MHG_OK;
  nfree(n, n->part);
  n->part = nullptr;
  n->part_blocks = 0;
  const int r = nalloc(n, &n->part, (size_t)blocks * (kMlpParams + 1));
  if (!r) n->part_blocks = blocks;
  return r;
}

std::vector<unsigned short> to_bf16(const float* x, size_t count) {
  std::vector<unsigned short> h(count);
  for (size_t i = 0; i < count; ++i) h[i] = f2bf(x[i]);
  return h;
}

}  // namespace

namespace mhg {

// Inference of a context's coefficients from its current state (a step's
// first launch when a network is attached).
template <class R>
void enqueue_learned_alpha(mhg_ctx* c, float* Xout, float* Sout, bool write_alpha) {
  LearnedArgsT<R> A;
  A.rho = bp<R>(c->rho);
  A.e = bp<R>(c->e);
  A.ux = bp<R>(c->ux[c->ucur]);
  A.uy = bp<R>(c->uy[c->ucur]);
  A.mask = c->mask;
  A.env_f = bp<R>(c->env_f);
  A.alpha = write_alpha ? bp<R>(c->alpha) : nullptr;
  A.iface = c->iface_alpha;
  launch_learned_alpha_2d<R>(c->stream, c->P, A, mlp_weights(c->mlp_wb, c->mlp_prm), Xout, Sout);
}
template void enqueue_learned_alpha<double>(mhg_ctx*, float*, float*, bool);
template void enqueue_learned_alpha<float>(mhg_ctx*, float*, float*, bool);

}  // namespace mhg

extern "C" {

int mhg_mlp_create(mhg_mlp** out) {
  if (!out) return MHG_E_ARG;
  *out = nullptr;
  mhg_mlp* n = new mhg_mlp();
  if (hipGetDevice(&n->device) != hipSuccess || hipStreamCreate(&n->stream) != hipSuccess) {
    delete n;
    return MHG_E_HIP;
  }
  int r = nalloc(n, &n->prm, kMlpParams) | nalloc(n, &n->m1, kMlpParams) | nalloc(n, &n->m2, kMlpParams);
  r |= nalloc(n, &n->grad, kMlpParams + 1) | nalloc(n, &n->wb, kMlpParams);
  if (r) {
    mhg_mlp_destroy(n);
    return MHG_E_ALLOC;
  }
  *out = n;
  return MHG_OK;
}

int mhg_mlp_destroy(mhg_mlp* n) {
  if (!n) return MHG_OK;
  if (n->stream) hipStreamSynchronize(n->stream);
  for (void* p : n->owned) hipFree(p);
  if (n->stream) hipStreamDestroy(n->stream);
  delete n;
  return MHG_OK;
}

const char* mhg_mlp_last_error(const mhg_mlp* n) { return n ? n->err.c_str() : "null network"; }

int32_t mhg_mlp_param_count(void) { return kMlpParams; }

int mhg_mlp_set_params(mhg_mlp* n, const float* params) {
  if (!n || !params) return MHG_E_ARG;
  NHIP(n, hipMemcpy(n->prm, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  NHIP(n, hipMemset(n->m1, 0, kMlpParams * sizeof(float)));
  NHIP(n, hipMemset(n->m2, 0, kMlpParams * sizeof(float)));
  launch_mlp_to_bf16(n->stream, n->prm, n->wb);
  NHIP(n, hipStreamSynchronize(n->stream));
  n->t = 0;
  return MHG_OK;
}

int mhg_mlp_get_params(mhg_mlp* n, float* params) {
  if (!n || !params) return MHG_E_ARG;
  NHIP(n, hipStreamSynchronize(n->stream));
  NHIP(n, hipMemcpy(params, n->prm, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
  return MHG_OK;
}

int mhg_mlp_set_data(mhg_mlp* n, const float* X, const float* T, int64_t m) {
  if (!n || !X || !T || m < 1 || m > (1LL << 31) - 1) return MHG_E_ARG;
  NHIP(n, hipStreamSynchronize(n->stream));
  nfree(n, n->X);
  nfree(n, n->T);
  nfree(n, n->perm);
  n->X = nullptr;
  n->T = nullptr;
  n->perm = nullptr;
  n->rows = 0;
  int r = nalloc(n, &n->X, (size_t)m * kMlpIn) | nalloc(n, &n->T, (size_t)m * kMlpTgt) | nalloc(n, &n->perm, (size_t)m);
  if (r) return MHG_E_ALLOC;
  const std::vector<unsigned short> xb = to_bf16(X, (size_t)m * kMlpIn);
  NHIP(n, hipMemcpy(n->X, xb.data(), xb.size() * sizeof(unsigned short), hipMemcpyHostToDevice));
  NHIP(n, hipMemcpy(n->T, T, (size_t)m * kMlpTgt * sizeof(float), hipMemcpyHostToDevice));
  n->rows = m;
  return MHG_OK;
}

int mhg_mlp_forward(mhg_mlp* n, const float* X, int64_t m, float* Y) {
  if (!n || !X || !Y || m < 1) return MHG_E_ARG;
  unsigned short* xd = nullptr;
  float* yd = nullptr;
  int r = nalloc(n, &xd, (size_t)m * kMlpIn) | nalloc(n, &yd, (size_t)m * kMlpOut);
  if (r) return MHG_E_ALLOC;
  const std::vector<unsigned short> xb = to_bf16(X, (size_t)m * kMlpIn);
  hipError_t e = hipMemcpy(xd, xb.data(), xb.size() * sizeof(unsigned short), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    launch_mlp_rows(n->stream, false, xd, nullptr, nullptr, 0, m, mlp_weights(n->wb, n->prm), yd, nullptr, 0.f);
    e = hipStreamSynchronize(n->stream);
  }
  if (e == hipSuccess) e = hipMemcpy(Y, yd, (size_t)m * kMlpOut * sizeof(float), hipMemcpyDeviceToHost);
  nfree(n, xd);
  nfree(n, yd);
  NHIP(n, e);
  return MHG_OK;
}

int mhg_mlp_gradient(mhg_mlp* n, int64_t row0, int64_t m, float* loss, float* grad) {
  if (!n || m < 1 || row0 < 0 || row0 + m > n->rows) return MHG_E_ARG;
  const long nblk = (m + kMlpRows - 1) / kMlpRows;
  int r = ensure_part(n, nblk);
  if (r) return r;
  launch_mlp_rows(n->stream, true, n->X, n->T, nullptr, row0, row0 + m, mlp_weights(n->wb, n->prm), nullptr, n->part,
                  1.f / (float)m);
  launch_mlp_reduce(n->stream, n->part, (int)nblk, n->grad);
  NHIP(n, hipStreamSynchronize(n->stream));
  std::vector<float> g(kMlpParams + 1);
  NHIP(n, hipMemcpy(g.data(), n->grad, g.size() * sizeof(float), hipMemcpyDeviceToHost));
  if (grad) std::memcpy(grad, g.data(), kMlpParams * sizeof(float));
  if (loss) *loss = g[kMlpParams];
  return MHG_OK;
}

// mhg_mlp_gradient with its partial gradients in the fixed per-device array
// of mlp_noalloc_scratch and its result written to the caller's device
// memory: no allocation, no host copy.
int mhg_mlp_gradient_noalloc(mhg_mlp* n, int64_t row0, int64_t m, float* d_out) {
  if (!n || !d_out || m < 1 || row0 < 0 || row0 + m > n->rows) return MHG_E_ARG;
  if (m > kMlpNoallocRows)
    return nfail(n, MHG_E_ARG, "mlp_gradient_noalloc: at most 4096 rows per call");
  float* part = nullptr;
  const int r = mlp_noalloc_scratch(n->device, &part);
  if (r) return nfail(n, r, "mlp_gradient_noalloc: scratch address lookup failed");
  const long nblk = (m + kMlpRows - 1) / kMlpRows;
  launch_mlp_rows(n->stream, true, n->X, n->T, nullptr, row0, row0 + m, mlp_weights(n->wb, n->prm), nullptr, part,
                  1.f / (float)m);
  launch_mlp_reduce(n->stream, part, (int)nblk, d_out);
  NHIP(n, hipGetLastError());
  NHIP(n, hipStreamSynchronize(n->stream));
  return MHG_OK;
}

int mhg_mlp_train(mhg_mlp* n, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* epoch_loss) {
  if (!n || epochs < 1 || batch < 1 || !(lr > 0.f) || n->rows < 1) return MHG_E_ARG;
  const long nb = (n->rows + batch - 1) / batch;
  int r = ensure_part(n, (batch + kMlpRows - 1) / kMlpRows);
  if (r) return r;
  if (nb > n->losses_cap) {
    nfree(n, n->losses);
    n->losses = nullptr;
    n->losses_cap = 0;
    r = nalloc(n, &n->losses, (size_t)nb);
    if (r) return r;
    n->losses_cap = nb;
  }
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  std::vector<int> perm(n->rows);
  std::vector<float> lh(nb);
  const MlpWeights W = mlp_weights(n->wb, n->prm);
  for (int ep = 0; ep < epochs; ++ep) {
    for (long i = 0; i < n->rows; ++i) perm[i] = (int)i;
    std::mt19937_64 rng(seed + 0x9e3779b97f4a7c15ULL * (uint64_t)(ep + 1));
    for (long i = n->rows - 1; i > 0; --i) std::swap(perm[i], perm[rng() % (uint64_t)(i + 1)]);
    NHIP(n, hipMemcpyAsync(n->perm, perm.data(), n->rows * sizeof(int), hipMemcpyHostToDevice, n->stream));
    for (long b = 0; b < nb; ++b) {
      const long r0 = b * batch, r1 = std::min<long>(n->rows, r0 + batch);
      launch_mlp_rows(n->stream, true, n->X, n->T, n->perm, r0, r1, W, nullptr, n->part, 1.f / (float)(r1 - r0));
      launch_mlp_reduce(n->stream, n->part, (int)((r1 - r0 + kMlpRows - 1) / kMlpRows), n->grad);
      n->t += 1;
      const float c1 = 1.f - std::pow(beta1, (float)n->t), c2 = 1.f - std::pow(beta2, (float)n->t);
      launch_mlp_adam(n->stream, n->prm, n->m1, n->m2, n->grad, lr, beta1, beta2, eps, c1, c2, n->wb);
      NHIP(n, hipMemcpyAsync(n->losses + b, n->grad + kMlpParams, sizeof(float), hipMemcpyDeviceToDevice, n->stream));
    }
    NHIP(n, hipStreamSynchronize(n->stream));
    NHIP(n, hipMemcpy(lh.data(), n->losses, nb * sizeof(float), hipMemcpyDeviceToHost));
    double s = 0.;
    for (long b = 0; b < nb; ++b) s += (double)lh[b] * (double)(std::min<long>(n->rows, (b + 1) * batch) - b * batch);
    if (epoch_loss) epoch_loss[ep] = (float)(s / (double)n->rows);
  }
  NHIP(n, hipGetLastError());
  return MHG_OK;
}

int mhg_set_limiter_mlp(mhg_ctx* c, const mhg_mlp* n) {
  if (!c || !n) return MHG_E_ARG;
  if (!c->P.learned) return ctx_fail(c, MHG_E_UNSUPPORTED, "set_limiter_mlp: the context's limiter is not MHG_LIM_LEARNED");
  if (!c->mlp_prm) {
    int r = ctx_dalloc(c, &c->mlp_prm, kMlpParams) | ctx_dalloc(c, &c->mlp_wb, kMlpParams);
    if (r) return MHG_E_ALLOC;
  }
  HIPCHK(c, hipStreamSynchronize(n->stream));
  HIPCHK(c, hipMemcpy(c->mlp_prm, n->prm, kMlpParams * sizeof(float), hipMemcpyDeviceToDevice));
  HIPCHK(c, hipMemcpy(c->mlp_wb, n->wb, kMlpParams * sizeof(unsigned short), hipMemcpyDeviceToDevice));
  c->use_mlp = true;
  ctx_drop_graphs(c);  // captured steps do not run the inference
  return MHG_OK;
}

int mhg_set_learned_interface_alpha(mhg_ctx* c, double alpha) {
  if (!c) return MHG_E_ARG;
  if (!c->P.learned)
    return ctx_fail(c, MHG_E_UNSUPPORTED, "set_learned_interface_alpha: the context's limiter is not MHG_LIM_LEARNED");
  if (!(alpha <= 2.0)) return ctx_fail(c, MHG_E_ARG, "set_learned_interface_alpha: alpha must be <= 2 (< 0: classical)");
  c->iface_alpha = alpha < 0.0 ? -1.0 : alpha;
  ctx_drop_graphs(c);
  return MHG_OK;
}

int mhg_learned_features(mhg_ctx* c, float* X, float* scale) {
  if (!c || !X) return MHG_E_ARG;
  if (!c->initialised) return ctx_fail(c, MHG_E_STATE, "context not initialised");
  if (c->P.dim != 2) return ctx_fail(c, MHG_E_UNSUPPORTED, "learned features: 2D contexts only");
  if (!c->mlp_prm) {  // zero network: the features do not depend on it
    int r = ctx_dalloc(c, &c->mlp_prm, kMlpParams) | ctx_dalloc(c, &c->mlp_wb, kMlpParams);
    if (r) return MHG_E_ALLOC;
  }
  if (c->owner) {
    int r = batch_refresh_view(c);
    if (r) return r;
  }
  const long nc = c->P.nc;
  float *xd = nullptr, *sd = nullptr;
  HIPCHK(c, hipMalloc(&xd, (size_t)nc * kMlpIn * sizeof(float)));
  if (hipMalloc(&sd, (size_t)nc * kMlpChan * sizeof(float)) != hipSuccess) {
    hipFree(xd);
    return ctx_fail(c, MHG_E_ALLOC, "hipMalloc failed");
  }
  if (c->bulk == 4) enqueue_learned_alpha<float>(c, xd, sd, false);
  else enqueue_learned_alpha<double>(c, xd, sd, false);
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(X, xd, (size_t)nc * kMlpIn * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess && scale) e = hipMemcpy(scale, sd, (size_t)nc * kMlpChan * sizeof(float), hipMemcpyDeviceToHost);
  hipFree(xd);
  hipFree(sd);
  HIPCHK(c, e);
  return MHG_OK;
}

int mhg_learned_alpha(mhg_ctx* c, double* alpha) {
  if (!c || !alpha) return MHG_E_ARG;
  if (!c->P.learned) return ctx_fail(c, MHG_E_UNSUPPORTED, "learned_alpha: the context's limiter is not MHG_LIM_LEARNED");
  if (c->use_mlp) {
    if (c->bulk == 4) enqueue_learned_alpha<float>(c, nullptr, nullptr, true);
    else enqueue_learned_alpha<double>(c, nullptr, nullptr, true);
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t n = 2 * (size_t)c->P.nc;
  if (c->bulk == 8) {
    HIPCHK(c, hipMemcpy(alpha, c->alpha, n * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::vector<float> t(n);
    HIPCHK(c, hipMemcpy(t.data(), c->alpha, n * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) alpha[i] = t[i];
  }
  return MHG_OK;
}

}  // extern "C"
