// mlp.hip -- the surrogate network: fully connected layers with bf16
// weights (fp32 master copy for Adam) and fp32 activations, trained on the
// device by hand-written forward, backward and Adam kernels.  Every GEMM
// runs on the matrix cores with bf16 operands (activations and gradients
// rounded to nearest even as they enter it) and fp32 accumulation.
//
// Layer l: Z_l = A_{l-1} W_l^T + b_l, A_l = gelu(Z_l) on hidden layers,
// A_L = Z_L.  gelu(x) = 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))).
// Loss = mean over the batch of 0.5 (A_L - target)^2.  Adam follows Kingma
// & Ba (2015), Algorithm 1 (bias-corrected moments).
//
// GEMMs: 64x64 output tiles, 256 threads (four waves of 2x2 16x16 MFMA
// blocks), K in steps of 32 staged through LDS as bf16.  Weight gradients
// reduce over the batch in row splits whose partials are summed in order;
// bias gradients (column sums of the fp32 gradients) likewise, so every
// reduction has a fixed order and training is deterministic.
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "mercury_hip.h"
#include "mh_internal.h"
#include "mh_nn_kernels.h"

// Scratch of mh_mlp_forward_noalloc: the hidden activations of one call, in
// two halves that alternate between layers.  Every device holds its own
// copy (MH_MLP_SCRATCH_FLOATS floats, 16 MiB) until the process ends.  In
// host code this name stands for a host-side shadow, not a device address:
// kernels take the address hipGetSymbolAddress gives on each device.
static __device__ __align__(256) float scratch[MH_MLP_SCRATCH_FLOATS];

struct mh_mlp {
  int device = 0;
  float* scratch = nullptr;   // device address of ::scratch on `device`
  std::vector<int> sizes;     // n_layers + 1
  std::vector<size_t> w_off;  // offset of W_l in the parameter vector; b_l follows W_l
  size_t n_params = 0;
  Dev<float> params, grads, m, v;
  Dev<__hip_bfloat16> w16;    // bf16 copy of the whole parameter vector (weights used)
  std::vector<Dev<float>> Z, A, dA;  // per layer activations / gradients
  Dev<float> x, t, dY, terms, data_x, data_t;
  Dev<float> epoch_terms;     // loss terms of every batch of a training run's last epoch
  Dev<float> part;            // split-K partials of the weight and bias gradients
  Dev<int> idx;
  int batch_cap = 0;
  int adam_t = 0;
  hipStream_t stream = nullptr;
  ~mh_mlp() {
    if (stream) (void)hipStreamDestroy(stream);
  }
  int L() const { return static_cast<int>(sizes.size()) - 1; }
  size_t b_off(int l) const { return w_off[l] + static_cast<size_t>(sizes[l + 1]) * sizes[l]; }
};

namespace {

// weight and bias gradients reduce over the batch in up to this many row
// splits (at least 64 rows each), summed in order
constexpr int kMlpSplit = 8;

void ensure_batch(mh_mlp* net, int B) {
  if (B <= net->batch_cap) return;
  const int L = net->L();
  net->Z.resize(L);
  net->A.resize(L);
  net->dA.resize(L + 1);
  for (int l = 0; l < L; l++) {
    net->Z[l].ensure(static_cast<size_t>(B) * net->sizes[l + 1]);
    net->A[l].ensure(static_cast<size_t>(B) * net->sizes[l + 1]);
  }
  for (int l = 0; l <= L; l++) net->dA[l].ensure(static_cast<size_t>(B) * net->sizes[l]);
  net->x.ensure(static_cast<size_t>(B) * net->sizes[0]);
  net->t.ensure(static_cast<size_t>(B) * net->sizes[L]);
  net->dY.ensure(static_cast<size_t>(B) * net->sizes[L]);
  net->terms.ensure(static_cast<size_t>(B) * net->sizes[L]);
  net->batch_cap = B;
}

// Forward of a batch already in net->x (or `in`): activations in A[l].
void forward(mh_mlp* net, const float* in, int B) {
  const int L = net->L();
  const float* prev = in;
  for (int l = 0; l < L; l++) {
    const int ni = net->sizes[l], no = net->sizes[l + 1];
    const __hip_bfloat16* W = net->w16.p + net->w_off[l];
    const float* b = net->params.p + net->b_off(l);
    if (l < L - 1)
      hipLaunchKernelGGL((gemm<float, __hip_bfloat16, false, true, 1>), grid_for(B, no), dim3(256), 0,
                         net->stream, prev, ni, W, ni, net->A[l].p, no, B, no, ni, b, net->Z[l].p,
                         static_cast<const float*>(nullptr));
    else
      hipLaunchKernelGGL((gemm<float, __hip_bfloat16, false, true, 2>), grid_for(B, no), dim3(256), 0,
                         net->stream, prev, ni, W, ni, net->A[l].p, no, B, no, ni, b,
                         static_cast<float*>(nullptr), static_cast<const float*>(nullptr));
    hipcheck(hipGetLastError(), "mlp forward");
    prev = net->A[l].p;
  }
}

// Row splits of a batch of B for the weight and bias gradients.
int mlp_splits(int B) { return std::max(1, std::min(kMlpSplit, B / 64)); }

// Backward from dY (already (Y - T)/B): gradients into net->grads, or with
// to_adam their split partials into net->part at the parameters' offsets,
// n_params apart, for adam_step_splits to sum.
void backward(mh_mlp* net, const float* in, int B, bool to_adam = false) {
  const int L = net->L();
  const float* delta = net->dY.p;  // dLoss/dZ_l for the current layer
  for (int l = L - 1; l >= 0; l--) {
    const int ni = net->sizes[l], no = net->sizes[l + 1];
    const float* a_prev = l == 0 ? in : net->A[l - 1].p;
    float* gW = net->grads.p + net->w_off[l];
    float* gb = net->grads.p + net->b_off(l);
    // gW[no x ni] = delta^T[no x B] a_prev[B x ni] and gb = column sums of
    // delta, each as S row-split partials summed in order
    const int S = mlp_splits(B);
    const int rows = (B + S - 1) / S;
    float* pW = to_adam ? net->part.p + net->w_off[l] : net->part.p;
    float* pb = to_adam ? net->part.p + net->b_off(l) : net->part.p;
    const size_t sW = to_adam ? net->n_params : static_cast<size_t>(no) * ni;
    const size_t sb = to_adam ? net->n_params : static_cast<size_t>(no);
    hipLaunchKernelGGL((wgrad_l<StridedB<float, false>>), dim3((ni + kTile - 1) / kTile, (no + kTile - 1) / kTile, S),
                       dim3(256), 0, net->stream, delta, no, StridedB<float, false>{a_prev, ni}, ni, B, rows, pW,
                       sW);
    if (!to_adam)
      hipLaunchKernelGGL(sum_splits, dim3((no * ni + 255) / 256), dim3(256), 0, net->stream, net->part.p, no * ni,
                         S, gW);
    hipLaunchKernelGGL(col_partial, dim3((no + kColTile - 1) / kColTile, S), dim3(kColThreads), 0, net->stream,
                       delta, B, no, rows, pb, sb);
    if (!to_adam)
      hipLaunchKernelGGL(sum_splits, dim3((no + 255) / 256), dim3(256), 0, net->stream, net->part.p, no, S, gb);
    if (l > 0) {
      // delta_prev[B x ni] = (delta[B x no] W[no x ni]) * gelu'(Z_{l-1})
      const __hip_bfloat16* W = net->w16.p + net->w_off[l];
      hipLaunchKernelGGL((gemm<float, __hip_bfloat16, false, false, 3>), grid_for(B, ni), dim3(256), 0,
                         net->stream, delta, no, W, ni, net->dA[l].p, ni, B, ni, no,
                         static_cast<const float*>(nullptr), static_cast<float*>(nullptr),
                         static_cast<const float*>(net->Z[l - 1].p));
      delta = net->dA[l].p;
    }
    hipcheck(hipGetLastError(), "mlp backward");
  }
}

float loss_and_grad(mh_mlp* net, const float* in, const float* target, int B) {
  const int L = net->L();
  forward(net, in, B);
  const int n = B * net->sizes[L];
  hipLaunchKernelGGL(loss_grad, dim3((n + 255) / 256), dim3(256), 0, net->stream, net->A[L - 1].p, target, n,
                     1.0f / B, net->dY.p, net->terms.p);
  backward(net, in, B);
  std::vector<float> terms(n);
  hipcheck(hipMemcpyAsync(terms.data(), net->terms.p, n * sizeof(float), hipMemcpyDeviceToHost, net->stream),
           "D2H");
  hipcheck(hipStreamSynchronize(net->stream), "sync");
  double s = 0.0;
  for (float v : terms) s += v;
  return static_cast<float>(s / B);
}

int fail(const mh::Error& e) {
  mh::set_error(e.what());
  return e.code;
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch[kMaxDevices] = {};

// Device address of `scratch` on `device` (current), resolved once per device.
float* scratch_address(int device) {
  if (device < 0 || device >= kMaxDevices) throw mh::Error(MH_ERR_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    hipcheck(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratch[device] = static_cast<float*>(p);
  }
  return g_scratch[device];
}

}  // namespace

extern "C" {

int mh_mlp_create(int device, int n_layers, const int* sizes, uint64_t seed, mh_mlp** out) {
  try {
    if (!out || !sizes || n_layers < 1) throw mh::Error(MH_ERR_ARG, "invalid layer list");
    auto net = std::make_unique<mh_mlp>();
    net->device = device;
    hipcheck(hipSetDevice(device), "hipSetDevice");
    net->scratch = scratch_address(device);
    hipcheck(hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking), "stream");
    net->sizes.assign(sizes, sizes + n_layers + 1);
    for (int l = 0; l <= n_layers; l++)
      if (sizes[l] < 1) throw mh::Error(MH_ERR_ARG, "layer sizes must be >= 1");
    size_t off = 0;
    for (int l = 0; l < n_layers; l++) {
      net->w_off.push_back(off);
      off += static_cast<size_t>(sizes[l + 1]) * sizes[l] + sizes[l + 1];
    }
    net->n_params = off;
    size_t widest = 0;
    for (int l = 0; l < n_layers; l++) widest = std::max<size_t>(widest, static_cast<size_t>(sizes[l]) * sizes[l + 1]);
    // split partials of one tensor, or of every parameter for adam_step_splits
    net->part.ensure(static_cast<size_t>(kMlpSplit) * std::max(widest, net->n_params));
    // He-uniform weights (bound sqrt(6 / fan_in)), zero biases
    std::mt19937_64 rng(seed);
    std::vector<float> p(off, 0.f);
    for (int l = 0; l < n_layers; l++) {
      const float bound = std::sqrt(6.0f / sizes[l]);
      std::uniform_real_distribution<float> u(-bound, bound);
      const size_t nw = static_cast<size_t>(sizes[l + 1]) * sizes[l];
      for (size_t i = 0; i < nw; i++) p[net->w_off[l] + i] = u(rng);
    }
    net->params.ensure(off);
    net->grads.ensure(off);
    net->m.ensure(off);
    net->v.ensure(off);
    net->w16.ensure(off);
    std::vector<float> zeros(off, 0.f);
    hipcheck(hipMemcpy(net->params.p, p.data(), off * sizeof(float), hipMemcpyHostToDevice), "H2D");
    hipcheck(hipMemcpy(net->m.p, zeros.data(), off * sizeof(float), hipMemcpyHostToDevice), "H2D");
    hipcheck(hipMemcpy(net->v.p, zeros.data(), off * sizeof(float), hipMemcpyHostToDevice), "H2D");
    hipLaunchKernelGGL(to_bf16, dim3((off + 255) / 256), dim3(256), 0, net->stream, net->params.p, net->w16.p,
                       static_cast<int>(off));
    hipcheck(hipStreamSynchronize(net->stream), "init");
    *out = net.release();
    return MH_OK;
  } catch (const mh::Error& e) {
    return fail(e);
  }
}

int mh_mlp_destroy(mh_mlp* net) {
  delete net;
  return MH_OK;
}

int mh_mlp_param_count(const mh_mlp* net) { return net ? static_cast<int>(net->n_params) : MH_ERR_ARG; }

int mh_mlp_get_params(const mh_mlp* net, float* out) {
  try {
    if (!net || !out) throw mh::Error(MH_ERR_ARG, "null argument");
    hipcheck(hipMemcpy(out, net->params.p, net->n_params * sizeof(float), hipMemcpyDeviceToHost), "D2H");
    return MH_OK;
  } catch (const mh::Error& e) {
    return fail(e);
  }
}

int mh_mlp_set_params(mh_mlp* net, const float* in) {
  try {
    if (!net || !in) throw mh::Error(MH_ERR_ARG, "null argument");
    hipcheck(hipSetDevice(net->device), "hipSetDevice");
    hipcheck(hipMemcpy(net->params.p, in, net->n_params * sizeof(float), hipMemcpyHostToDevice), "H2D");
    hipLaunchKernelGGL(to_bf16, dim3((net->n_params + 255) / 256), dim3(256), 0, net->stream, net->params.p,
                       net->w16.p, static_cast<int>(net->n_params));
    hipcheck(hipStreamSynchronize(net->stream), "set params");
    net->adam_t = 0;
    std::vector<float> zeros(net->n_params, 0.f);
    hipcheck(hipMemcpy(net->m.p, zeros.data(), zeros.size() * sizeof(float), hipMemcpyHostToDevice), "H2D");
    hipcheck(hipMemcpy(net->v.p, zeros.data(), zeros.size() * sizeof(float), hipMemcpyHostToDevice), "H2D");
    return MH_OK;
  } catch (const mh::Error& e) {
    return fail(e);
  }
}

int mh_mlp_forward(mh_mlp* net, const float* x, int batch, float* y) {
  try {
    if (!net || (batch > 0 && (!x || !y)) || batch < 0) throw mh::Error(MH_ERR_ARG, "null argument");
    if (batch == 0) return MH_OK;
    hipcheck(hipSetDevice(net->device), "hipSetDevice");
    const int L = net->L();
    const int chunk = 8192;
    ensure_batch(net, std::min(batch, chunk));
    for (int b0 = 0; b0 < batch; b0 += chunk) {
      const int B = std::min(chunk, batch - b0);
      hipcheck(hipMemcpyAsync(net->x.p, x + static_cast<size_t>(b0) * net->sizes[0],
                              static_cast<size_t>(B) * net->sizes[0] * sizeof(float), hipMemcpyHostToDevice,
                              net->stream), "H2D");
      forward(net, net->x.p, B);
      hipcheck(hipMemcpyAsync(y + static_cast<size_t>(b0) * net->sizes[L], net->A[L - 1].p,
                              static_cast<size_t>(B) * net->sizes[L] * sizeof(float), hipMemcpyDeviceToHost,
                              net->stream), "D2H");
    }
    hipcheck(hipStreamSynchronize(net->stream), "sync");
    return MH_OK;
  } catch (const mh::Error& e) {
    return fail(e);
  }
}

int mh_mlp_forward_noalloc(mh_mlp* net, const float* x, int batch, float* y, void* stream) {
  try {
    if (!net || batch < 0 || (batch > 0 && (!x || !y))) throw mh::Error(MH_ERR_ARG, "null argument");
    if (batch == 0) return MH_OK;
    const int L = net->L();
    size_t widest = 0;
    for (int l = 1; l < L; l++) widest = std::max<size_t>(widest, net->sizes[l]);
    const size_t half = MH_MLP_SCRATCH_FLOATS / 2;
    if (static_cast<size_t>(batch) * widest > half)
      throw mh::Error(MH_ERR_RANGE, "batch x widest hidden layer (" + std::to_string(batch) + " x " +
                                        std::to_string(widest) + ") over the scratch array's half (" +
                                        std::to_string(half) + " floats)");
    hipcheck(hipSetDevice(net->device), "hipSetDevice");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : net->stream;
    const float* prev = x;
    for (int l = 0; l < L; l++) {
      const int ni = net->sizes[l], no = net->sizes[l + 1];
      const __hip_bfloat16* W = net->w16.p + net->w_off[l];
      const float* b = net->params.p + net->b_off(l);
      float* out = l == L - 1 ? y : net->scratch + (l % 2) * half;
      if (l < L - 1)
        hipLaunchKernelGGL((gemm<float, __hip_bfloat16, false, true, 4>), grid_for(batch, no), dim3(256), 0, s,
                           prev, ni, W, ni, out, no, batch, no, ni, b, static_cast<float*>(nullptr),
                           static_cast<const float*>(nullptr));
      else
        hipLaunchKernelGGL((gemm<float, __hip_bfloat16, false, true, 2>), grid_for(batch, no), dim3(256), 0, s,
                           prev, ni, W, ni, out, no, batch, no, ni, b, static_cast<float*>(nullptr),
                           static_cast<const float*>(nullptr));
      hipcheck(hipGetLastError(), "mlp forward (no-alloc)");
      prev = out;
    }
    return MH_OK;
  } catch (const mh::Error& e) {
    return fail(e);
  }
}

int mh_mlp_loss_grad(mh_mlp* net, const float* x, const float* target, int batch, float* loss, float* grad) {
  try {
    if (!net || !x || !target || batch < 1) throw mh::Error(MH_ERR_ARG, "null argument");
    hipcheck(hipSetDevice(net->device), "hipSetDevice");
    ensure_batch(net, batch);
    const int L = net->L();
    hipcheck(hipMemcpyAsync(net->x.p, x, static_cast<size_t>(batch) * net->sizes[0] * sizeof(float),
                            hipMemcpyHostToDevice, net->stream), "H2D");
    hipcheck(hipMemcpyAsync(net->t.p, target, static_cast<size_t>(batch) * net->sizes[L] * sizeof(float),
                            hipMemcpyHostToDevice, net->stream), "H2D");
    const float l = loss_and_grad(net, net->x.p, net->t.p, batch);
    if (loss) *loss = l;
    if (grad)
      hipcheck(hipMemcpy(grad, net->grads.p, net->n_params * sizeof(float), hipMemcpyDeviceToHost), "D2H");
    return MH_OK;
  } catch (const mh::Error& e) {
    return fail(e);
  }
}

int mh_mlp_train(mh_mlp* net, const float* x, const float* target, int n, int batch, int n_epochs, float lr,
                 uint64_t seed, float* final_loss) {
  try {
    if (!net || !x || !target || n < 1 || batch < 1 || n_epochs < 0) throw mh::Error(MH_ERR_ARG, "bad arguments");
    hipcheck(hipSetDevice(net->device), "hipSetDevice");
    const int L = net->L();
    const int ni = net->sizes[0], no = net->sizes[L];
    batch = std::min(batch, n);
    ensure_batch(net, batch);
    net->data_x.ensure(static_cast<size_t>(n) * ni);
    net->data_t.ensure(static_cast<size_t>(n) * no);
    net->idx.ensure(n);
    hipcheck(hipMemcpyAsync(net->data_x.p, x, static_cast<size_t>(n) * ni * sizeof(float), hipMemcpyHostToDevice,
                            net->stream), "H2D");
    hipcheck(hipMemcpyAsync(net->data_t.p, target, static_cast<size_t>(n) * no * sizeof(float),
                            hipMemcpyHostToDevice, net->stream), "H2D");
    std::mt19937_64 rng(seed);
    std::vector<int> perm(n);
    // the last epoch keeps every batch's loss terms on the device, so no
    // batch waits for a copy to the host; they are summed after the loop
    const int n_batches = n / batch;
    const size_t nt = static_cast<size_t>(batch) * no;
    net->epoch_terms.ensure(static_cast<size_t>(n_batches) * nt);
    int batches = 0;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < n_epochs; ep++) {
      for (int i = 0; i < n; i++) perm[i] = i;
      std::shuffle(perm.begin(), perm.end(), rng);
      hipcheck(hipMemcpyAsync(net->idx.p, perm.data(), n * sizeof(int), hipMemcpyHostToDevice, net->stream), "H2D");
      const bool last = ep == n_epochs - 1;
      for (int b0 = 0; b0 + batch <= n; b0 += batch) {
        hipLaunchKernelGGL(gather_rows, dim3(batch), dim3(64), 0, net->stream, net->data_x.p, net->idx.p + b0,
                           batch, ni, net->x.p);
        hipLaunchKernelGGL(gather_rows, dim3(batch), dim3(64), 0, net->stream, net->data_t.p, net->idx.p + b0,
                           batch, no, net->t.p);
        forward(net, net->x.p, batch);
        float* terms = last ? net->epoch_terms.p + batches * nt : net->terms.p;
        hipLaunchKernelGGL(loss_grad, dim3((static_cast<int>(nt) + 255) / 256), dim3(256), 0, net->stream,
                           net->A[L - 1].p, net->t.p, static_cast<int>(nt), 1.0f / batch, net->dY.p, terms);
        backward(net, net->x.p, batch, true);
        net->adam_t++;
        const float bc1 = 1.f - std::pow(b1, static_cast<float>(net->adam_t));
        const float bc2 = 1.f - std::pow(b2, static_cast<float>(net->adam_t));
        hipLaunchKernelGGL(adam_step_splits, dim3((net->n_params + 255) / 256), dim3(256), 0, net->stream,
                           net->params.p, net->part.p, mlp_splits(batch), net->n_params, net->m.p, net->v.p,
                           net->w16.p, static_cast<int>(net->n_params), lr, b1, b2, eps, bc1, bc2);
        hipcheck(hipGetLastError(), "train step");
        if (last) batches++;
      }
    }
    // the last epoch's mean over batches of each batch's summed terms / batch
    std::vector<float> terms(static_cast<size_t>(batches) * nt);
    if (batches)
      hipcheck(hipMemcpyAsync(terms.data(), net->epoch_terms.p, terms.size() * sizeof(float),
                              hipMemcpyDeviceToHost, net->stream), "D2H");
    hipcheck(hipStreamSynchronize(net->stream), "sync");
    double epoch_loss = 0.0;
    for (int bi = 0; bi < batches; bi++) {
      double s = 0.0;
      for (size_t j = 0; j < nt; j++) s += terms[bi * nt + j];
      epoch_loss += s / batch;
    }
    if (batches) epoch_loss /= batches;
    if (final_loss) *final_loss = static_cast<float>(epoch_loss);
    return MH_OK;
  } catch (const mh::Error& e) {
    return fail(e);
  }
}

}  // extern "C"
