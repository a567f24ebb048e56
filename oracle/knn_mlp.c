/*
 * knn_mlp.c -- oracle KNN neighbour lists and FP64 surrogate MLP.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * orc_knn follows common/evolvables/knn.cu:147-190 (KNN_CPU): distances are
 * Eigen Vector3f norms (sqrtf(dx*dx + (dy*dy + dz*dz))), a zero distance
 * (the mass itself, or a coincident mass) is +inf, rows are sorted
 * ascending and cut to K.  KNN_CPU uses std::sort, whose order of equal
 * distances is unspecified; the GPU path of the reference (cub radix sort
 * of rows whose ids start ascending, knn.cu:26-60) keeps equal distances in
 * index order, which is the order used here.
 * Pinned by: brute-force recomputation in numpy with (distance, index) order
 * (tests/test_oracle_knn_mlp.py).
 *
 * MLP: y = W_L g(... g(W_1 x + b_1) ...) + b_L with g = GELU, tanh form
 * g(x) = 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))).  Loss is
 * mean_b 0.5 (y_b - target_b)^2.  Adam is Kingma & Ba (2015) Algorithm 1
 * with bias correction.  With bf16_operands set, every product operand
 * that is an activation or a back-propagated gradient is rounded as the
 * device's matrix-core GEMMs round it (to float, then to bf16, nearest
 * even): the layer inputs in W a, both factors of the weight gradient
 * delta a^T, and delta in W^T delta; bias gradients sum the unrounded
 * deltas.  The weights are taken as given (the caller passes bf16 values).
 * Pinned by: central finite differences of the loss, a one-layer closed
 * form (linear regression gradient), and Adam's first step equal to
 * -lr * sign(g) (tests/test_oracle_knn_mlp.py); the bf16_operands mode by
 * a numpy restatement with its own rounding, which differs from the
 * unrounded network by far more than the two disagree.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

typedef struct { float d; int j; } nd_t;

static int cmp_nd(const void* a, const void* b) {
  const nd_t* x = (const nd_t*)a;
  const nd_t* y = (const nd_t*)b;
  if (x->d < y->d) return -1;
  if (x->d > y->d) return 1;
  return x->j - y->j;
}

int orc_knn(const float* pos, int n, int K, uint16_t* nbr, float* dist) {
  if (K > n) return -1;
  nd_t* row = (nd_t*)malloc(sizeof(nd_t) * (size_t)n);
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < n; j++) {
      float dx = pos[3 * i] - pos[3 * j];
      float dy = pos[3 * i + 1] - pos[3 * j + 1];
      float dz = pos[3 * i + 2] - pos[3 * j + 2];
      float yz = dy * dy + dz * dz;
      float d = sqrtf(dx * dx + yz);
      row[j].d = (d == 0.0f) ? INFINITY : d;
      row[j].j = j;
    }
    qsort(row, (size_t)n, sizeof(nd_t), cmp_nd);
    for (int k = 0; k < K; k++) {
      nbr[(size_t)i * K + k] = (uint16_t)row[k].j;
      dist[(size_t)i * K + k] = row[k].d;
    }
  }
  free(row);
  return 0;
}

static const double kGeluC = 0.7978845608028654; /* sqrt(2/pi) */

static double gelu(double x) {
  double u = kGeluC * (x + 0.044715 * x * x * x);
  return 0.5 * x * (1.0 + tanh(u));
}
static double gelu_d(double x) {
  double u = kGeluC * (x + 0.044715 * x * x * x);
  double th = tanh(u);
  return 0.5 * (1.0 + th) +
         0.5 * x * (1.0 - th * th) * kGeluC * (1.0 + 3.0 * 0.044715 * x * x);
}

/* v rounded to float, then to bf16 (round to nearest even), as a double */
static double bf16r(double v) {
  float f = (float)v;
  uint32_t u;
  memcpy(&u, &f, 4);
  u = ((u + 0x7fffu + ((u >> 16) & 1u)) >> 16) << 16;
  memcpy(&f, &u, 4);
  return (double)f;
}

/* a copy of n values, rounded by bf16r when rb is set */
static double* operand(const double* v, size_t n, int rb) {
  double* o = (double*)malloc(sizeof(double) * (n ? n : 1));
  for (size_t i = 0; i < n; i++) o[i] = rb ? bf16r(v[i]) : v[i];
  return o;
}

/* activations buffer: pre-activations z_l and outputs a_l for every layer */
static void forward_all(int nl, const int* sz, const double* const* W,
                        const double* const* b, int B, const double* x,
                        double** z, double** a, int rb) {
  const double* in = x;
  for (int l = 0; l < nl; l++) {
    int ni = sz[l], no = sz[l + 1];
    double* op = operand(in, (size_t)B * ni, rb);
    for (int r = 0; r < B; r++)
      for (int o = 0; o < no; o++) {
        double s = b[l][o];
        const double* w = W[l] + (size_t)o * ni;
        const double* xi = op + (size_t)r * ni;
        for (int i = 0; i < ni; i++) s += w[i] * xi[i];
        z[l][(size_t)r * no + o] = s;
        a[l][(size_t)r * no + o] = (l < nl - 1) ? gelu(s) : s;
      }
    free(op);
    in = a[l];
  }
}

int orc_mlp_forward(int n_layers, const int* sizes, const double* const* W,
                    const double* const* b, int batch, const double* x,
                    double* y, int bf16_operands) {
  if (n_layers < 1 || batch < 0) return -1;
  double** z = (double**)malloc(sizeof(double*) * (size_t)n_layers);
  double** a = (double**)malloc(sizeof(double*) * (size_t)n_layers);
  for (int l = 0; l < n_layers; l++) {
    z[l] = (double*)malloc(sizeof(double) * (size_t)batch * sizes[l + 1]);
    a[l] = (double*)malloc(sizeof(double) * (size_t)batch * sizes[l + 1]);
  }
  forward_all(n_layers, sizes, W, b, batch, x, z, a, bf16_operands);
  memcpy(y, a[n_layers - 1], sizeof(double) * (size_t)batch * sizes[n_layers]);
  for (int l = 0; l < n_layers; l++) { free(z[l]); free(a[l]); }
  free(z); free(a);
  return 0;
}

double orc_mlp_loss_grad(int n_layers, const int* sizes,
                         const double* const* W, const double* const* b,
                         int batch, const double* x, const double* target,
                         double* const* gW, double* const* gb, int bf16_operands) {
  int nl = n_layers;
  double** z = (double**)malloc(sizeof(double*) * (size_t)nl);
  double** a = (double**)malloc(sizeof(double*) * (size_t)nl);
  for (int l = 0; l < nl; l++) {
    z[l] = (double*)malloc(sizeof(double) * (size_t)batch * sizes[l + 1]);
    a[l] = (double*)malloc(sizeof(double) * (size_t)batch * sizes[l + 1]);
  }
  forward_all(nl, sizes, W, b, batch, x, z, a, bf16_operands);
  int nout = sizes[nl];
  double loss = 0.0;
  /* delta of the last layer: dL/dz = (y - t) / B */
  double* delta = (double*)malloc(sizeof(double) * (size_t)batch * nout);
  for (int r = 0; r < batch; r++)
    for (int o = 0; o < nout; o++) {
      double e = a[nl - 1][(size_t)r * nout + o] - target[(size_t)r * nout + o];
      loss += 0.5 * e * e;
      delta[(size_t)r * nout + o] = e / batch;
    }
  loss /= batch;
  for (int l = nl - 1; l >= 0; l--) {
    int ni = sizes[l], no = sizes[l + 1];
    double* in = operand((l == 0) ? x : a[l - 1], (size_t)batch * ni, bf16_operands);
    double* dr = operand(delta, (size_t)batch * no, bf16_operands);
    memset(gW[l], 0, sizeof(double) * (size_t)ni * no);
    memset(gb[l], 0, sizeof(double) * (size_t)no);
    for (int r = 0; r < batch; r++)
      for (int o = 0; o < no; o++) {
        gb[l][o] += delta[(size_t)r * no + o];
        double d = dr[(size_t)r * no + o];
        for (int i = 0; i < ni; i++) gW[l][(size_t)o * ni + i] += d * in[(size_t)r * ni + i];
      }
    free(in);
    if (l > 0) {
      double* nd = (double*)malloc(sizeof(double) * (size_t)batch * ni);
      for (int r = 0; r < batch; r++)
        for (int i = 0; i < ni; i++) {
          double s = 0.0;
          for (int o = 0; o < no; o++) s += W[l][(size_t)o * ni + i] * dr[(size_t)r * no + o];
          nd[(size_t)r * ni + i] = s * gelu_d(z[l - 1][(size_t)r * ni + i]);
        }
      free(delta);
      delta = nd;
    }
    free(dr);
  }
  free(delta);
  for (int l = 0; l < nl; l++) { free(z[l]); free(a[l]); }
  free(z); free(a);
  return loss;
}

void orc_adam(int n, double* p, const double* g, double* m, double* v,
              double lr, double beta1, double beta2, double eps, int t) {
  double bc1 = 1.0 - pow(beta1, t), bc2 = 1.0 - pow(beta2, t);
  for (int i = 0; i < n; i++) {
    m[i] = beta1 * m[i] + (1.0 - beta1) * g[i];
    v[i] = beta2 * v[i] + (1.0 - beta2) * g[i] * g[i];
    double mh = m[i] / bc1, vh = v[i] / bc2;
    p[i] -= lr * mh / (sqrt(vh) + eps);
  }
}
