// mlp.hip -- the surrogate network of the warm start: a small MLP trained
// and evaluated on the GPU with hand-written MFMA kernels.
//
// Layers: n_in inputs (zero-padded to 32) -> n_hidden layers of `width` GELU
// units -> n_out linear outputs.  Weights live as FP32 masters (Adam
// updates them) with BF16 copies the matrix kernels read; products
// accumulate in FP32 (v_mfma_f32_32x32x8_bf16: one 32x32 output tile per
// wavefront, K in steps of 8).  Activations feeding a product are stored
// in BF16, pre-activations and gradients in FP32.  Batches are padded to a
// multiple of 32 rows; padded rows carry zero loss and zero gradient.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "dfargo_internal.h"

#define IN_PAD 32 // the input layer's K, padded to a whole output tile of dW

#define MCHK(call)                                                                                       \
  do {                                                                                                   \
    hipError_t e_ = (call);                                                                              \
    if (e_ != hipSuccess)                                                                                \
      return dfargo_set_error(DFARGO_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),     \
                              __FILE__, __LINE__);                                                       \
  } while (0)

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef uint16_t bf16;

__device__ __forceinline__ bf16 to_bf16(float f) { // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}

// C[M x N] = op(A)[M x K] op(B)[K x N] (+ bias[N]); op(A)(m, k) is A[m*lda + k]
// or, transposed, A[k*lda + m]; op(B)(k, n) is B[k*ldb + n] or B[n*ldb + k].
// M, N multiples of 32 and K of 8 (checked by the host).  One wavefront per
// 32x32 tile of C.  Operand layout of the 32x32x8 instruction: lane l holds
// op(A)(l % 32, 4 (l / 32) + v) and op(B)(4 (l / 32) + v, l % 32), v < 4; its
// 16 results are C(8 (v / 4) + 4 (l / 32) + v % 4, l % 32).
template <bool TA, bool TB>
__global__ void __launch_bounds__(64) k_gemm(const bf16 *__restrict__ A, const bf16 *__restrict__ B,
                                             float *__restrict__ C, const float *__restrict__ bias, int K, int lda,
                                             int ldb, int ldc) {
  const int l = threadIdx.x, m0 = blockIdx.y * 32, n0 = blockIdx.x * 32;
  const int r = l & 31, kq = (l >> 5) * 4;
  floatx16 acc;
  for (int v = 0; v < 16; v++) acc[v] = 0.0f;
  for (int k0 = 0; k0 < K; k0 += 8) {
    bf16x4 a, b;
    for (int v = 0; v < 4; v++) {
      const int k = k0 + kq + v;
      a[v] = (short)(TA ? A[(long)k * lda + m0 + r] : A[(long)(m0 + r) * lda + k]);
      b[v] = (short)(TB ? B[(long)(n0 + r) * ldb + k] : B[(long)k * ldb + n0 + r]);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
  }
  const float bb = bias ? bias[n0 + r] : 0.0f;
  for (int v = 0; v < 16; v++) {
    const int row = m0 + 8 * (v >> 2) + (l >> 5) * 4 + (v & 3);
    C[(long)row * ldc + n0 + r] = acc[v] + bb;
  }
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
  return cdf + z * pdf;
}

__global__ void k_gelu_fwd(const float *z, bf16 *h, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) h[i] = to_bf16(gelu(z[i]));
}
// dZ = dH GELU'(Z), FP32 for the bias sums and BF16 for the products
__global__ void k_gelu_bwd(const float *dh, const float *z, float *dz, bf16 *dzb, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const float g = dh[i] * gelu_grad(z[i]);
    dz[i] = g;
    dzb[i] = to_bf16(g);
  }
}
// the padded input rows: x (batch x n_in, FP32) into X (rows x IN_PAD, BF16)
__global__ void k_input(const float *x, bf16 *X, int batch, int n_in, int rows) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * IN_PAD) return;
  const int r = i / IN_PAD, c = i % IN_PAD;
  X[i] = (r < batch && c < n_in) ? to_bf16(x[r * n_in + c]) : (bf16)0;
}
// mean squared error over the first `batch` rows and its gradient
// dY = 2 (Y - T) / (batch n_out); padded rows get 0.  One workgroup; the
// loss is a fixed-order tree sum.
__global__ void __launch_bounds__(256) k_mse(const float *y, const float *t, float *dy, bf16 *dyb, int batch,
                                             int n_out, int rows, float *loss) {
  __shared__ float sh[256];
  const int tid = threadIdx.x;
  const float scale = 2.0f / (float)(batch * n_out);
  float s = 0.0f;
  for (long i = tid; i < (long)rows * n_out; i += 256) {
    const bool in = i / n_out < batch;
    const float d = in ? y[i] - t[i] : 0.0f;
    s += d * d;
    dy[i] = d * scale;
    dyb[i] = to_bf16(d * scale);
  }
  sh[tid] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) sh[tid] += sh[tid + w];
    __syncthreads();
  }
  if (tid == 0) *loss = sh[0] / (float)(batch * n_out);
}
// db[j] = sum over the rows of dZ[., j]
__global__ void k_colsum(const float *dz, float *db, int rows, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  float s = 0.0f;
  for (int r = 0; r < rows; r++) s += dz[(long)r * n + j];
  db[j] = s;
}
// Adam on the FP32 masters, refreshing the BF16 copy (if any)
__global__ void k_adam(float *w, const float *g, float *m, float *v, bf16 *wb, long n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi, vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = to_bf16(wi);
}
__global__ void k_weights_bf16(const float *w, bf16 *wb, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) wb[i] = to_bf16(w[i]);
}

// ------------------------------------------------------------------------

struct Layer {
  int k, n;          // inputs (padded) x outputs
  float *w, *b;      // FP32 masters: w[k][n] (padded rows zero), b[n]
  bf16 *wb;          // BF16 copy of w
  float *gw, *gb;    // gradients
  float *mw, *vw, *mb, *vb;
  float *z, *dh, *dz; // pre-activation, gradients (rows x n)
  bf16 *h, *dzb;     // activation (BF16), gradient (BF16)
};

struct dfargo_mlp {
  int n_in, width, n_hidden, n_out, max_rows, dev;
  long step = 0;
  hipStream_t stream = nullptr;
  std::vector<Layer> L;
  bf16 *X = nullptr;
  float *x_in = nullptr, *t_in = nullptr, *loss = nullptr;
  std::vector<void *> allocs;
};

static int malloc_zero(dfargo_mlp *m, void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return dfargo_set_error(DFARGO_ERR_ALLOC, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  m->allocs.push_back(*p);
  MCHK(hipMemset(*p, 0, bytes));
  return DFARGO_OK;
}

static unsigned long long splitmix(unsigned long long *s) {
  unsigned long long z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

extern "C" int dfargo_mlp_destroy(dfargo_mlp *m) {
  if (!m) return DFARGO_OK;
  hipSetDevice(m->dev);
  if (m->stream) hipStreamSynchronize(m->stream);
  for (void *p : m->allocs) hipFree(p);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
  return DFARGO_OK;
}

extern "C" size_t dfargo_mlp_param_count(const dfargo_mlp *m) {
  if (!m) return 0;
  size_t n = 0;
  for (size_t l = 0; l < m->L.size(); l++) {
    const int kin = l == 0 ? m->n_in : m->L[l].k;
    n += (size_t)kin * m->L[l].n + m->L[l].n;
  }
  return n;
}

extern "C" int dfargo_mlp_create(int n_in, int width, int n_hidden, int n_out, int max_batch, unsigned long seed,
                                 int device, dfargo_mlp **out) {
  if (!out) return dfargo_set_error(DFARGO_ERR_ARG, "null output");
  *out = nullptr;
  if (n_in < 1 || n_in > IN_PAD || width < 32 || width % 32 || n_out < 32 || n_out % 32 || n_hidden < 1 ||
      n_hidden > 16 || max_batch < 1 || max_batch > 65536)
    return dfargo_set_error(DFARGO_ERR_ARG, "need 1 <= n_in <= 32, width and n_out multiples of 32, "
                                            "1 <= n_hidden <= 16, 1 <= max_batch <= 65536");
  MCHK(hipSetDevice(device));
  dfargo_mlp *m = new dfargo_mlp();
  m->n_in = n_in;
  m->width = width;
  m->n_hidden = n_hidden;
  m->n_out = n_out;
  m->max_rows = (max_batch + 31) / 32 * 32;
  m->dev = device;
  int rc = DFARGO_OK;
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
    delete m;
    return dfargo_set_error(DFARGO_ERR_HIP, "hipStreamCreate failed");
  }
  const size_t R = m->max_rows;
  unsigned long long st = seed;
  for (int l = 0; l <= n_hidden && !rc; l++) {
    Layer y;
    memset(&y, 0, sizeof y);
    y.k = l == 0 ? IN_PAD : width;
    y.n = l == n_hidden ? n_out : width;
    const size_t nw = (size_t)y.k * y.n;
    if ((rc = malloc_zero(m, (void **)&y.w, 4 * nw)) || (rc = malloc_zero(m, (void **)&y.b, 4 * y.n)) ||
        (rc = malloc_zero(m, (void **)&y.wb, 2 * nw)) || (rc = malloc_zero(m, (void **)&y.gw, 4 * nw)) ||
        (rc = malloc_zero(m, (void **)&y.gb, 4 * y.n)) || (rc = malloc_zero(m, (void **)&y.mw, 4 * nw)) ||
        (rc = malloc_zero(m, (void **)&y.vw, 4 * nw)) || (rc = malloc_zero(m, (void **)&y.mb, 4 * y.n)) ||
        (rc = malloc_zero(m, (void **)&y.vb, 4 * y.n)) || (rc = malloc_zero(m, (void **)&y.z, 4 * R * y.n)) ||
        (rc = malloc_zero(m, (void **)&y.dh, 4 * R * y.n)) || (rc = malloc_zero(m, (void **)&y.dz, 4 * R * y.n)) ||
        (rc = malloc_zero(m, (void **)&y.h, 2 * R * y.n)) || (rc = malloc_zero(m, (void **)&y.dzb, 2 * R * y.n)))
      break;
    // Glorot-uniform weights over the real inputs, zero biases
    const int kin = l == 0 ? n_in : y.k;
    const double lim = sqrt(6.0 / (double)(kin + y.n));
    std::vector<float> w(nw, 0.0f);
    for (int i = 0; i < kin; i++)
      for (int j = 0; j < y.n; j++)
        w[(size_t)i * y.n + j] = (float)(lim * (2.0 * (double)(splitmix(&st) >> 11) / 9007199254740992.0 - 1.0));
    if (hipMemcpy(y.w, w.data(), 4 * nw, hipMemcpyHostToDevice) != hipSuccess) {
      rc = dfargo_set_error(DFARGO_ERR_HIP, "weight upload failed");
      break;
    }
    hipLaunchKernelGGL(k_weights_bf16, dim3((nw + 255) / 256), dim3(256), 0, m->stream, y.w, y.wb, (long)nw);
    m->L.push_back(y);
  }
  if (!rc && !(rc = malloc_zero(m, (void **)&m->X, 2 * R * IN_PAD)) &&
      !(rc = malloc_zero(m, (void **)&m->x_in, 4 * R * n_in)) && !(rc = malloc_zero(m, (void **)&m->t_in, 4 * R * n_out)))
    rc = malloc_zero(m, (void **)&m->loss, 4);
  if (!rc && hipStreamSynchronize(m->stream) != hipSuccess)
    rc = dfargo_set_error(DFARGO_ERR_HIP, "initialisation failed");
  if (rc) {
    dfargo_mlp_destroy(m);
    return rc;
  }
  *out = m;
  return DFARGO_OK;
}

// forward pass over `rows` padded rows (m->X filled)
static int forward(dfargo_mlp *m, int rows) {
  const bf16 *in = m->X;
  for (size_t l = 0; l < m->L.size(); l++) {
    Layer &y = m->L[l];
    hipLaunchKernelGGL((k_gemm<false, false>), dim3(y.n / 32, rows / 32), dim3(64), 0, m->stream, in, y.wb, y.z,
                       y.b, y.k, y.k, y.n, y.n);
    const long n = (long)rows * y.n;
    if (l + 1 < m->L.size())
      hipLaunchKernelGGL(k_gelu_fwd, dim3((n + 255) / 256), dim3(256), 0, m->stream, y.z, y.h, n);
    in = y.h;
  }
  MCHK(hipGetLastError());
  return DFARGO_OK;
}

static int load_inputs(dfargo_mlp *m, int batch, const float *x, int *rows) {
  if (batch < 1 || batch > m->max_rows) return dfargo_set_error(DFARGO_ERR_ARG, "batch %d outside [1, %d]", batch, m->max_rows);
  *rows = (batch + 31) / 32 * 32;
  MCHK(hipMemcpyAsync(m->x_in, x, 4 * (size_t)batch * m->n_in, hipMemcpyHostToDevice, m->stream));
  const int n = *rows * IN_PAD;
  hipLaunchKernelGGL(k_input, dim3((n + 255) / 256), dim3(256), 0, m->stream, m->x_in, m->X, batch, m->n_in, *rows);
  MCHK(hipGetLastError());
  return DFARGO_OK;
}

extern "C" int dfargo_mlp_forward(dfargo_mlp *m, int batch, const float *x, float *y) {
  if (!m || !x || !y) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  MCHK(hipSetDevice(m->dev));
  int rows, rc;
  if ((rc = load_inputs(m, batch, x, &rows)) || (rc = forward(m, rows))) return rc;
  MCHK(hipMemcpyAsync(y, m->L.back().z, 4 * (size_t)batch * m->n_out, hipMemcpyDeviceToHost, m->stream));
  MCHK(hipStreamSynchronize(m->stream));
  return DFARGO_OK;
}

// loss and gradients of the targets in m->t_in (rows x n_out)
static int backward(dfargo_mlp *m, int batch, int rows) {
  Layer &o = m->L.back();
  hipLaunchKernelGGL(k_mse, dim3(1), dim3(256), 0, m->stream, o.z, m->t_in, o.dz, o.dzb, batch, m->n_out, rows,
                     m->loss);
  for (int l = (int)m->L.size() - 1; l >= 0; l--) {
    Layer &y = m->L[l];
    if (l + 1 < (int)m->L.size()) {
      const long n = (long)rows * y.n;
      hipLaunchKernelGGL(k_gelu_bwd, dim3((n + 255) / 256), dim3(256), 0, m->stream, y.dh, y.z, y.dz, y.dzb, n);
    }
    const bf16 *in = l == 0 ? m->X : m->L[l - 1].h;
    // dW = in^T dZ (k x n, contraction over the rows), db = column sums
    hipLaunchKernelGGL((k_gemm<true, false>), dim3(y.n / 32, y.k / 32), dim3(64), 0, m->stream,
                       in, y.dzb, y.gw, (const float *)nullptr, rows, y.k, y.n, y.n);
    hipLaunchKernelGGL(k_colsum, dim3((y.n + 255) / 256), dim3(256), 0, m->stream, y.dz, y.gb, rows, y.n);
    // dH of the layer below = dZ W^T
    if (l > 0) {
      Layer &p = m->L[l - 1];
      hipLaunchKernelGGL((k_gemm<false, true>), dim3(p.n / 32, rows / 32), dim3(64), 0, m->stream, y.dzb, y.wb,
                         p.dh, (const float *)nullptr, y.n, y.n, y.n, p.n);
    }
  }
  MCHK(hipGetLastError());
  return DFARGO_OK;
}

static int load_targets(dfargo_mlp *m, int batch, const float *t) {
  MCHK(hipMemcpyAsync(m->t_in, t, 4 * (size_t)batch * m->n_out, hipMemcpyHostToDevice, m->stream));
  return DFARGO_OK;
}

// the parameters (or their gradients) in the flat order W0, b0, W1, b1, ...
// with W_l as n_in(l) x n(l) row-major over the real inputs
static int copy_params(dfargo_mlp *m, float *host, int grads, int to_device) {
  size_t off = 0;
  for (size_t l = 0; l < m->L.size(); l++) {
    Layer &y = m->L[l];
    const int kin = l == 0 ? m->n_in : y.k;
    float *w = grads ? y.gw : y.w, *b = grads ? y.gb : y.b;
    const size_t nw = (size_t)kin * y.n;
    if (to_device) {
      MCHK(hipMemcpyAsync(w, host + off, 4 * nw, hipMemcpyHostToDevice, m->stream));
      MCHK(hipMemcpyAsync(b, host + off + nw, 4 * (size_t)y.n, hipMemcpyHostToDevice, m->stream));
      const size_t all = (size_t)y.k * y.n;
      hipLaunchKernelGGL(k_weights_bf16, dim3((all + 255) / 256), dim3(256), 0, m->stream, y.w, y.wb, (long)all);
    } else {
      MCHK(hipMemcpyAsync(host + off, w, 4 * nw, hipMemcpyDeviceToHost, m->stream));
      MCHK(hipMemcpyAsync(host + off + nw, b, 4 * (size_t)y.n, hipMemcpyDeviceToHost, m->stream));
    }
    off += nw + y.n;
  }
  MCHK(hipStreamSynchronize(m->stream));
  return DFARGO_OK;
}

extern "C" int dfargo_mlp_get_params(dfargo_mlp *m, float *out) {
  if (!m || !out) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  MCHK(hipSetDevice(m->dev));
  return copy_params(m, out, 0, 0);
}

extern "C" int dfargo_mlp_set_params(dfargo_mlp *m, const float *in) {
  if (!m || !in) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  MCHK(hipSetDevice(m->dev));
  return copy_params(m, (float *)in, 0, 1);
}

extern "C" int dfargo_mlp_gradients(dfargo_mlp *m, int batch, const float *x, const float *t, float *loss,
                                    float *grads) {
  if (!m || !x || !t || !loss || !grads) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  MCHK(hipSetDevice(m->dev));
  int rows, rc;
  if ((rc = load_inputs(m, batch, x, &rows)) || (rc = load_targets(m, batch, t)) || (rc = forward(m, rows)) ||
      (rc = backward(m, batch, rows)))
    return rc;
  MCHK(hipMemcpyAsync(loss, m->loss, 4, hipMemcpyDeviceToHost, m->stream));
  return copy_params(m, grads, 1, 0);
}

extern "C" int dfargo_mlp_train(dfargo_mlp *m, int batch, const float *x, const float *t, int steps, float lr,
                                float *loss) {
  if (!m || !x || !t) return dfargo_set_error(DFARGO_ERR_ARG, "null argument");
  if (steps < 0 || !(lr > 0.0f)) return dfargo_set_error(DFARGO_ERR_ARG, "need steps >= 0 and lr > 0");
  MCHK(hipSetDevice(m->dev));
  int rows, rc;
  if ((rc = load_inputs(m, batch, x, &rows)) || (rc = load_targets(m, batch, t))) return rc;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int s = 0; s < steps; s++) {
    if ((rc = forward(m, rows)) || (rc = backward(m, batch, rows))) return rc;
    m->step++;
    const float c1 = 1.0f - powf(b1, (float)m->step), c2 = 1.0f - powf(b2, (float)m->step);
    for (Layer &y : m->L) {
      const long nw = (long)y.k * y.n;
      hipLaunchKernelGGL(k_adam, dim3((nw + 255) / 256), dim3(256), 0, m->stream, y.w, y.gw, y.mw, y.vw, y.wb, nw,
                         lr, b1, b2, eps, c1, c2);
      hipLaunchKernelGGL(k_adam, dim3((y.n + 255) / 256), dim3(256), 0, m->stream, y.b, y.gb, y.mb, y.vb,
                         (bf16 *)nullptr, (long)y.n, lr, b1, b2, eps, c1, c2);
    }
    MCHK(hipGetLastError());
  }
  if (loss) {
    // the loss of the parameters after the last step
    if ((rc = forward(m, rows)) || (rc = backward(m, batch, rows))) return rc;
    MCHK(hipMemcpyAsync(loss, m->loss, 4, hipMemcpyDeviceToHost, m->stream));
  }
  MCHK(hipStreamSynchronize(m->stream));
  return DFARGO_OK;
}
