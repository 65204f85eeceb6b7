// sg.hip -- self-gravity by log-polar FFT convolution on CDNA4
// (compute_selfgravity, src/sgmain.c:3-35; compute_fftdensity,
// src/sgdens.c:3-86; compute_sgacc, src/sgacc.c:3-162; update_sgvelocity,
// src/sgupdate.c:3-73).
//
// The (2 Nrad) x Nsec real transforms of the reference are done as one
// complex transform of Sr + i St (the two reduced densities share it), with
// the zero padding of rows >= Nrad applied on load.  The spectral product
// separates the two spectra from Z(k) and conj(Z(-k)), multiplies by the
// kernel spectra and recombines Ar + i At, so one complex inverse transform
// returns both accelerations.  The 1-D transforms are Stockham radix-4/3/2
// in LDS, one workgroup per row or column; twiddles come from tables the
// host fills in FP64.
#include <hip/hip_runtime.h>

#include "kernels.h"

template <typename T> struct cx {
  T x, y;
};
template <typename T> __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T> __device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) {
  return {a.x + b.x, a.y + b.y};
}
template <typename T> __device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) {
  return {a.x - b.x, a.y - b.y};
}
// (s * i) * a for real s
template <typename T> __device__ __forceinline__ cx<T> cmuli(cx<T> a, T s) {
  return {-s * a.y, s * a.x};
}

// DFT of length 4 in the direction dir (-1 forward, +1 inverse)
template <typename T> __device__ __forceinline__ void dft4(const cx<T> (&x)[4], cx<T> (&y)[4], T dir) {
  const cx<T> a0 = cadd(x[0], x[2]), a1 = csub(x[0], x[2]);
  const cx<T> a2 = cadd(x[1], x[3]), a3 = cmuli(csub(x[1], x[3]), dir);
  y[0] = cadd(a0, a2);
  y[2] = csub(a0, a2);
  y[1] = cadd(a1, a3);
  y[3] = csub(a1, a3);
}

// One radix-R pass of a Stockham transform: nb = L/R butterflies, each
// reading R inputs at stride nb and writing R outputs at stride Ns.  The
// quotient j / Ns comes from a 32-bit reciprocal (mg = 2^32/Ns rounded up,
// exact for j < 2^12 and Ns <= 2^12), not an integer division.
template <typename T, int R>
__device__ __forceinline__ void fft_pass(const cx<T> *in, cx<T> *out, int L, int Ns, unsigned mg, int span,
                                         const cx<T> *tw, int dir) {
  const int nb = L / R;
  const T s3 = (T)(dir * 0.86602540378443864676); // dir * sqrt(3)/2
  const T c8 = (T)0.70710678118654752440;         // sqrt(1/2)
  const T fd = (T)dir;
  for (int j = threadIdx.x; j < nb; j += blockDim.x) {
    const int g = (Ns == 1) ? j : (int)__umulhi((unsigned)j, mg);
    const int k = j - g * Ns;
    cx<T> v[R], y[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = in[j + r * nb];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; r++) {
        cx<T> w = tw[r * k * span];
        if (dir > 0) w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
    }
    if constexpr (R == 2) {
      y[0] = cadd(v[0], v[1]);
      y[1] = csub(v[0], v[1]);
    } else if constexpr (R == 4) {
      dft4<T>(v, y, fd);
    } else if constexpr (R == 8) {
      // y[2m] = DFT4(v[k] + v[k+4]), y[2m+1] = DFT4((v[k] - v[k+4]) W8^k)
      cx<T> e[4], o[4], E[4], O[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        e[q] = cadd(v[q], v[q + 4]);
        o[q] = csub(v[q], v[q + 4]);
      }
      // W8 = exp(dir i pi/4): (c8, dir c8); W8^2 = dir i; W8^3 = (-c8, dir c8)
      o[1] = {c8 * (o[1].x - fd * o[1].y), c8 * (o[1].y + fd * o[1].x)};
      o[2] = cmuli(o[2], fd);
      o[3] = {-c8 * (o[3].x + fd * o[3].y), c8 * (fd * o[3].x - o[3].y)};
      dft4<T>(e, E, fd);
      dft4<T>(o, O, fd);
#pragma unroll
      for (int q = 0; q < 4; q++) {
        y[2 * q] = E[q];
        y[2 * q + 1] = O[q];
      }
    } else {
      cx<T> t1 = cadd(v[1], v[2]);
      cx<T> t2 = {v[0].x - (T)0.5 * t1.x, v[0].y - (T)0.5 * t1.y};
      cx<T> t3 = cmuli(csub(v[1], v[2]), s3);
      y[0] = cadd(v[0], t1);
      y[1] = cadd(t2, t3);
      y[2] = csub(t2, t3);
    }
    const int idxD = g * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; r++) out[idxD + r * Ns] = y[r];
  }
}

// Stockham transform of length L = 2^a 3^b <= 4096 held in LDS buffers A
// and B (radix 3 passes first, then radix 8, then one radix 4 or 2 if
// needed); returns the buffer holding the result.  dir = -1 is the forward
// exp(-2 pi i jk/L), +1 the un-normalised inverse.  tw holds
// exp(-2 pi i m/L), m < L.
template <typename T>
__device__ __forceinline__ cx<T> *fft_lds(cx<T> *A, cx<T> *B, int L, const cx<T> *tw, int dir) {
  int Ns = 1, rem = L;
  cx<T> *in = A, *out = B;
  while (rem > 1) {
    const int R = (rem % 3 == 0) ? 3 : ((rem % 8 == 0) ? 8 : ((rem % 4 == 0) ? 4 : 2));
    const int span = L / (Ns * R);
    const unsigned mg = (Ns > 1) ? 0xFFFFFFFFu / (unsigned)Ns + 1u : 0u;
    if (R == 8) fft_pass<T, 8>(in, out, L, Ns, mg, span, tw, dir);
    else if (R == 4) fft_pass<T, 4>(in, out, L, Ns, mg, span, tw, dir);
    else if (R == 3) fft_pass<T, 3>(in, out, L, Ns, mg, span, tw, dir);
    else fft_pass<T, 2>(in, out, L, Ns, mg, span, tw, dir);
    __syncthreads();
    cx<T> *tmp = in;
    in = out;
    out = tmp;
    Ns *= R;
    rem /= R;
  }
  return in;
}

template <typename T> __device__ __forceinline__ double ldd(const void *p, long i) {
  return (double)((const T *)p)[i];
}

// compute_fftdensity (src/sgdens.c:54-70): Sr = Sigma sqrt(r/r0),
// St = Sr r/r0 on the first Nrad rows, then the azimuthal transform
template <typename T> __global__ void __launch_bounds__(256) k_sg_rows_fwd(KArgs a, int force) {
  extern __shared__ unsigned char smem[];
  if (!force && !a.ctrl->active) return;
  const int i = blockIdx.x, ns = a.ns;
  cx<T> *A = (cx<T> *)smem, *B = A + ns;
  for (int j = threadIdx.x; j < ns; j += blockDim.x) {
    double sr = ldd<T>(a.rho, (long)i * ns + j) * a.sg_sqrt[i];
    double stv = sr * a.Rmed[i] / a.Rmed[0];
    A[j] = {(T)sr, (T)stv};
  }
  __syncthreads();
  cx<T> *r = fft_lds<T>(A, B, ns, (const cx<T> *)a.tw_ns, -1);
  cx<T> *z = (cx<T> *)a.sg_z + (long)i * ns;
  for (int j = threadIdx.x; j < ns; j += blockDim.x) z[j] = r[j];
}

// Radial transforms and spectral product (src/sgacc.c:25-42) for one pair
// of azimuthal wavenumbers (k2, Ns-k2), whose spectra are tied by the
// Hermitian symmetry of the two real reduced densities.  One workgroup:
//   1. loads both columns of Z (rows >= Nrad are the zero padding),
//   2. transforms them over 2 Nrad,
//   3. separates S_r(k) = (Z(k) + conj Z(-k))/2, S_t(k) = (Z(k) - conj Z(-k))/2i,
//      multiplies by the kernel spectra (K(k1, Ns-k2) = conj K(-k1, k2), so
//      only wavenumbers 0..Ns/2 are stored, transposed) and recombines
//      A_r + i A_t,
//   4. inverse-transforms both columns and keeps rows < Nrad.
template <typename T> __global__ void __launch_bounds__(256) k_sg_cols(KArgs a, int force) {
  extern __shared__ unsigned char smem[];
  if (!force && !a.ctrl->active) return;
  const int ns = a.ns, nr = a.nr, L = 2 * nr, p = blockIdx.x;
  const int ka = p, kb = (ns - p) % ns;
  const bool self = (ka == kb);
  cx<T> *A1 = (cx<T> *)smem, *B1 = A1 + L, *A2 = B1 + L, *B2 = A2 + L;
  const cx<T> *z = (const cx<T> *)a.sg_z;
  const cx<T> zero = {(T)0, (T)0};
  for (int i = threadIdx.x; i < L; i += blockDim.x) {
    A1[i] = (i < nr) ? z[(long)i * ns + ka] : zero;
    if (!self) A2[i] = (i < nr) ? z[(long)i * ns + kb] : zero;
  }
  __syncthreads();
  const cx<T> *tw = (const cx<T> *)a.tw_2n;
  cx<T> *ra = fft_lds<T>(A1, B1, L, tw, -1);
  cx<T> *rb = self ? ra : fft_lds<T>(A2, B2, L, tw, -1);
  cx<T> *wa = (ra == A1) ? B1 : A1;
  cx<T> *wb = self ? wa : ((rb == A2) ? B2 : A2);
  const cx<T> *khr = (const cx<T> *)a.sg_kr + (long)p * L, *kht = (const cx<T> *)a.sg_kt + (long)p * L;
  const T G = (T)DF_G, h = (T)0.5;
  for (int k1 = threadIdx.x; k1 < L; k1 += blockDim.x) {
    const int m1 = (k1 == 0) ? 0 : L - k1;
    {
      const cx<T> zz = ra[k1], zp = rb[m1];
      const cx<T> sr = {h * (zz.x + zp.x), h * (zz.y - zp.y)};
      const cx<T> st = {h * (zz.y + zp.y), h * (zp.x - zz.x)};
      const cx<T> K1 = khr[k1], K2 = kht[k1];
      const cx<T> ar = {-G * (K1.x * sr.x - K1.y * sr.y), -G * (K1.x * sr.y + K1.y * sr.x)};
      const cx<T> at = {-G * (K2.x * st.x - K2.y * st.y), -G * (K2.x * st.y + K2.y * st.x)};
      wa[k1] = {ar.x - at.y, ar.y + at.x};
    }
    if (!self) {
      const cx<T> zz = rb[k1], zp = ra[m1];
      const cx<T> sr = {h * (zz.x + zp.x), h * (zz.y - zp.y)};
      const cx<T> st = {h * (zz.y + zp.y), h * (zp.x - zz.x)};
      const cx<T> K1 = {khr[m1].x, -khr[m1].y}, K2 = {kht[m1].x, -kht[m1].y};
      const cx<T> ar = {-G * (K1.x * sr.x - K1.y * sr.y), -G * (K1.x * sr.y + K1.y * sr.x)};
      const cx<T> at = {-G * (K2.x * st.x - K2.y * st.y), -G * (K2.x * st.y + K2.y * st.x)};
      wb[k1] = {ar.x - at.y, ar.y + at.x};
    }
  }
  __syncthreads();
  cx<T> *oa = fft_lds<T>(wa, (wa == A1) ? B1 : A1, L, tw, +1);
  cx<T> *ob = self ? oa : fft_lds<T>(wb, (wb == A2) ? B2 : A2, L, tw, +1);
  cx<T> *w = (cx<T> *)a.sg_w;
  for (int i = threadIdx.x; i < nr; i += blockDim.x) {
    w[(long)i * ns + ka] = oa[i];
    if (!self) w[(long)i * ns + kb] = ob[i];
  }
}

// inverse azimuthal transform, normalisation and self-force compensation
// (src/sgacc.c:137-161)
template <typename T> __global__ void __launch_bounds__(256) k_sg_rows_inv(KArgs a, int force) {
  extern __shared__ unsigned char smem[];
  if (!force && !a.ctrl->active) return;
  const int i = blockIdx.x, ns = a.ns;
  cx<T> *A = (cx<T> *)smem, *B = A + ns;
  const cx<T> *w = (const cx<T> *)a.sg_w + (long)i * ns;
  for (int j = threadIdx.x; j < ns; j += blockDim.x) A[j] = w[j];
  __syncthreads();
  cx<T> *r = fft_lds<T>(A, B, ns, (const cx<T> *)a.tw_ns, +1);
  for (int j = threadIdx.x; j < ns; j += blockDim.x) {
    const long l = (long)i * ns + j;
    double accr = (double)r[j].x * a.sg_normr[i];
    double acct = (double)r[j].y * a.sg_normt[i];
    accr += DF_G * ldd<T>(a.rho, l) * a.sg_rstep * a.sg_tstep / a.sg_eps;
    ((T *)a.sgr)[l] = (T)accr;
    ((T *)a.sgt)[l] = (T)acct;
  }
}

// update_sgvelocity (src/sgupdate.c:30-72) on VradInt / VthetaInt
template <typename T> __global__ void __launch_bounds__(256) k_sg_kick(KArgs a) {
  const dfargo_ctrl *c = a.ctrl;
  if (!c->active) return;
  const int i = blockIdx.y, ns = a.ns, j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ns) return;
  const long l = (long)i * ns + j;
  const double Dt = c->dt;
  if (i > 0) {
    double radsgacc = ((a.Rinf[i] - a.Rmed[i - 1]) * ldd<T>(a.sgr, l) +
                       (a.Rmed[i] - a.Rinf[i]) * ldd<T>(a.sgr, l - ns)) * a.InvDiffRmed[i];
    ((T *)a.vri)[l] = (T)(ldd<T>(a.vri, l) + Dt * radsgacc);
  }
  const long lm1 = (long)i * ns + (j == 0 ? ns - 1 : j - 1);
  double azisgacc = 0.5 * (ldd<T>(a.sgt, l) + ldd<T>(a.sgt, lm1));
  ((T *)a.vti)[l] = (T)(ldd<T>(a.vti, l) + Dt * azisgacc);
}

// generic in-place batched transforms (kernel spectra at set-up)
template <typename T>
__global__ void __launch_bounds__(256) k_fft_rows(cx<T> *data, int L, const cx<T> *tw, int dir) {
  extern __shared__ unsigned char smem[];
  cx<T> *A = (cx<T> *)smem, *B = A + L, *row = data + (long)blockIdx.x * L;
  for (int j = threadIdx.x; j < L; j += blockDim.x) A[j] = row[j];
  __syncthreads();
  cx<T> *r = fft_lds<T>(A, B, L, tw, dir);
  for (int j = threadIdx.x; j < L; j += blockDim.x) row[j] = r[j];
}
template <typename T>
__global__ void __launch_bounds__(256) k_fft_cols(cx<T> *data, int ncols, int L, const cx<T> *tw, int dir) {
  extern __shared__ unsigned char smem[];
  cx<T> *A = (cx<T> *)smem, *B = A + L;
  const int c = blockIdx.x;
  for (int i = threadIdx.x; i < L; i += blockDim.x) A[i] = data[(long)i * ncols + c];
  __syncthreads();
  cx<T> *r = fft_lds<T>(A, B, L, tw, dir);
  for (int i = threadIdx.x; i < L; i += blockDim.x) data[(long)i * ncols + c] = r[i];
}
__global__ void k_convert_d2f(const double *src, float *dst, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (float)src[i];
}

// ------------------------------------------------------------------------

template <typename K> static void allow_lds(K kernel, size_t bytes) {
  // dynamic LDS above 64 KiB must be requested (gfx950 has 160 KiB per CU)
  if (bytes > 65536) hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <typename T> static void sg_pipeline(int id, const KArgs &a, hipStream_t s, int force) {
  const size_t rows_lds = 2 * sizeof(cx<T>) * a.ns, cols_lds = 4 * sizeof(cx<T>) * 2 * a.nr;
  switch (id) {
    case KID_SG_ROWS_FWD:
      hipLaunchKernelGGL(k_sg_rows_fwd<T>, dim3(a.nr), dim3(256), rows_lds, s, a, force);
      break;
    case KID_SG_COLS:
      hipLaunchKernelGGL(k_sg_cols<T>, dim3(a.ns / 2 + 1), dim3(256), cols_lds, s, a, force);
      break;
    case KID_SG_ROWS_INV:
      hipLaunchKernelGGL(k_sg_rows_inv<T>, dim3(a.nr), dim3(256), rows_lds, s, a, force);
      break;
    case KID_SG_KICK: {
      int bx = a.ns < 256 ? a.ns : 256;
      hipLaunchKernelGGL(k_sg_kick<T>, dim3((a.ns + bx - 1) / bx, a.nr), dim3(bx), 0, s, a);
      break;
    }
  }
}

// Raise the dynamic LDS limit of the kernels that need more than 64 KiB,
// once at set-up (not inside a stream capture).
template <typename T> void sg_prepare(const KArgs &a) {
  const size_t rows_lds = 2 * sizeof(cx<T>) * a.ns, cols_lds = 4 * sizeof(cx<T>) * 2 * a.nr;
  allow_lds(k_sg_rows_fwd<T>, rows_lds);
  allow_lds(k_sg_cols<T>, cols_lds);
  allow_lds(k_sg_rows_inv<T>, rows_lds);
}

template <typename T> void launch_sg_kernel(int id, const KArgs &a, hipStream_t s) {
  sg_pipeline<T>(id, a, s, 0);
}
template <typename T> void launch_sg_accel(const KArgs &a, hipStream_t s) {
  sg_pipeline<T>(KID_SG_ROWS_FWD, a, s, 1);
  sg_pipeline<T>(KID_SG_COLS, a, s, 1);
  sg_pipeline<T>(KID_SG_ROWS_INV, a, s, 1);
}
template <typename T>
void launch_fft_rows(void *data, int nrows, int L, const void *tw, int dir, hipStream_t s) {
  allow_lds(k_fft_rows<T>, 2 * sizeof(cx<T>) * L);
  hipLaunchKernelGGL(k_fft_rows<T>, dim3(nrows), dim3(256), 2 * sizeof(cx<T>) * L, s,
                     (cx<T> *)data, L, (const cx<T> *)tw, dir);
}
template <typename T>
void launch_fft_cols(void *data, int ncols, int L, const void *tw, int dir, hipStream_t s) {
  allow_lds(k_fft_cols<T>, 2 * sizeof(cx<T>) * L);
  hipLaunchKernelGGL(k_fft_cols<T>, dim3(ncols), dim3(256), 2 * sizeof(cx<T>) * L, s,
                     (cx<T> *)data, ncols, L, (const cx<T> *)tw, dir);
}
void launch_convert_d2f(const double *src, float *dst, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_convert_d2f, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
}

template void sg_prepare<double>(const KArgs &);
template void sg_prepare<float>(const KArgs &);
template void launch_sg_kernel<double>(int, const KArgs &, hipStream_t);
template void launch_sg_kernel<float>(int, const KArgs &, hipStream_t);
template void launch_sg_accel<double>(const KArgs &, hipStream_t);
template void launch_sg_accel<float>(const KArgs &, hipStream_t);
template void launch_fft_rows<double>(void *, int, int, const void *, int, hipStream_t);
template void launch_fft_cols<double>(void *, int, int, const void *, int, hipStream_t);
template void launch_fft_rows<float>(void *, int, int, const void *, int, hipStream_t);
template void launch_fft_cols<float>(void *, int, int, const void *, int, hipStream_t);
