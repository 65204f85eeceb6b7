// fft_lds.h -- complex transforms in LDS (Stockham radix 8/4/3/2, one
// workgroup per transform) shared by the self-gravity kernels (sg.hip) and
// the end-of-sub-step kernel that forms the next sub-step's row spectra
// (step.hip).
#ifndef DFARGO_FFT_LDS_H
#define DFARGO_FFT_LDS_H

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

// LDS layout of a transform buffer: one spare slot after every
// 2^DF_FFT_PAD elements, so that the power-of-two strides of the Stockham
// passes spread over the LDS banks.  A buffer of L elements takes fpsz(L)
// slots.
#ifndef DF_FFT_PAD
#define DF_FFT_PAD 4
#endif
__device__ __forceinline__ int fp(int i) { return i + (i >> DF_FFT_PAD); }
__host__ __device__ inline int fpsz(int L) { return L + (L >> DF_FFT_PAD) + 1; }

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
// fft_bfly computes butterfly j into y and returns the index of its first
// output (the outputs go to idxD + r Ns); fft_butterfly also stores them.
template <typename T, int R>
__device__ __forceinline__ int fft_bfly(const cx<T> *in, int j, int nb, int Ns, unsigned mg, int span,
                                        const cx<T> *tw, int dir, cx<T> (&y)[R]) {
  const T s3 = (T)(dir * 0.86602540378443864676); // dir * sqrt(3)/2
  const T c8 = (T)0.70710678118654752440;         // sqrt(1/2)
  const T fd = (T)dir;
  {
    const int g = (Ns == 1) ? j : (int)__umulhi((unsigned)j, mg);
    const int k = j - g * Ns;
    cx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = in[fp(j + r * nb)];
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
    return g * Ns * R + k;
  }
}
template <typename T, int R>
__device__ __forceinline__ void fft_butterfly(const cx<T> *in, cx<T> *out, int j, int nb, int Ns, unsigned mg,
                                              int span, const cx<T> *tw, int dir) {
  cx<T> y[R];
  const int idxD = fft_bfly<T, R>(in, j, nb, Ns, mg, span, tw, dir, y);
#pragma unroll
  for (int r = 0; r < R; r++) out[fp(idxD + r * Ns)] = y[r];
}

template <typename T, int R>
__device__ __forceinline__ void fft_pass(const cx<T> *in, cx<T> *out, int L, int Ns, unsigned mg, int span,
                                         const cx<T> *tw, int dir) {
  const int nb = L / R;
  for (int j = threadIdx.x; j < nb; j += blockDim.x) fft_butterfly<T, R>(in, out, j, nb, Ns, mg, span, tw, dir);
}

// the same pass over two independent transforms of the same length at once:
// the workgroup's threads share their butterflies, and the two transforms
// share the pass's barrier
template <typename T, int R>
__device__ __forceinline__ void fft_pass2(const cx<T> *in0, cx<T> *out0, const cx<T> *in1, cx<T> *out1, int L,
                                          int Ns, unsigned mg, int span, const cx<T> *tw, int dir) {
  const int nb = L / R;
  for (int jj = threadIdx.x; jj < 2 * nb; jj += blockDim.x) {
    const bool second = jj >= nb;
    fft_butterfly<T, R>(second ? in1 : in0, second ? out1 : out0, second ? jj - nb : jj, nb, Ns, mg, span, tw,
                        dir);
  }
}

// Stockham transform of length L = 2^a 3^b <= 4096 held in LDS buffers A
// and B (padded layout fp) (radix 3 passes first, then radix 8, then one radix 4 or 2 if
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

// two transforms of length L in lockstep (buffers A0/B0 and A1/B1); the
// results are in *r0 and *r1
template <typename T>
__device__ __forceinline__ void fft_lds2(cx<T> *A0, cx<T> *B0, cx<T> *A1, cx<T> *B1, int L, const cx<T> *tw,
                                         int dir, cx<T> **r0, cx<T> **r1) {
  int Ns = 1, rem = L;
  cx<T> *in0 = A0, *out0 = B0, *in1 = A1, *out1 = B1;
  while (rem > 1) {
    const int R = (rem % 3 == 0) ? 3 : ((rem % 8 == 0) ? 8 : ((rem % 4 == 0) ? 4 : 2));
    const int span = L / (Ns * R);
    const unsigned mg = (Ns > 1) ? 0xFFFFFFFFu / (unsigned)Ns + 1u : 0u;
    if (R == 8) fft_pass2<T, 8>(in0, out0, in1, out1, L, Ns, mg, span, tw, dir);
    else if (R == 4) fft_pass2<T, 4>(in0, out0, in1, out1, L, Ns, mg, span, tw, dir);
    else if (R == 3) fft_pass2<T, 3>(in0, out0, in1, out1, L, Ns, mg, span, tw, dir);
    else fft_pass2<T, 2>(in0, out0, in1, out1, L, Ns, mg, span, tw, dir);
    __syncthreads();
    cx<T> *t0 = in0, *t1 = in1;
    in0 = out0;
    out0 = t0;
    in1 = out1;
    out1 = t1;
    Ns *= R;
    rem /= R;
  }
  *r0 = in0;
  *r1 = in1;
}

// ---- the same transforms in one buffer per transform (in place) ----
// Each pass reads all of a thread's butterflies into registers, waits for
// the workgroup, then writes the outputs over the inputs and waits again:
// half the LDS of the two-buffer form for one more barrier per pass, so that
// more workgroups share a CU.  A thread holds at most FFT_IP_ELEMS elements
// of a pass: the workgroup needs fft_ip_min_threads(ntr, L) threads or
// more (ntr transforms of length L in lockstep).  The butterflies' arithmetic
// is that of the two-buffer form, so the results are the same bit for bit.
#ifndef FFT_IP_ELEMS
#define FFT_IP_ELEMS 8
#endif
__host__ __device__ inline int fft_ip_min_threads(int ntr, int L) { return (ntr * L + FFT_IP_ELEMS - 1) / FFT_IP_ELEMS; }

template <typename T, int R, int NT>
__device__ __forceinline__ void fft_pass_ip(cx<T> *b0, cx<T> *b1, int L, int Ns, unsigned mg, int span,
                                            const cx<T> *tw, int dir) {
  constexpr int MB = (FFT_IP_ELEMS + R - 1) / R; // butterflies per thread
  const int nb = L / R, tot = NT * nb, TB = blockDim.x;
  cx<T> y[MB][R];
  int dst[MB];
#pragma unroll
  for (int b = 0; b < MB; b++) {
    const int jj = threadIdx.x + b * TB;
    if (jj < tot) {
      const bool second = NT == 2 && jj >= nb;
      dst[b] = fft_bfly<T, R>(second ? b1 : b0, second ? jj - nb : jj, nb, Ns, mg, span, tw, dir, y[b]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < MB; b++) {
    const int jj = threadIdx.x + b * TB;
    if (jj < tot) {
      cx<T> *out = (NT == 2 && jj >= nb) ? b1 : b0;
#pragma unroll
      for (int r = 0; r < R; r++) out[fp(dst[b] + r * Ns)] = y[b][r];
    }
  }
  __syncthreads();
}

// NT (1 or 2) transforms of length L = 2^a 3^b <= 4096 in place in b0 (and
// b1), the passes in the order of fft_lds; the results stay in b0 and b1.
template <typename T, int NT>
__device__ __forceinline__ void fft_lds_ip(cx<T> *b0, cx<T> *b1, int L, const cx<T> *tw, int dir) {
  int Ns = 1, rem = L;
  while (rem > 1) {
    const int R = (rem % 3 == 0) ? 3 : ((rem % 8 == 0) ? 8 : ((rem % 4 == 0) ? 4 : 2));
    const int span = L / (Ns * R);
    const unsigned mg = (Ns > 1) ? 0xFFFFFFFFu / (unsigned)Ns + 1u : 0u;
    if (R == 8) fft_pass_ip<T, 8, NT>(b0, b1, L, Ns, mg, span, tw, dir);
    else if (R == 4) fft_pass_ip<T, 4, NT>(b0, b1, L, Ns, mg, span, tw, dir);
    else if (R == 3) fft_pass_ip<T, 3, NT>(b0, b1, L, Ns, mg, span, tw, dir);
    else fft_pass_ip<T, 2, NT>(b0, b1, L, Ns, mg, span, tw, dir);
    Ns *= R;
    rem /= R;
  }
}

// The transform kernels take up to SG_TB_MAX threads; sg_threads picks the
// block size at launch.
#define SG_TB_MAX 1024

template <typename T> __device__ __forceinline__ double ldd(const void *p, long i) {
  return (double)((const T *)p)[i];
}

// Ring i's reduced densities for the transform (compute_fftdensity,
// src/sgdens.c:54-70): Sr = Sigma sqrt(r/r0) and St = Sr r/r0 carried as
// one complex value.
template <typename T> __device__ __forceinline__ cx<T> sg_reduced(const KArgs &a, int i, double rho) {
  const double sr = rho * a.sg_sqrt[i];
  const double stv = sr * a.Rmed[i] / a.Rmed[0];
  return {(T)sr, (T)stv};
}

// The forward azimuthal transform of ring i's reduced densities, which the
// workgroup has written to A (padded layout), stored transposed into zt
// ([Nsec][Nrad]); with NB = 2 buffers B is the second one (unused with 1).
template <typename T, int NB = 2>
__device__ __forceinline__ void sg_row_transform(const KArgs &a, int i, cx<T> *A, cx<T> *B, cx<T> *zt) {
  __syncthreads();
  cx<T> *r = A;
  if constexpr (NB == 1) fft_lds_ip<T, 1>(A, A, a.ns, (const cx<T> *)a.tw_ns, -1);
  else r = fft_lds<T>(A, B, a.ns, (const cx<T> *)a.tw_ns, -1);
  for (int j = threadIdx.x; j < a.ns; j += blockDim.x) zt[(long)j * a.nr + i] = r[fp(j)];
}

// LDS bytes of one row transform in nb buffers for element size csz
static inline size_t sg_rows_lds_bytes(const KArgs &a, size_t csz, int nb = 2) { return nb * csz * fpsz(a.ns); }

// Block size of a transform kernel of nwg workgroups using lds bytes each:
// the LDS lets 160 KiB / lds workgroups share a CU, the grid gives each CU
// nwg / CUs of them, and together they should hold about 16 wavefronts, so
// that the SIMDs stay busy through the barriers of the passes (256 ..
// SG_TB_MAX threads, a power of two, and no more than the butterflies of
// the widest pass, cap).
// In-place transforms also need tmin threads or more (fft_ip_min_threads);
// the result is 0 when SG_TB_MAX threads are not enough.
static inline int sg_threads(const KArgs &a, size_t lds, int nwg, int cap, int tmin = 0) {
  const size_t per_cu = 160 * 1024;
  int groups = (int)(per_cu / (lds ? lds : 1));
  const int per_grid = (nwg + a.num_cu - 1) / a.num_cu;
  if (per_grid < groups) groups = per_grid;
  const int want = 16 * 64 / (groups > 0 ? groups : 1);
  int t = 256;
  while (t < want && t < SG_TB_MAX && 2 * t <= cap) t *= 2;
  while (t < tmin && t < SG_TB_MAX) t *= 2;
  return (t < tmin) ? 0 : t;
}

#endif
