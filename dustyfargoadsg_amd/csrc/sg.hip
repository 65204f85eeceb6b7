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

#include <atomic>

#include "fft_lds.h"
#include "kernels.h"


// compute_fftdensity (src/sgdens.c:54-70) and the azimuthal transform of
// one ring of the stored density.  Z is stored transposed, [Nsec][Nrad]
// (zt), so that the radial transforms read contiguous columns.  The rings
// of one XCD are contiguous (ring_of_block): the 16 (FP32) or 8 (FP64)
// rings that share a line of zt are written in one L2, which merges the
// lines before they go to HBM.
// NB: LDS buffers per transform (2: Stockham ping-pong; 1: in place, half
// the LDS).
template <typename T, int NB>
__global__ void __launch_bounds__(SG_TB_MAX) k_sg_rows_fwd(KArgs a, int force, cx<T> *zt) {
  extern __shared__ unsigned char smem[];
  if (!force && !a.ctrl->go) return;
  const int i = ring_of_block(a), ns = a.ns;
  cx<T> *A = (cx<T> *)smem, *B = (NB == 2) ? A + fpsz(ns) : A;
  for (int j = threadIdx.x; j < ns; j += blockDim.x) A[fp(j)] = sg_reduced<T>(a, i, ldd<T>(a.rho, (long)i * ns + j));
  sg_row_transform<T, NB>(a, i, A, B, zt);
}

// The spectral product of wavenumber k1 of the pair (src/sgacc.c:25-42):
// from the transformed columns ra (wavenumber k2) and rb (Ns-k2), the new
// values of both at k1 (wb only when the pair is two columns).
template <typename T>
__device__ __forceinline__ void sg_spectral(const cx<T> *ra, const cx<T> *rb, int k1, int L, const cx<T> *khr,
                                            const cx<T> *kht, bool self, cx<T> &wa, cx<T> &wb) {
  const T G = (T)DF_G, h = (T)0.5;
  const int m1 = (k1 == 0) ? 0 : L - k1;
  {
    const cx<T> zz = ra[fp(k1)], zp = rb[fp(m1)];
    const cx<T> sr = {h * (zz.x + zp.x), h * (zz.y - zp.y)};
    const cx<T> st = {h * (zz.y + zp.y), h * (zp.x - zz.x)};
    const cx<T> K1 = khr[k1], K2 = kht[k1];
    const cx<T> ar = {-G * (K1.x * sr.x - K1.y * sr.y), -G * (K1.x * sr.y + K1.y * sr.x)};
    const cx<T> at = {-G * (K2.x * st.x - K2.y * st.y), -G * (K2.x * st.y + K2.y * st.x)};
    wa = {ar.x - at.y, ar.y + at.x};
  }
  if (!self) {
    const cx<T> zz = rb[fp(k1)], zp = ra[fp(m1)];
    const cx<T> sr = {h * (zz.x + zp.x), h * (zz.y - zp.y)};
    const cx<T> st = {h * (zz.y + zp.y), h * (zp.x - zz.x)};
    const cx<T> K1 = {khr[m1].x, -khr[m1].y}, K2 = {kht[m1].x, -kht[m1].y};
    const cx<T> ar = {-G * (K1.x * sr.x - K1.y * sr.y), -G * (K1.x * sr.y + K1.y * sr.x)};
    const cx<T> at = {-G * (K2.x * st.x - K2.y * st.y), -G * (K2.x * st.y + K2.y * st.x)};
    wb = {ar.x - at.y, ar.y + at.x};
  }
}

// Radial transforms and spectral product (src/sgacc.c:25-42) for one pair
// of azimuthal wavenumbers (k2, Ns-k2), whose spectra are tied by the
// Hermitian symmetry of the two real reduced densities.  Z and W are stored
// transposed here ([Nsec][Nrad]), so each column is contiguous.  One
// workgroup:
//   1. loads both columns of Z (rows >= Nrad are the zero padding),
//   2. transforms them over 2 Nrad,
//   3. separates S_r(k) = (Z(k) + conj Z(-k))/2, S_t(k) = (Z(k) - conj Z(-k))/2i,
//      multiplies by the kernel spectra (K(k1, Ns-k2) = conj K(-k1, k2), so
//      only wavenumbers 0..Ns/2 are stored, transposed) and recombines
//      A_r + i A_t,
//   4. inverse-transforms both columns and keeps rows < Nrad.
template <typename T, int NB>
__global__ void __launch_bounds__(SG_TB_MAX) k_sg_cols(KArgs a, int force, const cx<T> *zt, cx<T> *wt) {
  extern __shared__ unsigned char smem[];
  if (!force && !a.ctrl->go) return;
  const int ns = a.ns, nr = a.nr, L = 2 * nr, p = blockIdx.x, S = fpsz(L);
  const int ka = p, kb = (ns - p) % ns;
  const bool self = (ka == kb);
  // two buffers per column (A1/B1, A2/B2), or one (A1, A2)
  cx<T> *A1 = (cx<T> *)smem, *B1 = A1 + S, *A2 = (NB == 2) ? B1 + S : A1 + S, *B2 = A2 + S;
  const cx<T> *za = zt + (long)ka * nr, *zb = zt + (long)kb * nr;
  const cx<T> zero = {(T)0, (T)0};
  for (int i = threadIdx.x; i < L; i += blockDim.x) {
    A1[fp(i)] = (i < nr) ? za[i] : zero;
    if (!self) A2[fp(i)] = (i < nr) ? zb[i] : zero;
  }
  __syncthreads();
  const cx<T> *tw = (const cx<T> *)a.tw_2n;
  const cx<T> *khr = (const cx<T> *)a.sg_kr + (long)p * L, *kht = (const cx<T> *)a.sg_kt + (long)p * L;
  cx<T> *oa, *ob;
  if constexpr (NB == 1) {
    if (self) fft_lds_ip<T, 1>(A1, A1, L, tw, -1);
    else fft_lds_ip<T, 2>(A1, A2, L, tw, -1);
    cx<T> *rb = self ? A1 : A2;
    // wavenumbers k1 and L-k1 of both columns read and write the same four
    // slots, so one thread takes both: all four read before any is stored
    for (int k1 = threadIdx.x; k1 <= L / 2; k1 += blockDim.x) {
      const int m1 = (k1 == 0) ? 0 : L - k1;
      cx<T> va, vb, wa, wb;
      sg_spectral<T>(A1, rb, k1, L, khr, kht, self, va, vb);
      if (m1 != k1) sg_spectral<T>(A1, rb, m1, L, khr, kht, self, wa, wb);
      A1[fp(k1)] = va;
      if (!self) rb[fp(k1)] = vb;
      if (m1 != k1) {
        A1[fp(m1)] = wa;
        if (!self) rb[fp(m1)] = wb;
      }
    }
    __syncthreads();
    if (self) fft_lds_ip<T, 1>(A1, A1, L, tw, +1);
    else fft_lds_ip<T, 2>(A1, A2, L, tw, +1);
    oa = A1;
    ob = rb;
  } else {
    cx<T> *ra, *rb;
    if (self) {
      ra = rb = fft_lds<T>(A1, B1, L, tw, -1);
    } else {
      fft_lds2<T>(A1, B1, A2, B2, L, tw, -1, &ra, &rb);
    }
    cx<T> *wa = (ra == A1) ? B1 : A1;
    cx<T> *wb = self ? wa : ((rb == A2) ? B2 : A2);
    for (int k1 = threadIdx.x; k1 < L; k1 += blockDim.x) {
      cx<T> va, vb;
      sg_spectral<T>(ra, rb, k1, L, khr, kht, self, va, vb);
      wa[fp(k1)] = va;
      if (!self) wb[fp(k1)] = vb;
    }
    __syncthreads();
    if (self) {
      oa = ob = fft_lds<T>(wa, (wa == A1) ? B1 : A1, L, tw, +1);
    } else {
      fft_lds2<T>(wa, (wa == A1) ? B1 : A1, wb, (wb == A2) ? B2 : A2, L, tw, +1, &oa, &ob);
    }
  }
  cx<T> *wpa = wt + (long)ka * nr, *wpb = wt + (long)kb * nr;
  for (int i = threadIdx.x; i < nr; i += blockDim.x) {
    wpa[i] = oa[fp(i)];
    if (!self) wpb[i] = ob[fp(i)];
  }
}

// ring i's accelerations from its inverse-transformed row r (padded
// layout): normalisation and self-force compensation
template <typename T> __device__ __forceinline__ void sg_accel_store(const KArgs &a, int i, const cx<T> *r) {
  const int ns = a.ns;
  for (int j = threadIdx.x; j < ns; j += blockDim.x) {
    const long l = (long)i * ns + j;
    double accr = (double)r[fp(j)].x * a.sg_normr[i];
    double acct = (double)r[fp(j)].y * a.sg_normt[i];
    accr += DF_G * ldd<T>(a.rho, l) * a.sg_rstep * a.sg_tstep / a.sg_eps;
    ((T *)a.sgr)[l] = (T)accr;
    ((T *)a.sgt)[l] = (T)acct;
  }
}

// inverse azimuthal transform, normalisation and self-force compensation
// (src/sgacc.c:137-161); W stored transposed ([Nsec][Nrad], wt), read
// with the rings of one XCD contiguous as in k_sg_rows_fwd
template <typename T, int NB>
__global__ void __launch_bounds__(SG_TB_MAX) k_sg_rows_inv(KArgs a, int force, const cx<T> *wt) {
  extern __shared__ unsigned char smem[];
  if (!force && !a.ctrl->go) return;
  const int i = ring_of_block(a), ns = a.ns, nr = a.nr;
  cx<T> *A = (cx<T> *)smem, *B = A + fpsz(ns);
  for (int j = threadIdx.x; j < ns; j += blockDim.x) A[fp(j)] = wt[(long)j * nr + i];
  __syncthreads();
  cx<T> *r = A;
  if constexpr (NB == 1) fft_lds_ip<T, 1>(A, A, ns, (const cx<T> *)a.tw_ns, +1);
  else r = fft_lds<T>(A, B, ns, (const cx<T> *)a.tw_ns, +1);
  sg_accel_store<T>(a, i, r);
}

// Two neighbouring rings per workgroup, each transformed in place in its
// own buffer: the passes of both share their barriers and the transposed
// accesses of zt / wt read or write both rings' elements of a wavenumber
// side by side.  The butterflies are those of the other forms, so the
// results are the same bit for bit.  Needs Nrad even (self-gravity takes a
// power of two).
template <typename T>
__global__ void __launch_bounds__(SG_TB_MAX) k_sg_rows_fwd_pair(KArgs a, int force, cx<T> *zt) {
  extern __shared__ unsigned char smem[];
  if (!force && !a.ctrl->go) return;
  const int i = 2 * ring_of_block(a), ns = a.ns, nr = a.nr;
  cx<T> *A0 = (cx<T> *)smem, *A1 = A0 + fpsz(ns);
  for (int j = threadIdx.x; j < ns; j += blockDim.x) {
    A0[fp(j)] = sg_reduced<T>(a, i, ldd<T>(a.rho, (long)i * ns + j));
    A1[fp(j)] = sg_reduced<T>(a, i + 1, ldd<T>(a.rho, (long)(i + 1) * ns + j));
  }
  __syncthreads();
  fft_lds_ip<T, 2>(A0, A1, ns, (const cx<T> *)a.tw_ns, -1);
  for (int j = threadIdx.x; j < ns; j += blockDim.x) {
    zt[(long)j * nr + i] = A0[fp(j)];
    zt[(long)j * nr + i + 1] = A1[fp(j)];
  }
}
template <typename T>
__global__ void __launch_bounds__(SG_TB_MAX) k_sg_rows_inv_pair(KArgs a, int force, const cx<T> *wt) {
  extern __shared__ unsigned char smem[];
  if (!force && !a.ctrl->go) return;
  const int i = 2 * ring_of_block(a), ns = a.ns, nr = a.nr;
  cx<T> *A0 = (cx<T> *)smem, *A1 = A0 + fpsz(ns);
  for (int j = threadIdx.x; j < ns; j += blockDim.x) {
    A0[fp(j)] = wt[(long)j * nr + i];
    A1[fp(j)] = wt[(long)j * nr + i + 1];
  }
  __syncthreads();
  fft_lds_ip<T, 2>(A0, A1, ns, (const cx<T> *)a.tw_ns, +1);
  sg_accel_store<T>(a, i, A0);
  sg_accel_store<T>(a, i + 1, A1);
}

// generic in-place batched transforms (kernel spectra at set-up)
template <typename T>
__global__ void __launch_bounds__(256) k_fft_rows(cx<T> *data, int L, const cx<T> *tw, int dir) {
  extern __shared__ unsigned char smem[];
  cx<T> *A = (cx<T> *)smem, *B = A + fpsz(L), *row = data + (long)blockIdx.x * L;
  for (int j = threadIdx.x; j < L; j += blockDim.x) A[fp(j)] = row[j];
  __syncthreads();
  cx<T> *r = fft_lds<T>(A, B, L, tw, dir);
  for (int j = threadIdx.x; j < L; j += blockDim.x) row[j] = r[fp(j)];
}
template <typename T>
__global__ void __launch_bounds__(256) k_fft_cols(cx<T> *data, int ncols, int L, const cx<T> *tw, int dir) {
  extern __shared__ unsigned char smem[];
  cx<T> *A = (cx<T> *)smem, *B = A + fpsz(L);
  const int c = blockIdx.x;
  for (int i = threadIdx.x; i < L; i += blockDim.x) A[fp(i)] = data[(long)i * ncols + c];
  __syncthreads();
  cx<T> *r = fft_lds<T>(A, B, L, tw, dir);
  for (int i = threadIdx.x; i < L; i += blockDim.x) data[(long)i * ncols + c] = r[fp(i)];
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

static size_t rows_lds_bytes(const KArgs &a, size_t csz, int nb) { return sg_rows_lds_bytes(a, csz, nb); }
static size_t cols_lds_bytes(const KArgs &a, size_t csz, int nb) { return 2 * nb * csz * fpsz(2 * a.nr); }

// VGPRs per lane of the transform kernels, [FP64][column pairs][NB - 1]
// (the larger of the two row kernels), read at set-up; 0 before
static std::atomic<int> s_vgprs[2][2][2];

// Workgroups of tb threads of a transform kernel that share a CU: as many
// as its LDS (160 KiB) and its registers (512 VGPRs per SIMD lane, at most
// 8 wavefronts per SIMD; a workgroup spreads its wavefronts over the 4
// SIMDs) allow, and no more than the grid gives each CU.
static int sg_groups(const KArgs &a, size_t lds, int nwg, int tb, int vgprs) {
  int g = (int)((160 * 1024) / lds);
  const int per_grid = (nwg + a.num_cu - 1) / a.num_cu;
  if (per_grid < g) g = per_grid;
  if (vgprs > 0 && tb > 0) {
    int per_simd = 512 / ((vgprs + 7) / 8 * 8);
    if (per_simd > 8) per_simd = 8;
    const int need = ((tb + 63) / 64 + 3) / 4;
    if (per_simd / need < g) g = per_simd / need;
  }
  return g;
}
// Buffers per transform of the row (cols = 0) or column-pair (cols = 1)
// kernels: the tuning's choice (a.sg_fft_buffers 1 or 2), or by default
// one buffer where that lets more workgroups share a CU while they hold
// no more than 16 wavefronts of it.  The transforms run beside the CFL and
// control kernels: filling every wavefront slot of the CUs with them
// (1024x3072, FP32) delays those kernels by more than the transforms gain.
static int sg_nb(const KArgs &a, size_t csz, int cols) {
  const int L = cols ? 2 * a.nr : a.ns, ntr = cols ? 2 : 1, nwg = cols ? a.ns / 2 + 1 : a.nr;
  const int cap = cols ? 2 * a.nr : a.ns / 2, fp64 = csz == sizeof(cx<double>);
  const size_t l2 = cols ? cols_lds_bytes(a, csz, 2) : rows_lds_bytes(a, csz, 2);
  const int t1 = sg_threads(a, l2 / 2, nwg, cap, fft_ip_min_threads(ntr, L)), t2 = sg_threads(a, l2, nwg, cap);
  if (a.sg_fft_buffers == 1 && t1 > 0) return 1;
  if (a.sg_fft_buffers != 0) return 2;
  const int g1 = sg_groups(a, l2 / 2, nwg, t1, s_vgprs[fp64][cols][0].load());
  const int g2 = sg_groups(a, l2, nwg, t2, s_vgprs[fp64][cols][1].load());
  return (t1 > 0 && g1 > g2 && g1 * t1 <= 16 * 64) ? 1 : 2;
}


// ------------------------------------------------------------------------
// SELFGRAVITY Z: compute_SGZeroMode (src/sgzero.c:3-84), the axisymmetric
// part only.  The reference convolves, cyclically over 2 Nrad log-radius
// cells, the reduced ring-mean density Sr[k] = <Sigma>_k sqrt(Rmed_k/Rmed_0)
// (zero beyond Nrad) with the reduced kernel summed over the azimuths,
// through real FFTs of length 2 Nrad.  Only Nrad outputs and Nrad non-zero
// inputs count, so the direct sum -- Nrad^2 products, one thread per output
// ring with Sr staged in LDS and the kernel read from L2 -- is the cheaper
// form on the GPU; it equals the transform product in exact arithmetic.
// The result goes into every cell of its ring of the SG_Accr grid, so the
// velocity update reads it as it reads the full field; SG_Acct stays zero.
// ------------------------------------------------------------------------
enum { SGZ_TB = 256, SGZ_CHUNK = 2048 };

__device__ __forceinline__ double sgz_wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ring means of Sigma (mpi_make1Dprofile), one workgroup per ring
template <typename T> __global__ void __launch_bounds__(SGZ_TB) k_sgz_mean(KArgs a, int force) {
  if (!force && !a.ctrl->go) return;
  const int i = blockIdx.x, ns = a.ns, tid = threadIdx.x;
  __shared__ double part[SGZ_TB / 64];
  double s = 0.0;
  for (int j = tid; j < ns; j += SGZ_TB) s += ldd<T>(a.rho, (long)i * ns + j);
  s = sgz_wave_sum(s);
  if ((tid & 63) == 0) part[tid >> 6] = s;
  __syncthreads();
  if (tid == 0) {
    double m = 0.0;
    for (int w = 0; w < SGZ_TB / 64; w++) m += part[w];
    a.sgz_mean[i] = m / (double)ns;
  }
}

// SG_Accr[i] = -G rstep (K * Sr)[i] / sqrt(Rmed_i/Rmed_0)
//              + G <Sigma>_i rstep integrale
template <typename T> __global__ void __launch_bounds__(SGZ_TB) k_sgz_conv(KArgs a, int force) {
  if (!force && !a.ctrl->go) return;
  __shared__ double sS[SGZ_CHUNK];
  const int nr = a.nr, n2 = 2 * nr, tid = threadIdx.x, i = blockIdx.x * SGZ_TB + tid;
  double acc = 0.0;
  for (int k0 = 0; k0 < nr; k0 += SGZ_CHUNK) {
    const int kn = (nr - k0 < SGZ_CHUNK) ? nr - k0 : SGZ_CHUNK;
    __syncthreads();
    for (int k = tid; k < kn; k += SGZ_TB) sS[k] = a.sgz_mean[k0 + k] * a.sg_sqrt[k0 + k];
    __syncthreads();
    if (i < nr)
      for (int k = 0; k < kn; k++) {
        int d = i - (k0 + k);
        d += (d < 0) ? n2 : 0;
        acc += a.sgz_kr[d] * sS[k];
      }
  }
  if (i < nr) {
    double v = -DF_G * acc * a.sg_rstep / a.sg_sqrt[i];
    v += DF_G * a.sgz_mean[i] * a.sg_rstep * a.sgz_integrale;
    a.sgz_acc[i] = v;
  }
}

// the ring values into the SG_Accr grid
template <typename T> __global__ void __launch_bounds__(256) k_sgz_fill(KArgs a, int force) {
  if (!force && !a.ctrl->go) return;
  const int i = blockIdx.y, ns = a.ns, j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ns) return;
  ((T *)a.sgr)[(long)i * ns + j] = (T)a.sgz_acc[i];
}

template <typename T> static void sgz_pipeline(int id, const KArgs &a, hipStream_t s, int force) {
  const int bx = a.ns < 256 ? a.ns : 256;
  switch (id) {
    case KID_SGZ_MEAN: hipLaunchKernelGGL(k_sgz_mean<T>, dim3(a.nr), dim3(SGZ_TB), 0, s, a, force); break;
    case KID_SGZ_CONV:
      hipLaunchKernelGGL(k_sgz_conv<T>, dim3((a.nr + SGZ_TB - 1) / SGZ_TB), dim3(SGZ_TB), 0, s, a, force);
      break;
    case KID_SGZ_FILL:
      hipLaunchKernelGGL(k_sgz_fill<T>, dim3((a.ns + bx - 1) / bx, a.nr), dim3(bx), 0, s, a, force);
      break;
  }
}

// rows_fwd: Z^T (sg_z) -> pair transforms -> W^T (sg_w) -> rows_inv
// Rings per workgroup of the row transforms: the tuning's choice
// (a.sg_rows_pair 1 or 2), or by default two where the rows run in place
// anyway (sg_nb), Nrad is even and the thread count fits a workgroup.
// Measured (ms per sub-step, one against two rings): 1024x3072 FP64 (rows
// in place) 0.620 / 0.613, FP32 (two buffers) 0.461 / 0.465; 512x1536
// within 0.3 % either way.
static int sg_rows_per_group(const KArgs &a, size_t csz) {
  const bool fits = (a.nr % 2 == 0) && fft_ip_min_threads(2, a.ns) <= SG_TB_MAX;
  if (!fits || a.sg_rows_pair == 1) return 1;
  if (a.sg_rows_pair == 2) return 2;
  return sg_nb(a, csz, 0) == 1 ? 2 : 1;
}

template <typename T> static void sg_pipeline(int id, const KArgs &a, hipStream_t s, int force) {
  const size_t csz = sizeof(cx<T>);
  if ((id == KID_SG_ROWS_FWD || id == KID_SG_ROWS_INV) && sg_rows_per_group(a, csz) == 2) {
    const size_t lds = rows_lds_bytes(a, csz, 2);
    const int tb = sg_threads(a, lds, a.nr / 2, a.ns, fft_ip_min_threads(2, a.ns));
    if (id == KID_SG_ROWS_FWD)
      hipLaunchKernelGGL(k_sg_rows_fwd_pair<T>, dim3(a.nr / 2), dim3(tb), lds, s, a, force, (cx<T> *)a.sg_z);
    else
      hipLaunchKernelGGL(k_sg_rows_inv_pair<T>, dim3(a.nr / 2), dim3(tb), lds, s, a, force, (const cx<T> *)a.sg_w);
    return;
  }
  const int rnb = sg_nb(a, csz, 0), cnb = sg_nb(a, csz, 1);
  const size_t rows_lds = rows_lds_bytes(a, csz, rnb), cols_lds = cols_lds_bytes(a, csz, cnb);
  const int rows_tb = sg_threads(a, rows_lds, a.nr, a.ns / 2, rnb == 1 ? fft_ip_min_threads(1, a.ns) : 0);
  const int cols_tb =
      sg_threads(a, cols_lds, a.ns / 2 + 1, 2 * a.nr, cnb == 1 ? fft_ip_min_threads(2, 2 * a.nr) : 0);
  cx<T> *z = (cx<T> *)a.sg_z, *w = (cx<T> *)a.sg_w;
  switch (id) {
    case KID_SG_ROWS_FWD: {
      auto kf = (rnb == 1) ? k_sg_rows_fwd<T, 1> : k_sg_rows_fwd<T, 2>;
      hipLaunchKernelGGL(kf, dim3(a.nr), dim3(rows_tb), rows_lds, s, a, force, z);
      break;
    }
    case KID_SG_COLS: {
      auto kf = (cnb == 1) ? k_sg_cols<T, 1> : k_sg_cols<T, 2>;
      hipLaunchKernelGGL(kf, dim3(a.ns / 2 + 1), dim3(cols_tb), cols_lds, s, a, force, (const cx<T> *)z, w);
      break;
    }
    case KID_SG_ROWS_INV: {
      auto kf = (rnb == 1) ? k_sg_rows_inv<T, 1> : k_sg_rows_inv<T, 2>;
      hipLaunchKernelGGL(kf, dim3(a.nr), dim3(rows_tb), rows_lds, s, a, force, (const cx<T> *)w);
      break;
    }
  }
}

// Raise the dynamic LDS limit of the kernels that need more than 64 KiB,
// once at set-up (not inside a stream capture).
template <typename T> static int kernel_vgprs(const void *k) {
  hipFuncAttributes fa;
  return (hipFuncGetAttributes(&fa, k) == hipSuccess) ? fa.numRegs : 0;
}

template <typename T> void sg_prepare(const KArgs &a) {
  const size_t csz = sizeof(cx<T>);
  const int fp64 = sizeof(T) == sizeof(double);
  for (int nb = 1; nb <= 2; nb++) {
    const int rf = kernel_vgprs<T>(nb == 1 ? (const void *)k_sg_rows_fwd<T, 1> : (const void *)k_sg_rows_fwd<T, 2>);
    const int ri = kernel_vgprs<T>(nb == 1 ? (const void *)k_sg_rows_inv<T, 1> : (const void *)k_sg_rows_inv<T, 2>);
    s_vgprs[fp64][0][nb - 1] = rf > ri ? rf : ri;
    s_vgprs[fp64][1][nb - 1] = kernel_vgprs<T>(nb == 1 ? (const void *)k_sg_cols<T, 1> : (const void *)k_sg_cols<T, 2>);
  }
  allow_lds(k_sg_rows_fwd<T, 2>, rows_lds_bytes(a, csz, 2));
  allow_lds(k_sg_cols<T, 2>, cols_lds_bytes(a, csz, 2));
  allow_lds(k_sg_rows_inv<T, 2>, rows_lds_bytes(a, csz, 2));
  allow_lds(k_sg_rows_fwd<T, 1>, rows_lds_bytes(a, csz, 1));
  allow_lds(k_sg_cols<T, 1>, cols_lds_bytes(a, csz, 1));
  allow_lds(k_sg_rows_inv<T, 1>, rows_lds_bytes(a, csz, 1));
  allow_lds(k_sg_rows_fwd_pair<T>, rows_lds_bytes(a, csz, 2));
  allow_lds(k_sg_rows_inv_pair<T>, rows_lds_bytes(a, csz, 2));
}

template <typename T> void launch_sg_kernel(int id, const KArgs &a, hipStream_t s) {
  if (a.sg_zero) sgz_pipeline<T>(id, a, s, 0);
  else sg_pipeline<T>(id, a, s, 0);
}
template <typename T> void launch_sg_accel(const KArgs &a, hipStream_t s) {
  if (a.sg_zero) {
    for (int id : {KID_SGZ_MEAN, KID_SGZ_CONV, KID_SGZ_FILL}) sgz_pipeline<T>(id, a, s, 1);
    return;
  }
  for (int id : {KID_SG_ROWS_FWD, KID_SG_COLS, KID_SG_ROWS_INV})
    sg_pipeline<T>(id, a, s, 1);
}
// the row transforms of the stored density, whatever the loop condition
// (at the start of an AlgoGas call when k_finish forms them afterwards)
template <typename T> void launch_sg_rows_forced(const KArgs &a, hipStream_t s) {
  sg_pipeline<T>(KID_SG_ROWS_FWD, a, s, 1);
}
template <typename T>
void launch_fft_rows(void *data, int nrows, int L, const void *tw, int dir, hipStream_t s) {
  allow_lds(k_fft_rows<T>, 2 * sizeof(cx<T>) * fpsz(L));
  hipLaunchKernelGGL(k_fft_rows<T>, dim3(nrows), dim3(256), 2 * sizeof(cx<T>) * fpsz(L), s,
                     (cx<T> *)data, L, (const cx<T> *)tw, dir);
}
template <typename T>
void launch_fft_cols(void *data, int ncols, int L, const void *tw, int dir, hipStream_t s) {
  allow_lds(k_fft_cols<T>, 2 * sizeof(cx<T>) * fpsz(L));
  hipLaunchKernelGGL(k_fft_cols<T>, dim3(ncols), dim3(256), 2 * sizeof(cx<T>) * fpsz(L), s,
                     (cx<T> *)data, ncols, L, (const cx<T> *)tw, dir);
}
void launch_convert_d2f(const double *src, float *dst, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_convert_d2f, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
}

template void sg_prepare<double>(const KArgs &);
template void sg_prepare<float>(const KArgs &);
template void launch_sg_kernel<double>(int, const KArgs &, hipStream_t);
template void launch_sg_kernel<float>(int, const KArgs &, hipStream_t);
template void launch_sg_rows_forced<double>(const KArgs &, hipStream_t);
template void launch_sg_rows_forced<float>(const KArgs &, hipStream_t);
template void launch_sg_accel<double>(const KArgs &, hipStream_t);
template void launch_sg_accel<float>(const KArgs &, hipStream_t);
template void launch_fft_rows<double>(void *, int, int, const void *, int, hipStream_t);
template void launch_fft_cols<double>(void *, int, int, const void *, int, hipStream_t);
template void launch_fft_rows<float>(void *, int, int, const void *, int, hipStream_t);
template void launch_fft_cols<float>(void *, int, int, const void *, int, hipStream_t);
