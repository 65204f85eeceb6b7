// kernels.h -- kernel argument block and launcher declarations shared by the
// HIP translation units of libdfargo.
#pragma once

#include <hip/hip_runtime.h>

#include "dfargo_internal.h"

#define DF_G 1.0
#define DF_PI 3.14159265358979323846 /* M_PI of the host libm headers */
#define DF_CFLSECURITY 0.5           /* src/SourceEuler.c:10 */
#define DF_CVNR 1.41                 /* src/SourceEuler.c:13 */

// The particle state arrays (FP64, int ids).  Two sets: each sub-step
// reads the particles from one, in cell order, and writes them re-sorted by
// cell to the other; id is the particle's index in the reference's arrays.
struct DustSet {
  double *r, *th, *vr, *vth, *l, *size, *stokes;
  int *id;
};

// Everything a kernel needs, passed by value.  Field pointers are untyped:
// they hold T = double or float depending on the context's precision.
struct KArgs {
  int nr, ns;
  // switches decoded from dfargo_params
  int energy, visc_on, alpha_on, fast, floors, bc, sg, beta_cooling, viscous_heating;
  // frame (enum dfargo_frame) and whether it follows the outermost planet
  int frame, frame_outer;
  // wave-killing zones: rings 0 .. wkz_n_in-1 and wkz_i_out .. nr-1
  int wkz_n_in, wkz_i_out, damp_ini, zero_vrad;
  double wkz_rmin, wkz_rmax, damp_ratio;
  const double *VradMed;
  // planet accretion on (some planet has acc > 1e-10), and PMIN
  int accrete;
  double pmin;
  // SELFGRAVITY Z: ring means of Sigma, the radial acceleration per ring,
  // the reduced kernel (2 nr values, times tstep) and the ring's own term
  int sg_zero;
  double sgz_integrale;
  double *sgz_mean, *sgz_acc;
  const double *sgz_kr;
  // DAMPTOVISCOUS: the 1-D viscous disc (n1d cells), its boundary rule
  int n1d, bc1d, damp_visc;
  double *v1d_sigma, *v1d_vrad;
  const double *v1d_rmed, *v1d_nu;
  int gas_indirect, indirect_term, alt_indirect, exclude_hill, roche, forced_circular;
  int xcd_remap;
  // launch shapes the autotuner picks (results do not depend on them):
  // wavefronts per workgroup of the azimuthal transport (1, 2, 4) and
  // threads per workgroup of the velocity/boundary kernel (64, 128, 256)
  int tune_azi_waves, tune_finish_block;
  // LDS buffers per transform of the self-gravity kernels: 0 the default
  // (one where that lets more workgroups share a CU), 1 or 2
  int sg_fft_buffers;
  // rings per workgroup of the self-gravity row transforms: 0 the default
  // (two where the rows run in place), 1 or 2
  int sg_rows_pair;
  // with the full self-gravity: k_finish also forms the next sub-step's row
  // spectra (one workgroup per ring), and the self-gravity branch of a
  // sub-step starts with the radial transforms
  int sg_rows_in_finish;
  int num_cu; // compute units of the device
  // constants
  double pm, dphi, invdphi, gamma, alpha, visc_const, floor_dens;
  double sg_eps, sg_rstep, sg_tstep, vh_lr1, vh_lr2;
  double mass_taper_param, thickness_smoothing, roche_smoothing, indirect_smoothing;
  double exclude_hill_factor, release_date, release_radius, aspect_ratio, flaring;
  double transition_radius, transition_ratio, transition_width;
  // 1-D tables (FP64, device)
  const double *Rinf, *Rsup, *Rmed, *Surf, *InvRmed, *InvSurf, *InvDiffRmed, *InvDiffRsup;
  const double *InvRinf, *CosAz, *SinAz, *SigmaMed, *EnergyMed, *VthetaMed, *cs_iso;
  const double *powmin2, *r15, *visc_mult, *beta, *supp_torque, *sg_sqrt, *sg_normr, *sg_normt;
  const double *open_vt_in1, *open_vt_in2, *open_vt_out;
  const double *InvDxTheta; // 1 / (dphi Rmed[i]) per ring
  // fields (T)
  void *rho, *vr, *vt, *en, *cs, *vri, *vti, *trr, *trp, *tpp, *divv, *vh;
  void *t_rho, *t_rp, *t_rm, *t_tp, *t_tm, *t_en, *sgr, *sgt, *pot, *e_int;
  // the transported quantities after the azimuthal pass (the segments of a
  // ring read t_* around their cells while others store: separate arrays)
  void *z_rho, *z_rp, *z_rm, *z_tp, *z_tm, *z_en;
  // ring buffers (FP64 unless noted)
  double *vmoy, *ring_dt, *ring_part, *nu_ring, *gcs, *vmed, *lost_col;
  // per ring and DF_SRC_COLS-column tile of the source kernel, the sum of
  // vt dt / dt (nr x ceil(ns / DF_SRC_COLS))
  double *vt_part;
  long *nshift;
  dfargo_ctrl *ctrl;
  // dust particles (FP64 whatever the field type) and the grids they share
  // with the gas (T)
  int d_on, d_np, d_interp, d_feel_disk, d_feel_sg, d_feel_planets, d_feedback, d_sfta;
  int d_fixed_dt, log_grid;
  double d_dust_timestep, d_rho_code, d_unit_mass, d_len_m2; // d_len_m2 = unit_length^-2
  double aziinf0, azisup_last, delta_r, delta_phi, rmin, rmax;
  const double *Azimuth;
  DustSet pset[2];
  // per particle position of the current sub-step: drag, stopping time,
  // |dv|^2 for the heating
  double *p_raddrag, *p_azidrag, *p_ts, *p_dv2;
  // per particle the separable stencil weights of its three stencils:
  // p_w[(kind * 6 + q) * np + k], q < 3 radial, 3 + q azimuthal
  double *p_w;
  // buckets: particle cell and slot in it, per-cell counts and starts
  // (ring-local, then global), per-ring totals and offsets (ring_off[nr]:
  // the particles inside the grid); bucket entries are (id << 32 | position)
  int *p_cell, *p_slot, *cell_count, *cell_local, *cell_start, *ring_tot, *ring_off;
  unsigned long long *p_sorted;
  int *out_list; // positions of the particles outside the grid
  // first node of each stencil relative to the particle's cell, 2 + 2 bits
  // per stencil (ring, column; offset + 1)
  int *p_base;
  // minimum stopping time per particle block of the last interpolation
  // (bits of a positive double; all ones for a block without particles in
  // the grid)
  unsigned long long *p_tsmin;
  double d_ring_step; // log(Rinf[1]/Rinf[0]) on a logarithmic grid, Rinf[1]-Rinf[0] otherwise
  void *radgradp, *azigradp, *radindacc, *aziindacc, *radsgacc, *azisgacc;
  void *radfbacc, *azifbacc, *fbedot, *dustdens;
  // energy equation: the sound speed the sub-step started with, kept for
  // the particles when the CFL kernel recomputes cs
  void *cs_old;
  // the dust fluid (DUSTFLUID YES, FP64 storage): switches, grain
  // constants (internal density and radius in code units, the code-to-cgs
  // surface density factor), state, the sub-step's Stokes numbers,
  // DVradInt / DVthetaInt, stresses, the transport's buffers (dt_*, dz_*
  // as t_*, z_* for the gas), ring buffers and ring tables
  int df_on, df_sfta, df_feel_disk, df_feedback, df_feel_sg, df_diffusion, df_visc, df_open;
  double df_rho_int, df_radius, df_to_cgs, df_diff_inn;
  double *drho, *dvr, *dvt, *dstokes, *dvri, *dvti, *dtrr, *dtrp, *dtpp;
  double *dt_rho, *dt_rp, *dt_rm, *dt_tp, *dt_tm, *dz_rho, *dz_rp, *dz_rm, *dz_tp, *dz_tm;
  double *dvmoy, *dvmed, *dvt_part, *dlost_col, *df_axisg, *df_ring_dt;
  long *dnshift;
  const double *df_cs, *df_nu, *df_sup15, *df_cavsup, *DSigmaMed, *DVthetaMed, *DVradMed;
  // the transport kernels run on the dust fluid (Transportd): its lost
  // mass, no FARGO margin; in the short-friction-time approximation the
  // momenta only take the integer shift
  int tr_dust, tr_skip_mom;
  // self-gravity buffers (complex T)
  // sg_kr, sg_kt: kernel spectra for wavenumbers 0..Ns/2, [k2][k1] (transposed)
  void *sg_z, *sg_w, *sg_kr, *sg_kt, *tw_ns, *tw_2n;
};

// columns of a source-kernel tile: one wavefront per ring of the tile
#define DF_SRC_COLS 64

// launchers (step.hip, sg.hip); each enqueues on stream s
enum KernelId {
  KID_SG_ROWS_FWD = 0, KID_SG_COLS, KID_SG_ROWS_INV,
  KID_CFL, KID_CONTROL, KID_SOURCE, KID_ENERGY, KID_TRANSPORT_RAD, KID_TRANSPORT_AZI,
  KID_FINISH, KID_DUST_MOVE, KID_DUST_SCAN, KID_DUST_SCATTER, KID_DUST_SORT, KID_DUST_DEPOSIT, KID_DUST_REORDER,
  KID_EVANESCENT, KID_NONREFLECTING,
  KID_SGZ_MEAN, KID_SGZ_CONV, KID_SGZ_FILL, KID_EOS_STEP, KID_VISC1D,
  KID_DF_CFL, KID_DF_AXISG, KID_DF_SOURCE, KID_DF_STRESS, KID_DF_VISC, KID_DF_SUBSTEP2, KID_DF_TRANSPORT_RAD,
  KID_DF_TRANSPORT_AZI, KID_DF_FINISH, KID_COUNT
};
extern const char *kKernelNames[KID_COUNT];

// One block per ring; with xcd_remap the rings of one XCD (blocks are
// dealt to the 8 XCDs round-robin) are contiguous, so the radial neighbours
// a block reads, and the lines a ring-strided access shares, stay in one L2.
__device__ __forceinline__ int ring_of_block(const KArgs &a) {
  int b = blockIdx.x, n = gridDim.x;
  if (a.xcd_remap && (n % 8 == 0)) return (b % 8) * (n / 8) + b / 8;
  return b;
}

template <typename T> void launch_kernel(int id, const KArgs &a, hipStream_t s);
// one-time kernel attributes (dynamic LDS above 64 KiB); call before any capture
template <typename T> void prepare_kernels(const KArgs &a);
template <typename T> void sg_prepare(const KArgs &a);
template <typename T> void launch_sg_kernel(int id, const KArgs &a, hipStream_t s);
template <typename T> void launch_sg_rows_forced(const KArgs &a, hipStream_t s);
// dust.hip: the particle kernels of a sub-step (after the control kernel),
// and the start-up interpolation outside sub-steps
template <typename T> void launch_dust_kernel(int id, const KArgs &a, hipStream_t s);
template <typename T> void launch_dust_startup(const KArgs &a, double timestep, double omega_frame, hipStream_t s);
template <typename T> void launch_eos_pre(const KArgs &a, hipStream_t s);
template <typename T> void launch_potential(const KArgs &a, hipStream_t s);
template <typename T> void launch_sg_accel(const KArgs &a, hipStream_t s); // rows_fwd..rows_inv
template <typename T> void launch_visc_stress_state(const KArgs &a, hipStream_t s);
void launch_algogas_begin(const KArgs &a, double DT, hipStream_t s);
// per-DT diagnostics: the bodies whose disc force is logged (position,
// distance to the star, smoothing length, Hill radius)
struct DiagBodies {
  int n;
  double x[DF_MAXPL], y[DF_MAXPL], a[DF_MAXPL], rsm[DF_MAXPL], rh[DF_MAXPL];
};
#define DIAG_NPART(n) (3 + 4 * (n))
#define DIAG_NOUT(n) (3 + 8 * (n))
// ring partials into part (DIAG_NPART(DF_MAXPL) x nr doubles), sums into out
// (DIAG_NOUT(b.n) doubles)
template <typename T>
void launch_diagnostics(const KArgs &a, const DiagBodies &b, double *part, double *out, hipStream_t s);
int ring_kcheck(int ns); // threads-per-ring split used by the ring kernels, -1 if none

// FFT utilities (sg.hip)
template <typename T>
void launch_fft_rows(void *data, int nrows, int L, const void *tw, int dir, hipStream_t s);
template <typename T>
void launch_fft_cols(void *data, int ncols, int L, const void *tw, int dir, hipStream_t s);
void launch_convert_d2f(const double *src, float *dst, size_t n, hipStream_t s);
