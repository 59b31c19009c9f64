// Shared device/host types of the split-step engine.
//
// One *instance* is one two-component spinor on an (ny, nx) grid, stored as
// a contiguous [2][ny][nx] complex array (interleaved re/im, component 0
// first), in the k-space layout of the reference's `psik`
// (spinor_gpe/pspinor/tensor_tools.py:201-228: fft-shifted, x along axis 1).
// A *plan* groups instances of one grid shape and precision; every kernel is
// launched once per plan with blockIdx.z selecting the instance (slot).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgpe {

template <typename R> struct Cx;
template <> struct Cx<double> { typedef double2 T; };
template <> struct Cx<float> { typedef float2 T; };

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename C, typename R> __device__ __forceinline__ C cscale(C a, R s) { return C{a.x * s, a.y * s}; }
template <typename C> __device__ __forceinline__ C cconj(C a) { return C{a.x, -a.y}; }

// Loads and stores of one complex value through the global address space
// (global_load / global_store, counted by vmcnt only). The passes' state and
// table pointers come from slot descriptors (or a select with one), which the
// compiler would access with flat instructions, and flat accesses count
// against lgkmcnt as well: every later wait for a scalar load or for LDS
// traffic then also waited for the state loads to arrive (or for the
// stores to complete).
template <typename C> __device__ __forceinline__ C ld_global(const C* p) {
  typedef decltype(p->x) S;
  typedef S V2 __attribute__((ext_vector_type(2)));
  const V2 w = *(const __attribute__((address_space(1))) V2*)p;
  return C{w.x, w.y};
}
__device__ __forceinline__ double ld_global(const double* p) { return *(const __attribute__((address_space(1))) double*)p; }
template <typename C> __device__ __forceinline__ void st_global(C* p, C v) {
  typedef decltype(v.x) S;
  typedef S V2 __attribute__((ext_vector_type(2)));
  *(__attribute__((address_space(1))) V2*)p = V2{v.x, v.y};
}

// Inputs and outputs of the standalone 2D transforms: up to two arrays per
// launch (the two spinor components), passed by value as a kernel argument.
struct FftPair {
  const void* in[2];
  void* out[2];
};

// Sub-step durations of the 3-stage composition (tensor_propagator.py:101-103):
// index 0 = outer (gamma * dt), 1 = inner ((1 - 2 gamma) * dt).
enum { DUR_OUT = 0, DUR_IN = 1 };

// Form of the stored state of an instance.
enum {
  FORM_KSPACE = 0,  // psik, full k-space (possibly not normalized)
  FORM_HYBRID = 1,  // after a row pass: x in k-space, y in real space
};

// Column-pass modes.
enum {
  COL_MID = 0,    // DFT_y, K_prev, [S], K_cur, [NRM], IDFT_y (or entry if FORM_KSPACE)
  COL_EXIT = 1,   // DFT_y, K_prev, [S]          -> k-space (not normalized)
};

// Per-instance parameters, read-only for kernels (device memory).
// Complex tables are in the plan precision; f64 arrays stay f64.
struct InstDev {
  void* psi;        // [2][ny][nx] complex state
  double* part;     // column-pass partial sums, [PART_SLOTS][ntile]
  double* epart;    // energy partial sums, [EPART_SLOTS][estride]
  int estride;      // entries per energy partial slot
  int ny, nx;
  int imag;         // 1: imaginary time (t_step = -i dt), 0: real time
  int coupling;     // 0: none, 1: uniform Omega, 2: dense Omega array
  double atom_num, g_uu, g_dd, g_ud;
  double dv_r, dv_k, fft_c;      // dx*dy, dkx*dky, dx*dy/(2 pi)
  double dt_sub[2];              // |sub-step| for DUR_OUT / DUR_IN
  // Kinetic phase K_s = exp(-i T_s Delta/2), separable: kx[d][s][jx] * ky[d][s][jy]
  const void* kx[2][2];
  const void* ky[2][2];
  const double* t_dense[2];      // non-null: dense T_s (f64, ny*nx), evaluated per pixel
  // Potential phase P_s = exp(-i V_s Delta), separable: px[d][s][jx] * py[d][s][jy]
  const void* px[2][2];
  const void* py[2][2];
  const double* v_dense[2];      // non-null: dense V_s (f64)
  // Coupling C = [[alpha, beta e^{-i theta}], [beta e^{+i theta}, alpha]]
  double omega;                  // uniform Omega (coupling == 1)
  double c_alpha[2];             // uniform Omega: per duration
  double c_beta_re[2], c_beta_im[2];
  const double* omega_dense;     // coupling == 2: dense Omega (f64)
  const void* etheta;            // e^{-i theta(x)} table (nx) or null for theta = 0
  // Energy inputs (f64): T_s separable (tx, ty) or t_dense; V_s likewise
  const double* tx[2];
  const double* ty[2];
  const double* vx[2];
  const double* vy[2];
  const double* theta_x;         // theta(x) (nx) or null
};

// Per-instance mutable run state (device memory), written by kernels.
struct RunDev {
  double* pops;      // pops[2*i + s] for full step i of the current run, or null
  int pops_cap;      // capacity in full steps
  int step;          // full steps started in the current run
  int active;        // 0: kernels skip this instance
  int form;          // FORM_*
  int converged;     // set by the convergence kernel
  int n_checks;      // energy evaluations done in this run
  double conv_tol;   // |dE/E| per full step; <= 0 disables
  double e_prev;     // previous total energy (convergence test)
  int steps_prev;    // step count at previous energy evaluation
  int pad_;
  double energy[8];  // E_tot, E_kin, E_pot, E_int, E_cpl, mu, N, rel change
};

enum { PART_S = 0, PART_NRM = 1, PART_SLOTS = 2 };
// Energy partial sums: column-pass slots (KIN, S) per column tile, row-pass
// slots (POT, INT, CPL, NRM) per row block; each slot has estride entries.
enum { EPART_KIN = 0, EPART_S = 1, EPART_POT = 2, EPART_INT = 3, EPART_CPL = 4, EPART_NRM = 5, EPART_SLOTS = 6 };

}  // namespace sgpe
