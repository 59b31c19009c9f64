// Split-step passes, energy evaluation and standalone transforms for grid
// lengths without a compiled power-of-two kernel: any even length up to 2048
// (the reference only requires even mesh sizes,
// spinor_gpe/pspinor/pspinor.py:331-332).
//
// The length is a runtime argument. A workgroup loads its transforms (W
// columns, or both components of one row) into LDS and runs a mixed-radix
// Stockham transform there, ping-ponging between two LDS buffers, one stage
// per prime-power factor (radices 8, 4, 2, 3, 5, 7, then any larger prime
// factors), one barrier per stage; the butterflies of a stage are dealt to
// the workgroup's threads. Twiddles and the small DFT matrices come from one
// table of the length's roots of unity, root[m] = exp(-2 pi i m / n).
// Radices above 7 take O(R) work per output: such lengths are supported,
// not fast. The passes do what col_pass_kernel and
// row_pass_kernel do (same sub-step schedule, same partial sums, same state
// forms), so plans can mix them: a 96 x 128 grid takes the compiled
// 128-point column pass and this row pass.
#pragma once

#include "sgpe_kernels.h"

namespace sgpe {

constexpr int kGenMaxStages = 12;
constexpr int kGenThreads = 256;
constexpr int kGenMaxLen = 2048;

// Stage radices of one length (host-built; kernel argument).
struct GenFft {
  int n;
  int nst;
  int radix[kGenMaxStages];
};

// dst[e] = src(e) for e = threadIdx.x, threadIdx.x + blockDim.x, ... < n,
// four elements per thread read before any is written. Array pointers
// taken from the slot descriptor are generic addresses, which the compiler
// must assume may alias LDS, so a plain copy loop waits for every load
// before its store: a memory round trip per element.
template <typename C, class Src>
__device__ __forceinline__ void gen_copy_in(C* dst, int n, Src src) {
  for (int e0 = threadIdx.x; e0 < n; e0 += 4 * (int)blockDim.x) {
    C v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = e0 + k * (int)blockDim.x;
      if (e < n) v[k] = src(e);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = e0 + k * (int)blockDim.x;
      if (e < n) dst[e] = v[k];
    }
  }
}

// Small DFT of radix R (3, 5 or 7) from the root table: y_q = sum_r b_r w^{rq}
// with w = root[n / R] (conjugated for the inverse).
template <int R, bool INV, typename C>
__device__ __forceinline__ void gen_small_dft(C (&b)[R], const C* __restrict__ root, int n) {
  C w[R];
  w[0] = C{1, 0};
#pragma unroll
  for (int j = 1; j < R; ++j) {
    w[j] = root[j * (n / R)];
    if (INV) w[j].y = -w[j].y;
  }
  C y[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    C acc = b[0];
#pragma unroll
    for (int r = 1; r < R; ++r) acc = cadd(acc, cmul(b[r], w[(r * q) % R]));
    y[q] = acc;
  }
#pragma unroll
  for (int q = 0; q < R; ++q) b[q] = y[q];
}

// One Stockham stage of radix R and span L over ntr transforms in LDS:
// element i of transform m sits at base[m * ms + i * is]; reads src, writes
// dst (the same layout). Butterfly v of transform m takes inputs
// v + r n/R, twiddles them by w_{LR}^{r (v mod L)} and writes outputs to
// (v - k) R + k + q L.
template <int R, bool INV, typename C>
__device__ __forceinline__ void gen_stage_r(const C* src, C* dst, int n, int L, int ntr, int ms, int is,
                                            const C* __restrict__ root) {
  const int nb = n / R;
  const int per = n / (L * R);
  const int total = nb * ntr;
  for (int w = threadIdx.x; w < total; w += blockDim.x) {
    int v, m;
    if (is == 1) {  // row layout: consecutive threads take consecutive butterflies
      m = w / nb;
      v = w - m * nb;
    } else {  // column layout: consecutive threads take neighbouring columns
      v = w / ntr;
      m = w - v * ntr;
    }
    const int k = v % L;
    C b[R];
#pragma unroll
    for (int r = 0; r < R; ++r) b[r] = src[m * ms + (v + r * nb) * is];
    if (k) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        C tw = root[r * k * per];
        if (INV) tw.y = -tw.y;
        b[r] = cmul(b[r], tw);
      }
    }
    if constexpr (R == 2 || R == 4 || R == 8)
      dft<R, INV>(b);
    else
      gen_small_dft<R, INV>(b, root, n);
    const int base = (v - k) * R + k;
#pragma unroll
    for (int q = 0; q < R; ++q) dst[m * ms + (base + q * L) * is] = b[q];
  }
}

// The same stage for a prime radix R > 7 known only at run time: one output
// per thread, y[(v - k) R + k + q L] = sum_r x[v + r n/R] w_n^{r (k n/(L R) + q n/R)}
// (the twiddle and the DFT_R factor of a term fold into one root).
template <bool INV, typename C>
__device__ __forceinline__ void gen_stage_any(const C* src, C* dst, int n, int L, int R, int ntr, int ms, int is,
                                              const C* __restrict__ root) {
  const int nb = n / R;
  const int per = n / (L * R);
  const int total = n * ntr;
  for (int w = threadIdx.x; w < total; w += blockDim.x) {
    int o, m;
    if (is == 1) {
      m = w / n;
      o = w - m * n;
    } else {
      o = w / ntr;
      m = w - o * ntr;
    }
    const int k = o % L;
    const int q = (o / L) % R;
    const int v = (o / (L * R)) * L + k;
    const int e = (k * per + q * nb) % n;
    C acc = src[m * ms + v * is];
    int idx = 0;
    for (int r = 1; r < R; ++r) {
      idx += e;
      if (idx >= n) idx -= n;
      C tw = root[idx];
      if (INV) tw.y = -tw.y;
      acc = cadd(acc, cmul(src[m * ms + (v + r * nb) * is], tw));
    }
    dst[m * ms + o * is] = acc;
  }
}

// Whole transform of ntr transforms in LDS buffer `a` (scratch `b` of the
// same size). Returns the buffer holding the result. Every thread of the
// block calls it; the caller synchronizes before (data in `a`) and the
// result is published by the last stage's barrier.
template <bool INV, typename C>
__device__ __forceinline__ C* gen_fft(C* a, C* b, const GenFft& f, int ntr, int ms, int is,
                                      const C* __restrict__ root) {
  int L = 1;
  for (int s = 0; s < f.nst; ++s) {
    const int R = f.radix[s];
    switch (R) {
      case 2: gen_stage_r<2, INV>(a, b, f.n, L, ntr, ms, is, root); break;
      case 3: gen_stage_r<3, INV>(a, b, f.n, L, ntr, ms, is, root); break;
      case 4: gen_stage_r<4, INV>(a, b, f.n, L, ntr, ms, is, root); break;
      case 5: gen_stage_r<5, INV>(a, b, f.n, L, ntr, ms, is, root); break;
      case 7: gen_stage_r<7, INV>(a, b, f.n, L, ntr, ms, is, root); break;
      case 8: gen_stage_r<8, INV>(a, b, f.n, L, ntr, ms, is, root); break;
      default: gen_stage_any<INV>(a, b, f.n, L, R, ntr, ms, is, root); break;
    }
    __syncthreads();
    C* t = a;
    a = b;
    b = t;
    L *= R;
  }
  return a;
}

// Block-wide sums of two doubles for a whole-tile result (thread 0 gets it).
__device__ __forceinline__ void gen_sums(double (&v)[2], double* lds_tail) {
  block_sum<2>(v, lds_tail);
}

// ------------------------------------------------------------ column pass
// W columns jx0 .. jx0 + W - 1 of component blockIdx.y of slot blockIdx.z:
// the schedule of col_pass_kernel (DFT_y unless the state is in k-space,
// K_prev, S sums, K_cur, NRM sums, IDFT_y), partial sums per tile.
template <typename R>
__global__ __launch_bounds__(kGenThreads) void gen_col_pass_kernel(const InstDev* __restrict__ insts,
                                                                   RunDev* __restrict__ runs, GenFft f,
                                                                   const typename Cx<R>::T* __restrict__ root,
                                                                   int W, int mode, int d_prev, int d_cur) {
  typedef typename Cx<R>::T C;
  extern __shared__ double smem[];
  const int inst = blockIdx.z;
  if (!runs[inst].active) return;
  const InstDev& P = insts[inst];
  const int N = f.n;
  const int nx = P.nx;
  const int comp = blockIdx.y;
  const int tile = blockIdx.x;
  const int jx0 = tile * W;
  const int cnt = N * W;
  C* A = reinterpret_cast<C*>(smem);
  C* B = A + cnt;
  C* psi = reinterpret_cast<C*>(P.psi) + (size_t)comp * N * nx + jx0;
  gen_copy_in(A, cnt, [&](int e) { return psi[(size_t)(e / W) * nx + (e % W)]; });
  const bool kspace = runs[inst].form == FORM_KSPACE;
  __syncthreads();
  C* cur = kspace ? A : gen_fft<false>(A, B, f, W, 1, W, root);
  C* oth = cur == A ? B : A;
  const double* td = P.t_dense[comp];
  const C* kxp = reinterpret_cast<const C*>(P.kx[d_prev][comp]);
  const C* kyp = reinterpret_cast<const C*>(P.ky[d_prev][comp]);
  const C* kxc = reinterpret_cast<const C*>(P.kx[d_cur][comp]);
  const C* kyc = reinterpret_cast<const C*>(P.ky[d_cur][comp]);
  const R hp = (R)(0.5 * P.dt_sub[d_prev]);
  const R hc = (R)(0.5 * P.dt_sub[d_cur]);
  const bool mid = mode == COL_MID;
  double v[2] = {0.0, 0.0};
  // Four elements per thread at a time: their table entries are read before
  // any of them is rewritten (see gen_copy_in); same order of operations.
  for (int e0 = threadIdx.x; e0 < cnt; e0 += 4 * (int)blockDim.x) {
    C kp[4], kc[4];
    R tv[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = e0 + i * (int)blockDim.x;
      if (e >= cnt) continue;
      const int jy = e / W, jx = jx0 + e % W;
      if (td) {
        tv[i] = (R)td[(size_t)jy * nx + jx];
      } else {
        kp[i] = cmul(kxp[jx], kyp[jy]);
        kc[i] = cmul(kxc[jx], kyc[jy]);
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = e0 + i * (int)blockDim.x;
      if (e >= cnt) continue;
      C a = cur[e];
      if (!kspace) {
        C k;
        if (td)
          k = P.imag ? phase_op<R, true>(tv[i] * hp) : phase_op<R, false>(tv[i] * hp);
        else
          k = kp[i];
        a = cmul(a, k);
      }
      v[0] += norm2d<R>(a);
      if (mid) {
        C k;
        if (td)
          k = P.imag ? phase_op<R, true>(tv[i] * hc) : phase_op<R, false>(tv[i] * hc);
        else
          k = kc[i];
        a = cmul(a, k);
        v[1] += norm2d<R>(a);
      }
      cur[e] = a;
    }
  }
  __syncthreads();
  if (mid) cur = gen_fft<true>(cur, oth, f, W, 1, W, root);
  for (int e = threadIdx.x; e < cnt; e += blockDim.x) psi[(size_t)(e / W) * nx + (e % W)] = cur[e];
  gen_sums(v, reinterpret_cast<double*>(B + cnt));
  if (threadIdx.x == 0) {
    const int nt = nx / W;
    P.part[PART_S * 2 * nt + comp * nt + tile] = v[0];
    P.part[PART_NRM * 2 * nt + comp * nt + tile] = v[1];
  }
}

// -------------------------------------------------------------- row pass
// The real-space block of row_pass_kernel for one point at runtime x:
// normalization, U C P C U with U from the normalized densities (same U
// before and after, tensor_propagator.py:244-267), then the scale c of the
// forward transform. Separable or dense potentials, no, uniform or
// array coupling, its e^{-i theta(x)} phase.
template <typename R, bool IMAG>
__device__ __forceinline__ void gen_point(typename Cx<R>::T& p0, typename Cx<R>::T& p1, const InstDev& P, int dur,
                                          int y, int x, int nx, double lam2) {
  typedef typename Cx<R>::T C;
  const R lam = (R)sqrt(lam2);
  const R lam2r = (R)lam2;
  const R hdt = (R)(0.5 * P.dt_sub[dur]);
  const R dts = (R)P.dt_sub[dur];
  const R n0 = lam2r * (p0.x * p0.x + p0.y * p0.y);
  const R n1 = lam2r * (p1.x * p1.x + p1.y * p1.y);
  const C u0 = phase_op<R, IMAG>(((R)P.g_uu * n0 + (R)P.g_ud * n1) * hdt);
  const C u1 = phase_op<R, IMAG>(((R)P.g_dd * n1 + (R)P.g_ud * n0) * hdt);
  p0 = cscale(mul_phase<IMAG>(p0, u0), lam);
  p1 = cscale(mul_phase<IMAG>(p1, u1), lam);
  const bool apply = P.coupling != 0;
  R alpha = (R)P.c_alpha[dur];
  C b01 = C{(R)P.c_beta_re[dur], (R)P.c_beta_im[dur]}, b10 = b01;
  if (P.coupling == 2) {
    const R ang = (R)P.omega_dense[(size_t)y * nx + x] * dts * (R)0.25;
    if constexpr (IMAG) {
      const R e = ex<R>(ang), ei = (R)1 / e;
      alpha = (R)0.5 * (e + ei);
      b01 = C{(R)-0.5 * (e - ei), (R)0};
    } else {
      R s, c;
      sc<R>(ang, &s, &c);
      alpha = c;
      b01 = C{(R)0, -s};
    }
    b10 = b01;
  }
  if (apply) {
    const C e = reinterpret_cast<const C*>(P.etheta)[x];
    b10 = cmul(b01, cconj(e));
    b01 = cmul(b01, e);
    const C q0 = cadd(cscale(p0, alpha), cmul(b01, p1));
    const C q1 = cadd(cmul(b10, p0), cscale(p1, alpha));
    p0 = q0;
    p1 = q1;
  }
  const double* vd0 = P.v_dense[0];
  const double* vd1 = P.v_dense[1];
  const C f0 = vd0 ? phase_op<R, IMAG>((R)vd0[(size_t)y * nx + x] * dts)
                   : cmul(reinterpret_cast<const C*>(P.px[dur][0])[x], reinterpret_cast<const C*>(P.py[dur][0])[y]);
  const C f1 = vd1 ? phase_op<R, IMAG>((R)vd1[(size_t)y * nx + x] * dts)
                   : cmul(reinterpret_cast<const C*>(P.px[dur][1])[x], reinterpret_cast<const C*>(P.py[dur][1])[y]);
  p0 = cmul(p0, f0);
  p1 = cmul(p1, f1);
  if (apply) {
    const C q0 = cadd(cscale(p0, alpha), cmul(b01, p1));
    const C q1 = cadd(cmul(b10, p0), cscale(p1, alpha));
    p0 = q0;
    p1 = q1;
  }
  const R cf = (R)P.fft_c;
  p0 = cscale(mul_phase<IMAG>(p0, u0), cf);
  p1 = cscale(mul_phase<IMAG>(p1, u1), cf);
}

// Row blockIdx.x (both components) of slot blockIdx.z: the schedule of
// row_pass_kernel (populations and step count, real-space norm from the
// column partials by Parseval, IDFT_x, U C P C U, DFT_x).
template <typename R>
__global__ __launch_bounds__(kGenThreads) void gen_row_pass_kernel(const InstDev* __restrict__ insts,
                                                                   RunDev* __restrict__ runs, GenFft f,
                                                                   const typename Cx<R>::T* __restrict__ root,
                                                                   int dur, int first_sub, int col_w) {
  typedef typename Cx<R>::T C;
  extern __shared__ double smem[];
  const int inst = blockIdx.z;
  RunDev& run = runs[inst];
  if (!run.active) return;
  const InstDev& P = insts[inst];
  const int N = f.n;
  const int ny = P.ny;
  const int y = blockIdx.x;
  C* A = reinterpret_cast<C*>(smem);
  C* B = A + 2 * N;
  C* r0 = reinterpret_cast<C*>(P.psi) + (size_t)y * N;
  C* r1 = r0 + (size_t)ny * N;
  gen_copy_in(A, 2 * N, [&](int e) { return e < N ? r0[e] : r1[e - N]; });
  const int nt2 = 2 * (N / col_w);
  if (blockIdx.x == 0) {
    if (first_sub) {
      const int step = run.step;
      record_pops(P, run.pops, run.pops_cap, step, nt2);
      if (threadIdx.x == 0) run.step = step + 1;
    }
    if (threadIdx.x == 0) run.form = FORM_HYBRID;
  }
  const double nrm = wave_sum_strided(P.part + PART_NRM * nt2, nt2);
  const double lam2 = P.atom_num / (P.dv_r * (double)N * (double)ny * nrm);
  __syncthreads();
  C* cur = gen_fft<true>(A, B, f, 2, N, 1, root);
  C* oth = cur == A ? B : A;
  for (int x = threadIdx.x; x < N; x += blockDim.x) {
    C p0 = cur[x], p1 = cur[N + x];
    if (P.imag)
      gen_point<R, true>(p0, p1, P, dur, y, x, N, lam2);
    else
      gen_point<R, false>(p0, p1, P, dur, y, x, N, lam2);
    cur[x] = p0;
    cur[N + x] = p1;
  }
  __syncthreads();
  cur = gen_fft<false>(cur, oth, f, 2, N, 1, root);
  for (int x = threadIdx.x; x < N; x += blockDim.x) {
    r0[x] = cur[x];
    r1[x] = cur[N + x];
  }
}

// ----------------------------------------------------------------- energy
// In-place evaluation as energy_col_kernel / energy_row_kernel: the column
// part (k-space kinetic and norm sums, IDFT_y; or FWD: DFT_y, restoring
// psik), the row part (IDFT_x, real-space sums, DFT_x / (Nx Ny)).
template <typename R, bool FWD>
__global__ __launch_bounds__(kGenThreads) void gen_energy_col_kernel(const InstDev* __restrict__ insts,
                                                                     const RunDev* __restrict__ runs, GenFft f,
                                                                     const typename Cx<R>::T* __restrict__ root,
                                                                     int W) {
  typedef typename Cx<R>::T C;
  extern __shared__ double smem[];
  const int inst = blockIdx.z;
  if (!runs[inst].active) return;
  const InstDev& P = insts[inst];
  const int N = f.n;
  const int nx = P.nx;
  const int comp = blockIdx.y;
  const int tile = blockIdx.x;
  const int jx0 = tile * W;
  const int cnt = N * W;
  C* A = reinterpret_cast<C*>(smem);
  C* B = A + cnt;
  C* psi = reinterpret_cast<C*>(P.psi) + (size_t)comp * N * nx + jx0;
  double v[2] = {0.0, 0.0};
  const double* td = P.t_dense[comp];
  // Four elements per thread at a time, read before any is stored (see
  // gen_copy_in); the sums run in the same order.
  for (int e0 = threadIdx.x; e0 < cnt; e0 += 4 * (int)blockDim.x) {
    C a[4];
    double tv[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = e0 + i * (int)blockDim.x;
      if (e >= cnt) continue;
      a[i] = psi[(size_t)(e / W) * nx + (e % W)];
      if constexpr (!FWD) {
        const int jy = e / W, jx = jx0 + e % W;
        tv[i] = td ? td[(size_t)jy * nx + jx] : P.tx[comp][jx] + P.ty[comp][jy];
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = e0 + i * (int)blockDim.x;
      if (e >= cnt) continue;
      A[e] = a[i];
      if constexpr (!FWD) {
        const double m = norm2d<R>(a[i]);
        v[0] += tv[i] * m;
        v[1] += m;
      }
    }
  }
  __syncthreads();
  C* cur = gen_fft<!FWD>(A, B, f, W, 1, W, root);
  for (int e = threadIdx.x; e < cnt; e += blockDim.x) psi[(size_t)(e / W) * nx + (e % W)] = cur[e];
  if constexpr (!FWD) {
    gen_sums(v, reinterpret_cast<double*>(B + cnt));
    if (threadIdx.x == 0) {
      const int nt = nx / W;
      P.epart[EPART_KIN * P.estride + comp * nt + tile] = v[0];
      P.epart[EPART_S * P.estride + comp * nt + tile] = v[1];
    }
  }
}

template <typename R>
__global__ __launch_bounds__(kGenThreads) void gen_energy_row_kernel(const InstDev* __restrict__ insts,
                                                                     const RunDev* __restrict__ runs, GenFft f,
                                                                     const typename Cx<R>::T* __restrict__ root) {
  typedef typename Cx<R>::T C;
  extern __shared__ double smem[];
  const int inst = blockIdx.z;
  if (!runs[inst].active) return;
  const InstDev& P = insts[inst];
  const int N = f.n;
  const int ny = P.ny;
  const int y = blockIdx.x;
  C* A = reinterpret_cast<C*>(smem);
  C* B = A + 2 * N;
  C* r0 = reinterpret_cast<C*>(P.psi) + (size_t)y * N;
  C* r1 = r0 + (size_t)ny * N;
  gen_copy_in(A, 2 * N, [&](int e) { return e < N ? r0[e] : r1[e - N]; });
  __syncthreads();
  C* cur = gen_fft<true>(A, B, f, 2, N, 1, root);
  C* oth = cur == A ? B : A;
  const double* vd0 = P.v_dense[0];
  const double* vd1 = P.v_dense[1];
  const C* eth = reinterpret_cast<const C*>(P.etheta);
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int x = threadIdx.x; x < N; x += blockDim.x) {
    const C a0 = cur[x], a1 = cur[N + x];
    const double m0 = norm2d<R>(a0), m1 = norm2d<R>(a1);
    const double v0 = vd0 ? vd0[(size_t)y * N + x] : P.vx[0][x] + P.vy[0][y];
    const double v1 = vd1 ? vd1[(size_t)y * N + x] : P.vx[1][x] + P.vy[1][y];
    s[0] += v0 * m0 + v1 * m1;
    s[1] += P.g_uu * m0 * m0 + P.g_dd * m1 * m1 + 2.0 * P.g_ud * m0 * m1;
    s[3] += m0 + m1;
    if (P.coupling) {
      const double om = P.coupling == 2 ? P.omega_dense[(size_t)y * N + x] : P.omega;
      double cr = (double)a0.x * a1.x + (double)a0.y * a1.y;
      const double ci = (double)a0.x * a1.y - (double)a0.y * a1.x;
      if (eth) cr = cr * (double)eth[x].x - ci * (double)eth[x].y;
      s[2] += om * cr;
    }
  }
  __syncthreads();
  cur = gen_fft<false>(cur, oth, f, 2, N, 1, root);
  const R back = (R)(1.0 / ((double)N * (double)ny));
  for (int x = threadIdx.x; x < N; x += blockDim.x) {
    r0[x] = cscale(cur[x], back);
    r1[x] = cscale(cur[N + x], back);
  }
  block_sum<4>(s, reinterpret_cast<double*>(B + 2 * N));
  if (threadIdx.x == 0) {
    P.epart[EPART_POT * P.estride + blockIdx.x] = s[0];
    P.epart[EPART_INT * P.estride + blockIdx.x] = s[1];
    P.epart[EPART_CPL * P.estride + blockIdx.x] = s[2];
    P.epart[EPART_NRM * P.estride + blockIdx.x] = s[3];
  }
}

// ---------------------------------------------------- standalone fft_2d
// Rows (row blockIdx.x) of array blockIdx.z of the pair: optional pre-
// (forward) or post- (inverse) factor (-1)^(x+y) * scale, as fft_rows_kernel.
template <typename R>
__global__ __launch_bounds__(kGenThreads) void gen_fft_rows_kernel(FftPair io, GenFft f,
                                                                   const typename Cx<R>::T* __restrict__ root,
                                                                   int inv, double scale) {
  typedef typename Cx<R>::T C;
  extern __shared__ double smem[];
  const C* in = reinterpret_cast<const C*>(io.in[blockIdx.z]);
  C* out = reinterpret_cast<C*>(io.out[blockIdx.z]);
  const int N = f.n;
  const int y = blockIdx.x;
  C* A = reinterpret_cast<C*>(smem);
  C* B = A + N;
  const R sy = (R)((y & 1) ? -scale : scale);
  gen_copy_in(A, N, [&](int x) {
    const C a = in[(size_t)y * N + x];
    return inv ? a : cscale(a, (x & 1) ? -sy : sy);
  });
  __syncthreads();
  C* cur = inv ? gen_fft<true>(A, B, f, 1, N, 1, root) : gen_fft<false>(A, B, f, 1, N, 1, root);
  for (int x = threadIdx.x; x < N; x += blockDim.x) {
    const C a = cur[x];
    out[(size_t)y * N + x] = inv ? cscale(a, (x & 1) ? -sy : sy) : a;
  }
}

// Columns jx0 .. jx0 + W - 1 of array blockIdx.z of the pair (in place
// allowed).
template <typename R>
__global__ __launch_bounds__(kGenThreads) void gen_fft_cols_kernel(FftPair io, GenFft f,
                                                                   const typename Cx<R>::T* __restrict__ root,
                                                                   int nx, int W, int inv) {
  typedef typename Cx<R>::T C;
  extern __shared__ double smem[];
  const C* in = reinterpret_cast<const C*>(io.in[blockIdx.z]);
  C* out = reinterpret_cast<C*>(io.out[blockIdx.z]);
  const int N = f.n;
  const int jx0 = blockIdx.x * W;
  const int cnt = N * W;
  C* A = reinterpret_cast<C*>(smem);
  C* B = A + cnt;
  gen_copy_in(A, cnt, [&](int e) { return in[(size_t)(e / W) * nx + jx0 + (e % W)]; });
  __syncthreads();
  C* cur = inv ? gen_fft<true>(A, B, f, W, 1, W, root) : gen_fft<false>(A, B, f, W, 1, W, root);
  for (int e = threadIdx.x; e < cnt; e += blockDim.x) out[(size_t)(e / W) * nx + jx0 + (e % W)] = cur[e];
}

}  // namespace sgpe
