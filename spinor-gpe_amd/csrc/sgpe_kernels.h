// Split-step kernels. One split step (the reference's full_step,
// spinor_gpe/pspinor/tensor_propagator.py:214-222) is three sub-steps; each
// sub-step here is two sweeps over the state:
//
//   column pass  (per column jx, per component; y is the strided axis)
//       DFT_y -> K_prev (end of the previous sub-step) -> [S: sum |.|^2]
//             -> K_cur (start of this sub-step) -> [NRM: sum |.|^2] -> IDFT_y
//   row pass     (per row y, both components together)
//       IDFT_x -> normalize (from NRM) -> U C P C U -> DFT_x * dx dy/(2 pi)
//
// which is tensor_propagator.py:242-271 with the fftshift/ifftshift pairs and
// the FFT scale factors folded away: for even grid sizes the shifts only
// multiply real space by (-1)^(x+y), which commutes with every real-space
// factor, and the scale is absorbed by the renormalization that follows each
// inverse transform. The k-space renormalization at the end of a sub-step
// (tensor_propagator.py:271) is deferred: the next real-space renormalization
// absorbs it, the per-step populations are ratios of the S sums, and the
// explicit normalize kernel applies it when the state is handed back.
#pragma once

#include "sgpe_fft.h"

namespace sgpe {

// ---------------------------------------------------------------- helpers

// sin and cos by Cody-Waite reduction x = k pi/2 + r, |r| <= pi/4, with
// pi/2 = P1 + P2 + P3 (P1 pi/2 rounded to double, P2 and P3 the successive
// remainders): each fma(-k, Pi, r) rounds once, so r is within a few 1e-16
// of x - k pi/2 for |k| < 1e15 (|x| < 1.5e15; beyond, a double no longer
// resolves a radian of phase and the error grows with |x|). The quadrant is
// k mod 4, computed exactly in double. Then Taylor polynomials in r (terms
// through r^17 / r^18 for double, r^9 / r^10 for float: truncation below half
// an ulp); within 2 ulp of the correctly rounded values. There is no
// library fallback: inlined at every call site, the device library's
// large-argument code made the row pass's pointwise block twice as slow even
// where it never ran (instruction fetch across it).
__device__ __forceinline__ double cw_reduce(double x, int* quadrant) {
  const double k = rint(x * 0.6366197723675814);
  double r = fma(-k, 1.5707963267948966, x);
  r = fma(-k, 6.123233995736766e-17, r);
  r = fma(-k, -1.4973849048591698e-33, r);
  *quadrant = (int)fma(-4.0, floor(k * 0.25), k);
  return r;
}

template <typename R> __device__ __forceinline__ void quadrant_fix(int q, R sr, R cr, R* s, R* c) {
  const R ss = (q & 1) ? cr : sr;
  const R cc = (q & 1) ? sr : cr;
  *s = (q & 2) ? -ss : ss;
  *c = ((q + 1) & 2) ? -cc : cc;
}

template <typename R> __device__ __forceinline__ void sc(R x, R* s, R* c);
template <> __device__ __forceinline__ void sc<double>(double x, double* s, double* c) {
  int q;
  const double r = cw_reduce(x, &q);
  const double r2 = r * r;
  double ps = 2.8114572543455206e-15;
  ps = fma(ps, r2, -7.647163731819816e-13);
  ps = fma(ps, r2, 1.6059043836821613e-10);
  ps = fma(ps, r2, -2.505210838544172e-08);
  ps = fma(ps, r2, 2.7557319223985893e-06);
  ps = fma(ps, r2, -0.0001984126984126984);
  ps = fma(ps, r2, 0.008333333333333333);
  ps = fma(ps, r2, -0.16666666666666666);
  const double sr = fma(r * r2, ps, r);
  double pc = -1.5619206968586225e-16;
  pc = fma(pc, r2, 4.779477332387385e-14);
  pc = fma(pc, r2, -1.1470745597729725e-11);
  pc = fma(pc, r2, 2.08767569878681e-09);
  pc = fma(pc, r2, -2.755731922398589e-07);
  pc = fma(pc, r2, 2.48015873015873e-05);
  pc = fma(pc, r2, -0.001388888888888889);
  pc = fma(pc, r2, 0.041666666666666664);
  pc = fma(pc, r2, -0.5);
  const double cr = fma(r2, pc, 1.0);
  quadrant_fix<double>(q, sr, cr, s, c);
}
template <> __device__ __forceinline__ void sc<float>(float x, float* s, float* c) {
  // The reduction runs in double for every argument (exact for any float
  // x): a single-precision reduction for small arguments with the double one
  // as a fallback put two branches around every sin/cos, and the blocks they
  // cut the pointwise code into kept its independent sin/cos chains from
  // overlapping (the complex64 real-time row pass took 0.3 us longer).
  int q;
  const float r = (float)cw_reduce((double)x, &q);
  const float r2 = r * r;
  float ps = 2.7557319223985893e-06f;
  ps = fmaf(ps, r2, -0.0001984126984126984f);
  ps = fmaf(ps, r2, 0.008333333333333333f);
  ps = fmaf(ps, r2, -0.16666666666666666f);
  const float sr = fmaf(r * r2, ps, r);
  float pc = -2.755731922398589e-07f;
  pc = fmaf(pc, r2, 2.48015873015873e-05f);
  pc = fmaf(pc, r2, -0.001388888888888889f);
  pc = fmaf(pc, r2, 0.041666666666666664f);
  pc = fmaf(pc, r2, -0.5f);
  const float cr = fmaf(r2, pc, 1.0f);
  quadrant_fix<float>(q, sr, cr, s, c);
}
template <typename R> __device__ __forceinline__ R ex(R x);
template <> __device__ __forceinline__ double ex<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ float ex<float>(float x) { return expf(x); }

// exp(-i E Delta) for Delta = dt (real time) or -i dt (imaginary time).
template <typename R, bool IMAG>
__device__ __forceinline__ typename Cx<R>::T phase_op(R e_dt) {
  typedef typename Cx<R>::T C;
  if constexpr (IMAG) {
    return C{ex<R>(-e_dt), (R)0};
  } else {
    R s, c;
    sc<R>(-e_dt, &s, &c);
    return C{c, s};
  }
}

// p times a phase factor from phase_op: in imaginary time the factor is
// real (zero imaginary part), and scaling by its real part gives the same
// result as the complex product without its multiplications by zero.
template <bool IMAG, typename C> __device__ __forceinline__ C mul_phase(C p, C u) {
  if constexpr (IMAG)
    return cscale(p, u.x);
  else
    return cmul(p, u);
}

// The pass kernels select their pointwise code by uniform flags (time mode,
// dense or separable operators) OUTSIDE the loops, in separately compiled
// variants: inside one loop the compiler would evaluate the sin/cos or exp
// of every alternative for every point and select afterwards.

// The same with a non-temporal hint: every pass reads and writes each state
// element exactly once, and streaming them past the caches shortens a pass
// when a single instance leaves the chip mostly idle (for batched plans the
// normal policy can be faster; the plan chooses, see the streaming flags of
// sgpe_plan_set_streaming).
template <typename C> __device__ __forceinline__ C ld_stream(const C* p) {
  typedef decltype(p->x) S;
  typedef S V2 __attribute__((ext_vector_type(2)));
  const V2 w = __builtin_nontemporal_load((const __attribute__((address_space(1))) V2*)p);
  return C{w.x, w.y};
}
template <typename C> __device__ __forceinline__ void st_stream(C* p, C v) {
  typedef decltype(v.x) S;
  typedef S V2 __attribute__((ext_vector_type(2)));
  const V2 w = {v.x, v.y};
  __builtin_nontemporal_store(w, (__attribute__((address_space(1))) V2*)p);
}

// Device-coherent accesses (cache policy SC1: device scope) to data that
// other workgroups of the same launch write or read, through a buffer
// resource over [base, base + bytes): writes reach the device's coherence
// point before the wave's vmcnt drains, and reads do not return lines an
// L2 of another XCD may hold stale, so workgroups of one launch exchange
// data without cache-wide writebacks or invalidations. Out-of-range offsets
// read 0 and drop writes (the resource's bound). 0x00020000 is the
// resource's third word for gfx9 targets (raw buffer, 32-bit data format).
constexpr int kCpolSc1 = 16;
struct CoherentBuf {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ CoherentBuf(void* base, uint32_t bytes)
      : r(__builtin_amdgcn_make_buffer_rsrc(base, 0, (int)bytes, 0x00020000)) {}
  template <typename C> __device__ __forceinline__ C load(uint32_t off) const {
    if constexpr (sizeof(C) == 16) {
      typedef unsigned int V4 __attribute__((ext_vector_type(4)));
      const V4 w = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, kCpolSc1);
      return __builtin_bit_cast(C, w);
    } else {
      typedef unsigned int V2 __attribute__((ext_vector_type(2)));
      const V2 w = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, kCpolSc1);
      return __builtin_bit_cast(C, w);
    }
  }
  template <typename C> __device__ __forceinline__ void store(uint32_t off, C v) const {
    if constexpr (sizeof(C) == 16) {
      typedef unsigned int V4 __attribute__((ext_vector_type(4)));
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V4, v), r, (int)off, 0, kCpolSc1);
    } else {
      typedef unsigned int V2 __attribute__((ext_vector_type(2)));
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, (int)off, 0, kCpolSc1);
    }
  }
};
// Access modes of the pass bodies (CM): device-coherent state and partial
// sum reads (CM_LOAD) and writes (CM_STORE); 0 = ordinary accesses.
enum { CM_LOAD = 1, CM_STORE = 2 };
__device__ __forceinline__ double ld_coherent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coherent(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of n doubles by one wave: lane l adds entries l, l + 64, ..., then a
// butterfly across the lanes; every lane gets the same total, and every wave
// that calls it sums in the same order. COH: device-coherent reads.
// Entries are read four per lane at a time, so that their loads are in flight
// together (one load per iteration made every entry a dependent memory round
// trip), and added in the same order as one at a time.
template <bool COH = false>
__device__ __forceinline__ double wave_sum_strided(const double* __restrict__ p, int n) {
  const int lane = threadIdx.x & 63;
  double v = 0.0;
  for (int i = lane; i < n; i += 4 * 64) {
    double e[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int ik = i + 64 * k;
      e[k] = ik < n ? (COH ? ld_coherent(p + ik) : ld_global(p + ik)) : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (i + 64 * k < n) v += e[k];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <typename R> __device__ __forceinline__ double norm2d(typename Cx<R>::T a) {
  double x = (double)a.x, y = (double)a.y;
  return x * x + y * y;
}

// ------------------------------------------------------------ column pass

// Workgroups are dealt to the 8 XCDs (each with its own L2) round-robin in
// dispatch order. With n tiles a multiple of 8, XCD k gets the contiguous
// tiles [k n/8, (k+1) n/8), so neighbouring column tiles, which share cache
// lines when a tile row is narrower than a line (complex64, 8 columns),
// meet in one L2 instead of being fetched twice.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int b, int n) {
  return n % kXcds == 0 ? (b % kXcds) * (n / kXcds) + b / kXcds : b;
}

// Kinetic factors of the column pass on this thread's E points of column jx:
// the closing K of the previous sub-step (unless the state is already in
// k-space) with the S sum after it, then the opening K of this sub-step with
// the NRM sum after it (mid). DENSE: K from the dense T array, else from the
// separable tables. The two block-uniform choices (kspace, mid) are taken
// once, around whole point loops: tested per point, they put every table
// read in a block of its own with its own wait, a chain of 2E dependent LDS
// round trips (≈1 µs of a 512-point column pass). The sums run over the
// points in the same order either way. DENSE reads T in each loop (the second
// read hits the cache) rather than keeping E values live across both.
// KYG: the y tables are in global memory (else staged in LDS).
template <typename R, int E, int T, bool DENSE, bool IMAG, bool KYG, typename C>
__device__ __forceinline__ void col_kinetic(C (&a)[E], int t, int nx, int jx, const double* __restrict__ td, R hp,
                                            R hc, C kxp, C kxc, const C* kyp, const C* kyc, bool kspace, bool mid,
                                            double& s_acc, double& n_acc) {
  if (!kspace) {
#pragma unroll
    for (int j = 0; j < E; ++j) {
      C k;
      if constexpr (DENSE)
        k = phase_op<R, IMAG>((R)ld_global(td + (size_t)(t + j * T) * nx + jx) * hp);
      else
        k = cmul(kxp, KYG ? ld_global(kyp + t + j * T) : kyp[t + j * T]);
      a[j] = cmul(a[j], k);
    }
  }
#pragma unroll
  for (int j = 0; j < E; ++j) s_acc += norm2d<R>(a[j]);
  if (mid) {
#pragma unroll
    for (int j = 0; j < E; ++j) {
      C k;
      if constexpr (DENSE)
        k = phase_op<R, IMAG>((R)ld_global(td + (size_t)(t + j * T) * nx + jx) * hc);
      else
        k = cmul(kxc, KYG ? ld_global(kyc + t + j * T) : kyc[t + j * T]);
      a[j] = cmul(a[j], k);
      n_acc += norm2d<R>(a[j]);
    }
  }
}

// One column tile of one component (the body of col_pass_kernel and of the
// persistent step kernel). `gate` is called once every global load of the
// prologue has been issued: it returns -1 to drop the tile (inactive slot),
// else whether the stored state is already in k-space (1) or hybrid (0).
// DENSE: the plan has some slot with a dense kinetic array (the code then
// also handles table slots); otherwise every slot uses the separable tables.
template <typename R, int N, int E, int W, bool DENSE, int CM, class Gate, bool GKY = false>
__device__ __forceinline__ void col_item(const InstDev& P, typename Cx<R>::T* __restrict__ psi, int nx,
                                         const typename Cx<R>::T* __restrict__ tws, int mode, int d_prev,
                                         int d_cur, int stream, int comp, int tile, double* smem, Gate gate) {
  typedef typename Cx<R>::T C;
  typedef LdsView<W, 1 << 30, 0> V;
  constexpr int T = N / E;
  // LDS: [W*N exchange][N twiddles][N ky_prev][N ky_cur][sums] (tables and
  // a separate sum slot when they fit). GKY: the two y phase tables stay in
  // global memory (cache resident) even where they would fit, which leaves
  // LDS for more workgroups per CU in batched plans.
  constexpr size_t kData = (size_t)W * N * sizeof(C);
  constexpr bool kTw = tw_in_lds(kData, N, sizeof(C));
  constexpr bool kKy = !GKY && kTw && kData + 3 * (size_t)N * sizeof(C) <= kLdsBytes;
  constexpr size_t kBase = kData + (kTw ? (size_t)N * sizeof(C) : 0) + (kKy ? 2 * (size_t)N * sizeof(C) : 0);
  constexpr bool kSumSep = col_sum_separate(kBase);
  // The W*T threads of a tile interleave its W columns, so a column's
  // transform lies within one wave when the whole tile does.
  constexpr bool kWave = W * T <= 64;
  const int c = threadIdx.x % W;
  const int t = threadIdx.x / W;
  const int jx = tile * W + c;
  C* const sm = reinterpret_cast<C*>(smem);
  C* lds = sm + c;
  C* col = psi + (size_t)comp * N * nx + jx;

  // Every global load is issued before any computation: the state first,
  // then the kinetic tables and twiddles (staged into LDS).
  constexpr int kBlock = W * T;
  C a[E];
  // CM & CM_LOAD / CM_STORE: device-coherent state reads / writes; element
  // (comp, jy, jx) at byte ((comp N + jy) nx + jx) sizeof(C)
  const CoherentBuf cb(psi, CM ? (uint32_t)(2 * (size_t)N * nx * sizeof(C)) : 0u);
  const uint32_t cbase = (uint32_t)(((size_t)comp * N * nx + jx) * sizeof(C));
  if constexpr ((CM & CM_LOAD) != 0) {
#pragma unroll
    for (int j = 0; j < E; ++j) a[j] = cb.load<C>(cbase + (uint32_t)((t + j * T) * nx * sizeof(C)));
  } else if (stream) {
#pragma unroll
    for (int j = 0; j < E; ++j) a[j] = ld_stream(col + (size_t)(t + j * T) * nx);
  } else {
#pragma unroll
    for (int j = 0; j < E; ++j) a[j] = ld_global(col + (size_t)(t + j * T) * nx);
  }
  const double* td = P.t_dense[comp];
  C kxp = C{(R)1, (R)0}, kxc = C{(R)1, (R)0};
  TableStage<kKy, N, kBlock, C> sp, sc_;
  TableStage<kTw, N, kBlock, C> stw;
  if (!td) {
    kxp = ld_global(reinterpret_cast<const C*>(P.kx[d_prev][comp]) + jx);
    kxc = ld_global(reinterpret_cast<const C*>(P.kx[d_cur][comp]) + jx);
  }
  // The y-tables are staged unconditionally (a conditional fill puts the
  // staging registers in scratch); without them (dense kinetic operator) the
  // twiddle table, also N entries, stands in and the copies go unused.
  sp.load(td ? tws : reinterpret_cast<const C*>(P.ky[d_prev][comp]));
  sc_.load(td ? tws : reinterpret_cast<const C*>(P.ky[d_cur][comp]));
  stw.load(tws);
  const int gk = gate();
  if (gk < 0) return;
  const bool kspace = gk != 0;
  const C* kyp = sp.store(sm + W * N + N);
  const C* kyc = sc_.store(sm + W * N + 2 * N);
  const C* tw = stw.store(sm + W * N);

  if (!kspace) {
    if constexpr (kWave && (kTw || kKy)) __syncthreads();  // publishes the staged tables
    fft_run<N, E, false, V, kWave>(a, t, tw, lds);
  } else {
    __syncthreads();  // publishes the staged tables
  }

  double s_acc = 0.0, n_acc = 0.0;
  const R hp = (R)(0.5 * P.dt_sub[d_prev]);
  const R hc = (R)(0.5 * P.dt_sub[d_cur]);
  const bool mid = mode == COL_MID;
  if (!DENSE || !td)
    col_kinetic<R, E, T, false, false, !kKy>(a, t, nx, jx, nullptr, hp, hc, kxp, kxc, kyp, kyc, kspace, mid, s_acc,
                                           n_acc);
  else if constexpr (DENSE) {
    if (P.imag)
      col_kinetic<R, E, T, true, true, !kKy>(a, t, nx, jx, td, hp, hc, kxp, kxc, kyp, kyc, kspace, mid, s_acc, n_acc);
    else
      col_kinetic<R, E, T, true, false, !kKy>(a, t, nx, jx, td, hp, hc, kxp, kxc, kyp, kyc, kspace, mid, s_acc,
                                              n_acc);
  }
  // Columns of 1024 points and more: an opaque copy of t for the inverse
  // transform and the stores, so their addresses are recomputed rather than
  // held since the loads (as in row_item; 2-5 % per pass there, while
  // shorter batched columns lose 1-2 % to the recomputation).
  int tf = t;
  if constexpr (N >= 1024) asm volatile("" : "+v"(tf));
  if (mode == COL_MID) fft_run<N, E, true, V, kWave>(a, tf, tw, lds);
  if constexpr ((CM & CM_STORE) != 0) {
#pragma unroll
    for (int j = 0; j < E; ++j) cb.store<C>(cbase + (uint32_t)((tf + j * T) * nx * sizeof(C)), a[j]);
  } else if (stream) {
#pragma unroll
    for (int j = 0; j < E; ++j) st_stream(col + (size_t)(tf + j * T) * nx, a[j]);
  } else {
#pragma unroll
    for (int j = 0; j < E; ++j) st_global(col + (size_t)(tf + j * T) * nx, a[j]);
  }

  double v[2] = {s_acc, n_acc};
  if constexpr (kSumSep)
    block_sum_thread0<2>(v, reinterpret_cast<double*>(reinterpret_cast<char*>(smem) + kBase));
  else
    block_sum<2>(v, smem);
  if (threadIdx.x == 0) {
    const int nt = nx / W;
    double* ps = P.part + PART_S * 2 * nt + comp * nt + tile;
    double* pn = P.part + PART_NRM * 2 * nt + comp * nt + tile;
    if constexpr ((CM & CM_STORE) != 0) {
      st_coherent(ps, v[0]);
      st_coherent(pn, v[1]);
    } else {
      *ps = v[0];
      *pn = v[1];
    }
  }
}

template <typename R, int N, int E, int W, bool DENSE, bool GKY = false>
__global__ __launch_bounds__(W*(N / E)) void col_pass_kernel(const InstDev* __restrict__ insts,
                                                             RunDev* __restrict__ runs, void* direct_psi,
                                                             int direct_n, const typename Cx<R>::T* __restrict__ tws,
                                                             int mode, int d_prev, int d_cur, int stream) {
  typedef typename Cx<R>::T C;
  extern __shared__ double smem[];
  const int inst = blockIdx.z;
  const InstDev& P = insts[inst];
  // A single-slot plan passes its state's address and row length as kernel
  // arguments (direct_psi): the state loads are issued at once and the
  // active flag is checked after them. Otherwise the address and length are
  // read together with the active flag, so the state loads wait for one
  // round of scalar loads, not two.
  const bool direct = direct_psi != nullptr;
  C* const psi = reinterpret_cast<C*>(direct ? direct_psi : P.psi);
  const int nx = direct ? direct_n : P.nx;
  if (!direct && (!runs[inst].active || !psi || nx <= 0)) return;
  auto gate = [&]() -> int {
    if (direct) {
      // The prologue loads stay ahead of the flag test (an inactive slot's
      // state and tables are still allocated: the host passes direct_psi
      // only for a loaded slot).
      asm volatile("" ::: "memory");
      if (!runs[inst].active) return -1;
    }
    return runs[inst].form == FORM_KSPACE ? 1 : 0;
  };
  col_item<R, N, E, W, DENSE, 0, decltype(gate), GKY>(P, psi, nx, tws, mode, d_prev, d_cur, stream, blockIdx.y,
                                     xcd_tile(blockIdx.x, nx / W), smem, gate);
}

// Sum of one partial slot over all column tiles of both components
// (deterministic: same order in every block). All threads get the result.
__device__ __forceinline__ void part_sums(const double* part, int nt2, int slot, int split, double* out2,
                                          double* scratch) {
  double v[2] = {0.0, 0.0};
  const double* p = part + (size_t)slot * nt2;
  for (int i = threadIdx.x; i < nt2; i += blockDim.x) v[(split && i >= nt2 / 2) ? 1 : 0] += p[i];
  block_sum<2>(v, scratch);
  out2[0] = v[0];
  out2[1] = v[1];
}

// Real-space block of the row pass on this thread's points J0 <= j < J1 of row y, both
// components: U C P C U with the normalization folded in (tensor_propagator.
// py:244-267). FAST: separable potential (x and y tables) and no or uniform
// coupling, which is then always applied (the identity when off) with its
// e^{-i theta} table (ones in the rotated frame); otherwise dense potentials
// and coupling arrays where present.
template <typename R, int N, int E, int T, bool IMAG, bool FAST, int J0 = 0, int J1 = E, typename C>
__device__ __forceinline__ void row_pointwise(C (&a0)[E], C (&a1)[E], const InstDev& P, int dur, int y, int t,
                                              double lam2, C py0, C py1, const C (&tx0)[E], const C (&tx1)[E],
                                              const C (&tex)[E]) {
  const R lam = (R)sqrt(lam2);
  const R lam2r = (R)lam2;
  const R cf = (R)P.fft_c;
  const R guu = (R)P.g_uu, gdd = (R)P.g_dd, gud = (R)P.g_ud;
  const R hdt = (R)(0.5 * P.dt_sub[dur]);
  const R dts = (R)P.dt_sub[dur];
  const int cpl = P.coupling;
  const R ca = (R)P.c_alpha[dur];
  const C cb = C{(R)P.c_beta_re[dur], (R)P.c_beta_im[dur]};
  const double* omd = P.omega_dense;
  const double* vd0 = P.v_dense[0];
  const double* vd1 = P.v_dense[1];
#pragma unroll
  for (int j = J0; j < J1; ++j) {
    const int x = t + j * T;
    C p0 = a0[j], p1 = a1[j];
    // Interaction half step U_s = exp(-i (g n) Delta/2) from the normalized
    // densities (tensor_propagator.py:247-250); U is reused after C P C (:267).
    const R n0 = lam2r * (p0.x * p0.x + p0.y * p0.y);
    const R n1 = lam2r * (p1.x * p1.x + p1.y * p1.y);
    const C u0 = phase_op<R, IMAG>((guu * n0 + gud * n1) * hdt);
    const C u1 = phase_op<R, IMAG>((gdd * n1 + gud * n0) * hdt);
    p0 = cscale(mul_phase<IMAG>(p0, u0), lam);
    p1 = cscale(mul_phase<IMAG>(p1, u1), lam);
    R alpha = ca;
    C b01 = cb, b10 = cb;
    bool apply = true;
    if constexpr (FAST) {
      const C e = tex[j];
      b10 = cmul(b01, cconj(e));
      b01 = cmul(b01, e);
    } else {
      apply = cpl != 0;
      if (cpl == 2) {
        const R ang = (R)omd[(size_t)y * N + x] * dts * (R)0.25;
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
      if (cpl) {
        const C e = tex[j];
        b10 = cmul(b01, cconj(e));
        b01 = cmul(b01, e);
      }
    }
    // Coupling half step C (tensor_propagator.py:252-254).
    if (apply) {
      const C q0 = cadd(cscale(p0, alpha), cmul(b01, p1));
      const C q1 = cadd(cmul(b10, p0), cscale(p1, alpha));
      p0 = q0;
      p1 = q1;
    }
    // Potential step P_s = exp(-i V_s Delta) (tensor_propagator.py:256).
    C q0, q1;
    if constexpr (FAST) {
      q0 = cmul(tx0[j], py0);
      q1 = cmul(tx1[j], py1);
    } else {
      q0 = vd0 ? phase_op<R, IMAG>((R)vd0[(size_t)y * N + x] * dts) : cmul(tx0[j], py0);
      q1 = vd1 ? phase_op<R, IMAG>((R)vd1[(size_t)y * N + x] * dts) : cmul(tx1[j], py1);
    }
    p0 = cmul(p0, q0);
    p1 = cmul(p1, q1);
    // Second coupling half step (tensor_propagator.py:258-260).
    if (apply) {
      const C s0 = cadd(cscale(p0, alpha), cmul(b01, p1));
      const C s1 = cadd(cmul(b10, p0), cscale(p1, alpha));
      p0 = s0;
      p1 = s1;
    }
    a0[j] = cscale(mul_phase<IMAG>(p0, u0), cf);
    a1[j] = cscale(mul_phase<IMAG>(p1, u1), cf);
  }
}

// ---------------------------------------------------------------- row pass

// The x-table values (potential phases of both components, e^{-i theta})
// of this thread's points J0 <= j < J1 of a row; `stand_in` replaces an
// absent potential table. Global loads: issued before a transform's last
// LDS exchange, they stay in flight across its waits.
template <int E, int T, int J0 = 0, int J1 = E, typename C>
__device__ __forceinline__ void load_x_tables(C (&tx0)[E], C (&tx1)[E], C (&tex)[E], const C* px0, const C* px1,
                                              const C* eth, const C* stand_in, int t) {
  const C* s0 = px0 ? px0 : stand_in;
  const C* s1 = px1 ? px1 : stand_in;
#pragma unroll
  for (int j = J0; j < J1; ++j) {
    tx0[j] = ld_global(s0 + t + j * T);
    tx1[j] = ld_global(s1 + t + j * T);
    tex[j] = ld_global(eth + t + j * T);
  }
}

// The row pointwise block point by point from point J on, point J's tables
// already loaded: point J + 1's tables are loaded first, so their latency
// overlaps point J's arithmetic, and at most two points' tables are live.
template <typename R, int N, int E, int T, bool FAST, int J, typename C>
__device__ __forceinline__ void pointwise_points(C (&a0)[E], C (&a1)[E], const InstDev& P, int dur, int y, int t,
                                                 double lam2, C py0, C py1, C (&tx0)[E], C (&tx1)[E], C (&tex)[E],
                                                 const C* px0, const C* px1, const C* eth, const C* tws) {
  if constexpr (J < E) {
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (J + 1 < E) load_x_tables<E, T, J + 1, J + 2>(tx0, tx1, tex, px0, px1, eth, tws, t);
    if (P.imag)
      row_pointwise<R, N, E, T, true, FAST, J, J + 1>(a0, a1, P, dur, y, t, lam2, py0, py1, tx0, tx1, tex);
    else
      row_pointwise<R, N, E, T, false, FAST, J, J + 1>(a0, a1, P, dur, y, t, lam2, py0, py1, tx0, tx1, tex);
    pointwise_points<R, N, E, T, FAST, J + 1>(a0, a1, P, dur, y, t, lam2, py0, py1, tx0, tx1, tex, px0, px1, eth,
                                              tws);
  }
}

// Population record of one full step from the column pass's S sums
// (tensor_propagator.py:194): pops[2 (step - 1) + s] for step > 0 within
// the run's capacity. Every thread of the wave gets the sums; thread 0
// writes.
template <bool COH = false>
__device__ __forceinline__ void record_pops(const InstDev& P, double* pops, int cap, int step, int nt2) {
  const double* part = P.part;
  double ps[2] = {wave_sum_strided<COH>(part + PART_S * nt2, nt2 / 2),
                  wave_sum_strided<COH>(part + PART_S * nt2 + nt2 / 2, nt2 / 2)};
  if (threadIdx.x == 0 && step > 0 && pops && step - 1 < cap) {
    const double tot = ps[0] + ps[1];
    pops[2 * (step - 1) + 0] = P.atom_num * ps[0] / tot;
    pops[2 * (step - 1) + 1] = P.atom_num * ps[1] / tot;
  }
}

// One block of ROWS row pairs (both components) of one instance (the body of
// row_pass_kernel and of the persistent step kernel). `gate` is called once
// every global load of the prologue has been issued; false drops the block.
// GENERAL: some slot has a dense potential or a coupling array; otherwise
// every slot takes the separable fast path.
template <typename R, int N, int E, int ROWS, bool GENERAL, int CM, class Gate>
__device__ __forceinline__ void row_item(const InstDev& P, typename Cx<R>::T* __restrict__ psi, int ny,
                                         const typename Cx<R>::T* __restrict__ tws, int dur, int col_w, int stream,
                                         int blk, double* smem, Gate gate) {
  typedef typename Cx<R>::T C;
  typedef SwzView<sizeof(C)> V;
  constexpr int T = N / E;
  // LDS: [ROWS*2*N exchange][N twiddles][ROWS*2*N second exchange buffer,
  // when double-buffered]. The x phase tables stay in global memory (cache
  // resident): copies in LDS would cost occupancy.
  constexpr size_t kData = (size_t)ROWS * 2 * N * sizeof(C);
  constexpr bool kTw = tw_in_lds(kData, N, sizeof(C));
  constexpr bool kDb = row_double_buffer(E, kData, kTw ? (size_t)N * sizeof(C) : 0);
  // A row's T threads are consecutive lanes: with T <= 64 each row's
  // transforms lie within one wave.
  constexpr bool kWave = T <= 64;
  const int nt2 = 2 * (N / col_w);
  const int row = threadIdx.x / T;
  const int t = threadIdx.x % T;
  const int y = blk * ROWS + row;
  C* const sm = reinterpret_cast<C*>(smem);
  C* r0 = psi + (size_t)y * N;
  C* r1 = r0 + (size_t)ny * N;
  C* lds0 = sm + (size_t)row * 2 * N;
  C* lds1 = lds0 + N;

  // Global loads first: the two rows, the twiddles (held in registers until
  // they are stored to LDS) and the norm partials; the kernel then waits for
  // memory once.
  C a0[E], a1[E];
  // CM: as in col_item; element (comp, y, x) at byte ((comp ny + y) N + x) sizeof(C)
  const CoherentBuf cb(psi, CM ? (uint32_t)(2 * (size_t)ny * N * sizeof(C)) : 0u);
  const uint32_t c0 = (uint32_t)(((size_t)y * N + t) * sizeof(C));
  const uint32_t c1 = c0 + (uint32_t)((size_t)ny * N * sizeof(C));
  if constexpr ((CM & CM_LOAD) != 0) {
#pragma unroll
    for (int j = 0; j < E; ++j) {
      a0[j] = cb.load<C>(c0 + (uint32_t)(j * T * sizeof(C)));
      a1[j] = cb.load<C>(c1 + (uint32_t)(j * T * sizeof(C)));
    }
  } else if (stream) {
#pragma unroll
    for (int j = 0; j < E; ++j) {
      a0[j] = ld_stream(r0 + t + j * T);
      a1[j] = ld_stream(r1 + t + j * T);
    }
  } else {
#pragma unroll
    for (int j = 0; j < E; ++j) {
      a0[j] = ld_global(r0 + t + j * T);
      a1[j] = ld_global(r1 + t + j * T);
    }
  }
  constexpr int kBlock = ROWS * T;
  const double* vd0 = P.v_dense[0];
  const double* vd1 = P.v_dense[1];
  C py0 = C{(R)1, (R)0}, py1 = C{(R)1, (R)0};
  const C* px0 = nullptr;
  const C* px1 = nullptr;
  if (!vd0) {
    py0 = ld_global(reinterpret_cast<const C*>(P.py[dur][0]) + y);
    px0 = reinterpret_cast<const C*>(P.px[dur][0]);
  }
  if (!vd1) {
    py1 = ld_global(reinterpret_cast<const C*>(P.py[dur][1]) + y);
    px1 = reinterpret_cast<const C*>(P.px[dur][1]);
  }
  const C* eth = reinterpret_cast<const C*>(P.etheta);
  constexpr bool kEarlyTables = E <= 2;
  C tx0[E], tx1[E], tex[E];
  if constexpr (kEarlyTables) load_x_tables<E, T>(tx0, tx1, tex, px0, px1, eth, tws, t);
  TableStage<kTw, N, kBlock, C> stw;
  stw.load(tws);
  if (!gate()) return;

  // Real-space normalization (tensor_tools.py:301-305) from the column
  // pass's k-space sums by Parseval: sum_r |IDFT2 v|^2 = Nx Ny sum_k |v|^2.
  // Every wave sums all partials itself (same order in every wave and
  // block: no barrier, deterministic); the loads join the ones above. The
  // twiddles stored here are published by the transform's first barrier
  // (its first stage uses none).
  const double* part = P.part;
  double nrm[1] = {wave_sum_strided<(CM & CM_LOAD) != 0>(part + PART_NRM * nt2, nt2)};
  const C* tw = stw.store(sm + ROWS * 2 * N);
  if constexpr (kWave && kTw) __syncthreads();  // publishes the twiddles staged by every wave
  const double lam2 = P.atom_num / (P.dv_r * (double)N * (double)ny * nrm[0]);

  C* const alt0 = sm + ROWS * 2 * N + (kTw ? N : 0) + (size_t)row * 2 * N;
  C* const alt1 = alt0 + N;
  if constexpr (kDb)
    fft_run2_db<N, E, true, V, kWave>(a0, a1, t, tw, lds0, lds1, alt0, alt1);
  else
    fft_run2<N, E, true, V, kWave>(a0, a1, t, tw, lds0, lds1);

  // The x-table values of this thread's points (an absent potential table
  // is stood in for by the global twiddle table: N entries, unused). Radix
  // 2: loaded with the state at the start (few registers to hold them
  // through the inverse transform). Radix 4: after the inverse transform,
  // in two halves, the second loaded only after the first half of the
  // points is done (the scheduling barrier keeps the compiler from hoisting
  // those loads), so at most half of them are live at once: 216 -> 162
  // VGPRs for the complex128 pass (3 waves per SIMD instead of 2), paid for
  // with one more L2 round trip per row. Complex128 radix >= 8: point by
  // point, each point's tables loaded while the previous point is computed
  // (296 -> 224 VGPRs, 2 waves per SIMD; with the twiddles in global memory
  // two 2048-point workgroups share a CU). Complex64 radix >= 8 uses the
  // halves, which already fit 3 waves per SIMD (point by point would only
  // add latency).
  if constexpr (E >= 8 && sizeof(C) == 16) {
    load_x_tables<E, T, 0, 1>(tx0, tx1, tex, px0, px1, eth, tws, t);
    pointwise_points<R, N, E, T, !GENERAL, 0>(a0, a1, P, dur, y, t, lam2, py0, py1, tx0, tx1, tex, px0, px1, eth,
                                              tws);
  } else {
    constexpr int H = kEarlyTables ? E : E / 2;
    if constexpr (!kEarlyTables) load_x_tables<E, T, 0, H>(tx0, tx1, tex, px0, px1, eth, tws, t);
    if (P.imag)
      row_pointwise<R, N, E, T, true, !GENERAL, 0, H>(a0, a1, P, dur, y, t, lam2, py0, py1, tx0, tx1, tex);
    else
      row_pointwise<R, N, E, T, false, !GENERAL, 0, H>(a0, a1, P, dur, y, t, lam2, py0, py1, tx0, tx1, tex);
    if constexpr (H < E) {
      __builtin_amdgcn_sched_barrier(0);
      load_x_tables<E, T, H, E>(tx0, tx1, tex, px0, px1, eth, tws, t);
      if (P.imag)
        row_pointwise<R, N, E, T, true, !GENERAL, H, E>(a0, a1, P, dur, y, t, lam2, py0, py1, tx0, tx1, tex);
      else
        row_pointwise<R, N, E, T, false, !GENERAL, H, E>(a0, a1, P, dur, y, t, lam2, py0, py1, tx0, tx1, tex);
    }
  }
  // Large rows: the forward transform and the stores index by an opaque copy
  // of t, so their LDS and state addresses are recomputed here instead of
  // being held in registers since the loads and the inverse transform, which
  // compute the same addresses (held, they cost up to 70 VGPRs and spill the
  // 4096-point complex128 pass). It pays where registers are what limits
  // the resident workgroups: 4096 points, and 1024-2048 points where the
  // LDS admits three waves per SIMD (complex64 2048 points: 192 -> 130
  // VGPRs, 58 -> 50 us per pass). Where the LDS allows two waves anyway
  // (complex128 2048 points) or small batched rows, whose workgroups fill
  // the CUs' slots already, only the recomputation remains (+2 %).
  constexpr size_t kLdsWg = kData * (kDb ? 2 : 1) + (kTw ? (size_t)N * sizeof(C) : 0);
  constexpr int kLdsWaves = (int)(kLdsBytes / kLdsWg) * ((ROWS * T + 63) / 64) / 4;
  constexpr bool kOpaque = N >= 4096 || (N >= 1024 && kLdsWaves >= 3);
  int tf = t;
  if constexpr (kOpaque) asm volatile("" : "+v"(tf));
  // The forward transform starts on the buffer pair the inverse one read
  // last but one, so no barrier is needed between them.
  if constexpr (kDb && fft_exchanges<N, E>() % 2)
    fft_run2_db<N, E, false, V, kWave>(a0, a1, tf, tw, alt0, alt1, lds0, lds1);
  else if constexpr (kDb)
    fft_run2_db<N, E, false, V, kWave>(a0, a1, tf, tw, lds0, lds1, alt0, alt1);
  else
    fft_run2<N, E, false, V, kWave>(a0, a1, tf, tw, lds0, lds1);
  if constexpr ((CM & CM_STORE) != 0) {
#pragma unroll
    for (int j = 0; j < E; ++j) {
      cb.store<C>(c0 + (uint32_t)(j * T * sizeof(C)), a0[j]);
      cb.store<C>(c1 + (uint32_t)(j * T * sizeof(C)), a1[j]);
    }
  } else if (stream) {
#pragma unroll
    for (int j = 0; j < E; ++j) {
      st_stream(r0 + tf + j * T, a0[j]);
      st_stream(r1 + tf + j * T, a1[j]);
    }
  } else {
#pragma unroll
    for (int j = 0; j < E; ++j) {
      st_global(r0 + tf + j * T, a0[j]);
      st_global(r1 + tf + j * T, a1[j]);
    }
  }
}

template <typename R, int N, int E, int ROWS, bool GENERAL>
__global__ __launch_bounds__(ROWS*(N / E)) void row_pass_kernel(const InstDev* __restrict__ insts,
                                                                RunDev* __restrict__ runs, void* direct_psi,
                                                                int direct_n,
                                                                const typename Cx<R>::T* __restrict__ tws,
                                                                int dur, int first_sub, int col_w, int stream) {
  typedef typename Cx<R>::T C;
  extern __shared__ double smem[];
  const int inst = blockIdx.z;
  RunDev& run = runs[inst];
  const InstDev& P = insts[inst];
  // State address and column length: kernel arguments for a single-slot
  // plan (loads first, active flag checked after them), else read together
  // with the active flag (one round of scalar loads before the state loads).
  const bool direct = direct_psi != nullptr;
  C* const psi = reinterpret_cast<C*>(direct ? direct_psi : P.psi);
  const int ny = direct ? direct_n : P.ny;
  if (!direct && (!run.active || !psi || ny <= 0)) return;
  auto gate = [&]() -> bool {
    if (direct) {
      // The prologue loads stay ahead of the flag test (a loaded slot's
      // state and tables stay allocated when it is inactive).
      asm volatile("" ::: "memory");
      if (!run.active) return false;
    }
    // Block 0 of the first sub-step records the previous full step's
    // populations and counts the step; the state leaves this pass in hybrid
    // form (no other block of this kernel reads the flag; the next column
    // pass does).
    if (blockIdx.x == 0) {
      if (first_sub) {
        const int step = run.step;
        record_pops(P, run.pops, run.pops_cap, step, 2 * (N / col_w));
        if (threadIdx.x == 0) run.step = step + 1;
      }
      if (threadIdx.x == 0) run.form = FORM_HYBRID;
    }
    return true;
  };
  row_item<R, N, E, ROWS, GENERAL, 0>(P, psi, ny, tws, dur, col_w, stream, blockIdx.x, smem, gate);
}

// Row pass with the two components on separate waves (the split layout):
// ROWS rows per workgroup, each taken by two waves of T = N/E = 64 threads,
// wave c transforming component c with exchanges inside the wave (no
// workgroup barrier). The two waves meet in LDS for the pointwise block,
// which all 128 threads of the row share (N/128 points each), then each wave
// transforms its component back. Against the layout above (both components
// per thread, transforms spread over several waves) this halves the LDS
// exchanges per transform at 512 points: 10.7 vs 11.8 us per pass for one
// 512^2 complex128 instance, 9.2 vs 9.5 us in complex64 (same-host A/B). The
// tuner picks it per grid and batch size; rows per workgroup are encoded as
// -ROWS in the shape tables.
template <typename R, int N, int E, int ROWS, bool GENERAL>
__global__ __launch_bounds__(ROWS * 2 * (N / E)) void row_split_kernel(const InstDev* __restrict__ insts,
                                                                       RunDev* __restrict__ runs, void* direct_psi,
                                                                       int direct_n,
                                                                       const typename Cx<R>::T* __restrict__ tws,
                                                                       int dur, int first_sub, int col_w, int stream) {
  typedef typename Cx<R>::T C;
  typedef SwzView<sizeof(C)> V;
  constexpr int T = N / E;
  static_assert(T == 64, "one wave per component");
  constexpr int TP = 2 * T;   // threads of one row
  constexpr int EP = N / TP;  // pointwise points per thread
  extern __shared__ double smem[];
  const int inst = blockIdx.z;
  RunDev& run = runs[inst];
  const InstDev& P = insts[inst];
  // State address and column length as in row_pass_kernel.
  const bool direct = direct_psi != nullptr;
  C* const psi = reinterpret_cast<C*>(direct ? direct_psi : P.psi);
  const int ny = direct ? direct_n : P.ny;
  if (!direct && (!run.active || !psi || ny <= 0)) return;
  const int rw = threadIdx.x / TP;
  const int y = blockIdx.x * ROWS + rw;
  const int tid = threadIdx.x % TP;
  const int comp = tid / T;
  const int t = tid % T;
  // LDS: [ROWS][2][N] data (swizzled within each component), then the
  // twiddles when they fit (tw_in_lds).
  C* const sm = reinterpret_cast<C*>(smem) + rw * 2 * N;
  C* const lds = sm + comp * N;
  C* const row = psi + ((size_t)comp * ny + y) * N;
  constexpr size_t kData = (size_t)ROWS * 2 * N * sizeof(C);
  constexpr bool kTw = tw_in_lds(kData, N, sizeof(C));
  C a[E];
  if (stream) {
#pragma unroll
    for (int j = 0; j < E; ++j) a[j] = ld_stream(row + t + j * T);
  } else {
#pragma unroll
    for (int j = 0; j < E; ++j) a[j] = ld_global(row + t + j * T);
  }
  const double* vd0 = P.v_dense[0];
  const double* vd1 = P.v_dense[1];
  C py0 = C{(R)1, (R)0}, py1 = C{(R)1, (R)0};
  const C* px0 = nullptr;
  const C* px1 = nullptr;
  if (!vd0) {
    py0 = ld_global(reinterpret_cast<const C*>(P.py[dur][0]) + y);
    px0 = reinterpret_cast<const C*>(P.px[dur][0]);
  }
  if (!vd1) {
    py1 = ld_global(reinterpret_cast<const C*>(P.py[dur][1]) + y);
    px1 = reinterpret_cast<const C*>(P.px[dur][1]);
  }
  const C* eth = reinterpret_cast<const C*>(P.etheta);
  TableStage<kTw, N, ROWS * TP, C> stw;
  stw.load(tws);
  if (direct) {
    asm volatile("" ::: "memory");
    if (!run.active) return;
  }
  // Block 0 of the first sub-step records the previous full step's
  // populations and counts the step; the state leaves in hybrid form.
  const int nt2 = 2 * (N / col_w);
  if (blockIdx.x == 0) {
    if (first_sub) {
      const int step = run.step;
      record_pops(P, run.pops, run.pops_cap, step, nt2);
      if (threadIdx.x == 0) run.step = step + 1;
    }
    if (threadIdx.x == 0) run.form = FORM_HYBRID;
  }
  // Real-space normalization from the column pass's k-space sums (Parseval).
  const double nrm = wave_sum_strided(P.part + PART_NRM * nt2, nt2);
  const C* tw = stw.store(reinterpret_cast<C*>(smem) + ROWS * 2 * N);
  if constexpr (kTw) __syncthreads();  // publishes the twiddles to every wave
  const double lam2 = P.atom_num / (P.dv_r * (double)N * (double)ny * nrm);
  fft_run<N, E, true, V, true>(a, t, tw, lds);
  // The x tables of this thread's pointwise points, loaded here so that they
  // arrive while the workgroup meets at the barrier (loaded at the start,
  // they kept the prologue waiting for the slot descriptor's pointers).
  C tx0[EP], tx1[EP], tex[EP];
  load_x_tables<EP, TP>(tx0, tx1, tex, px0, px1, eth, tws, tid);
#pragma unroll
  for (int j = 0; j < E; ++j) lds[V::at(t + j * T)] = a[j];
  __syncthreads();
  C p0[EP], p1[EP];
#pragma unroll
  for (int j = 0; j < EP; ++j) {
    p0[j] = sm[V::at(tid + j * TP)];
    p1[j] = sm[N + V::at(tid + j * TP)];
  }
  if (P.imag)
    row_pointwise<R, N, EP, TP, true, !GENERAL>(p0, p1, P, dur, y, tid, lam2, py0, py1, tx0, tx1, tex);
  else
    row_pointwise<R, N, EP, TP, false, !GENERAL>(p0, p1, P, dur, y, tid, lam2, py0, py1, tx0, tx1, tex);
#pragma unroll
  for (int j = 0; j < EP; ++j) {
    sm[V::at(tid + j * TP)] = p0[j];
    sm[N + V::at(tid + j * TP)] = p1[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < E; ++j) a[j] = lds[V::at(t + j * T)];
  fft_run<N, E, false, V, true>(a, t, tw, lds);
  if (stream) {
#pragma unroll
    for (int j = 0; j < E; ++j) st_stream(row + t + j * T, a[j]);
  } else {
#pragma unroll
    for (int j = 0; j < E; ++j) st_global(row + t + j * T, a[j]);
  }
}

// --------------------------------------------------------------- normalize
// k-space renormalization to atom_num (tensor_propagator.py:271) from the
// exit column pass's S sums, plus the populations of the last step
// (tensor_propagator.py:194). Grid: (chunks, 1, n_inst).
template <typename R>
__global__ __launch_bounds__(256) void normalize_kernel(const InstDev* __restrict__ insts,
                                                        RunDev* __restrict__ runs, int col_w) {
  typedef typename Cx<R>::T C;
  __shared__ double scratch[16];
  const int inst = blockIdx.z;
  RunDev& run = runs[inst];
  if (!run.active) return;
  const InstDev& P = insts[inst];
  const int nt2 = 2 * (P.nx / col_w);
  double s[2];
  part_sums(P.part, nt2, PART_S, 1, s, scratch);
  const double tot = s[0] + s[1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int step = run.step;
    if (step > 0 && run.pops && step - 1 < run.pops_cap) {
      run.pops[2 * (step - 1) + 0] = P.atom_num * s[0] / tot;
      run.pops[2 * (step - 1) + 1] = P.atom_num * s[1] / tot;
    }
  }
  const R kappa = (R)sqrt(P.atom_num / (P.dv_k * tot));
  C* psi = reinterpret_cast<C*>(P.psi);
  const size_t n = (size_t)2 * P.ny * P.nx;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    psi[i] = cscale(psi[i], kappa);
}

// Marks the state as normalized k-space. Separate launch so that no block of
// the normalize kernel can observe a half-updated flag.
static __global__ void set_form_kernel(RunDev* __restrict__ runs, int n, int form) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && runs[i].active) runs[i].form = form;
}

static __global__ void set_active_kernel(RunDev* __restrict__ runs, int slot, int active) {
  if (threadIdx.x == 0) runs[slot].active = active;
}

// Run-state (re)configuration for one slot; launched instead of copying so
// that it is ordered on the stream and graph-safe.
static __global__ void config_run_kernel(RunDev* __restrict__ runs, int slot, double* pops, int cap, int reset_step,
                                  int active, int form, double conv_tol) {
  if (threadIdx.x != 0) return;
  RunDev& r = runs[slot];
  r.pops = pops;
  r.pops_cap = cap;
  if (reset_step) {
    r.step = 0;
    r.converged = 0;
    r.n_checks = 0;
    r.e_prev = 0.0;
    r.steps_prev = 0;
  }
  if (active >= 0) r.active = active;
  if (form >= 0) r.form = form;
  r.conv_tol = conv_tol;
}

// ------------------------------------------------------------------ energy
// GP energy functional of a normalized k-space state, float64 sums:
//   E_kin = dv_k sum_s sum_k T_s |psik_s|^2
//   E_pot = dv_r sum_s sum_r V_s n_s, E_int = dv_r/2 sum (g_uu n0^2 + g_dd n1^2 + 2 g_ud n0 n1)
//   E_cpl = dv_r sum Omega Re(psi0* psi1 e^{-i theta}),  mu = (E_kin + E_pot + E_cpl + 2 E_int)/N
// with real space from chi = IDFT2(psik): n = |chi|^2 / (Nx Ny c)^2 (the
// (-1)^(x+y) of the shift cancels in every term). The transforms work in
// place: the column part leaves IDFT_y psik, the row part sums over
// IDFT_x IDFT_y psik and writes back DFT_x of it scaled by 1/(Nx Ny), and
// the restoring column part (FWD) applies DFT_y, which brings back psik up
// to rounding; no second copy of the state is needed.

// Column part: kinetic and norm sums in k-space, then IDFT_y in place; or
// (FWD) DFT_y in place, the last step of the evaluation.
template <typename R, int N, int E, int W, bool FWD>
__global__ __launch_bounds__(W*(N / E)) void energy_col_kernel(const InstDev* __restrict__ insts,
                                                               const RunDev* __restrict__ runs,
                                                               const typename Cx<R>::T* __restrict__ tws) {
  typedef typename Cx<R>::T C;
  typedef LdsView<W, 1 << 30, 0> V;
  constexpr int T = N / E;
  extern __shared__ double smem[];
  const int inst = blockIdx.z;
  if (!runs[inst].active) return;
  const InstDev& P = insts[inst];
  const int nx = P.nx;
  const int comp = blockIdx.y;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);  // neighbouring tiles share an L2
  const int c = threadIdx.x % W;
  const int t = threadIdx.x / W;
  const int jx = tile * W + c;
  C* lds = reinterpret_cast<C*>(smem) + c;
  C* col = reinterpret_cast<C*>(P.psi) + (size_t)comp * N * nx + jx;
  constexpr size_t kData = (size_t)W * N * sizeof(C);
  const C* tw = stage_twiddles<tw_in_lds(kData, N, sizeof(C))>(tws, N, reinterpret_cast<C*>(smem) + W * N);
  C a[E];
#pragma unroll
  for (int j = 0; j < E; ++j) a[j] = ld_global(col + (size_t)(t + j * T) * nx);
  if constexpr (FWD) {
    fft_run<N, E, false, V>(a, t, tw, lds);
#pragma unroll
    for (int j = 0; j < E; ++j) st_global(col + (size_t)(t + j * T) * nx, a[j]);
    return;
  }
  const double* td = P.t_dense[comp];
  const double tx = td ? 0.0 : P.tx[comp][jx];
  double kin = 0.0, s = 0.0;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int jy = t + j * T;
    const double tv = td ? td[(size_t)jy * nx + jx] : tx + P.ty[comp][jy];
    const double m = norm2d<R>(a[j]);
    kin += tv * m;
    s += m;
  }
  fft_run<N, E, true, V>(a, t, tw, lds);
#pragma unroll
  for (int j = 0; j < E; ++j) st_global(col + (size_t)(t + j * T) * nx, a[j]);
  double v[2] = {kin, s};
  block_sum<2>(v, smem);
  if (threadIdx.x == 0) {
    const int nt = nx / W;
    P.epart[EPART_KIN * P.estride + comp * nt + tile] = v[0];
    P.epart[EPART_S * P.estride + comp * nt + tile] = v[1];
  }
}

// Row part: IDFT_x of the rows, real-space sums, then DFT_x / (Nx Ny) in place.
template <typename R, int N, int E, int ROWS>
__global__ __launch_bounds__(ROWS*(N / E)) void energy_row_kernel(const InstDev* __restrict__ insts,
                                                                  const RunDev* __restrict__ runs,
                                                                  const typename Cx<R>::T* __restrict__ tws) {
  typedef typename Cx<R>::T C;
  typedef LdsView<1, 1 << 30, 0> V;
  constexpr int T = N / E;
  extern __shared__ double smem[];
  const int inst = blockIdx.z;
  if (!runs[inst].active) return;
  const InstDev& P = insts[inst];
  const int ny = P.ny;
  const int row = threadIdx.x / T;
  const int t = threadIdx.x % T;
  const int y = blockIdx.x * ROWS + row;
  C* r0 = reinterpret_cast<C*>(P.psi) + (size_t)y * N;
  C* r1 = r0 + (size_t)ny * N;
  C* lds0 = reinterpret_cast<C*>(smem) + (size_t)row * 2 * N;
  C* lds1 = lds0 + N;
  constexpr size_t kData = (size_t)ROWS * 2 * N * sizeof(C);
  const C* tw = stage_twiddles<tw_in_lds(kData, N, sizeof(C))>(tws, N, reinterpret_cast<C*>(smem) + ROWS * 2 * N);
  C a0[E], a1[E];
#pragma unroll
  for (int j = 0; j < E; ++j) {
    a0[j] = ld_global(r0 + t + j * T);
    a1[j] = ld_global(r1 + t + j * T);
  }
  fft_run2<N, E, true, V>(a0, a1, t, tw, lds0, lds1);
  const double* vd0 = P.v_dense[0];
  const double* vd1 = P.v_dense[1];
  const double vy0 = vd0 ? 0.0 : P.vy[0][y];
  const double vy1 = vd1 ? 0.0 : P.vy[1][y];
  const C* eth = reinterpret_cast<const C*>(P.etheta);
  double pot = 0.0, inter = 0.0, cpl = 0.0, nrm = 0.0;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int x = t + j * T;
    const double m0 = norm2d<R>(a0[j]), m1 = norm2d<R>(a1[j]);
    const double v0 = vd0 ? vd0[(size_t)y * N + x] : P.vx[0][x] + vy0;
    const double v1 = vd1 ? vd1[(size_t)y * N + x] : P.vx[1][x] + vy1;
    pot += v0 * m0 + v1 * m1;
    inter += P.g_uu * m0 * m0 + P.g_dd * m1 * m1 + 2.0 * P.g_ud * m0 * m1;
    nrm += m0 + m1;
    if (P.coupling) {
      const double om = P.coupling == 2 ? P.omega_dense[(size_t)y * N + x] : P.omega;
      // Re(conj(a0) a1 e^{-i theta})
      double cr = (double)a0[j].x * a1[j].x + (double)a0[j].y * a1[j].y;
      double ci = (double)a0[j].x * a1[j].y - (double)a0[j].y * a1[j].x;
      if (eth) {
        const double er = (double)eth[x].x, ei = (double)eth[x].y;
        cr = cr * er - ci * ei;
      }
      cpl += om * cr;
    }
  }
  fft_run2<N, E, false, V>(a0, a1, t, tw, lds0, lds1);
  const R back = (R)(1.0 / ((double)N * (double)ny));
#pragma unroll
  for (int j = 0; j < E; ++j) {
    st_global(r0 + t + j * T, cscale(a0[j], back));
    st_global(r1 + t + j * T, cscale(a1[j], back));
  }
  double v[4] = {pot, inter, cpl, nrm};
  block_sum<4>(v, smem);
  if (threadIdx.x == 0) {
    P.epart[EPART_POT * P.estride + blockIdx.x] = v[0];
    P.epart[EPART_INT * P.estride + blockIdx.x] = v[1];
    P.epart[EPART_CPL * P.estride + blockIdx.x] = v[2];
    P.epart[EPART_NRM * P.estride + blockIdx.x] = v[3];
  }
}

// One block per slot: reduce the partials, write the energy terms and run
// the convergence test |E - E_prev| / |E| per full step <= conv_tol (a
// converged slot is deactivated, freezing its normalized k-space state).
static __global__ __launch_bounds__(256) void energy_final_kernel(const InstDev* __restrict__ insts,
                                                           RunDev* __restrict__ runs, int n_col_tiles,
                                                           int n_row_blocks, double* __restrict__ out) {
  __shared__ double scratch[32];
  const int inst = blockIdx.x;
  RunDev& run = runs[inst];
  if (!run.active) return;
  const InstDev& P = insts[inst];
  double v[6] = {0, 0, 0, 0, 0, 0};
  const int ncol = 2 * n_col_tiles;
  for (int i = threadIdx.x; i < ncol; i += blockDim.x) {
    v[0] += P.epart[EPART_KIN * P.estride + i];
    v[1] += P.epart[EPART_S * P.estride + i];
  }
  for (int i = threadIdx.x; i < n_row_blocks; i += blockDim.x) {
    v[2] += P.epart[EPART_POT * P.estride + i];
    v[3] += P.epart[EPART_INT * P.estride + i];
    v[4] += P.epart[EPART_CPL * P.estride + i];
    v[5] += P.epart[EPART_NRM * P.estride + i];
  }
  block_sum<6>(v, scratch);
  if (threadIdx.x != 0) return;
  const double kappa2 = P.atom_num / (P.dv_k * v[1]);
  const double nn = (double)P.nx * (double)P.ny * P.fft_c;
  const double a = kappa2 / (nn * nn);
  const double e_kin = kappa2 * P.dv_k * v[0];
  const double e_pot = P.dv_r * a * v[2];
  const double e_int = 0.5 * P.dv_r * a * a * v[3];
  const double e_cpl = P.dv_r * a * v[4];
  const double atoms = P.dv_r * a * v[5];
  const double e_tot = e_kin + e_pot + e_int + e_cpl;
  const double mu = (e_kin + e_pot + e_cpl + 2.0 * e_int) / atoms;
  double rel = -1.0;
  if (run.n_checks > 0) {
    const int steps = run.step - run.steps_prev;
    rel = fabs(e_tot - run.e_prev) / fabs(e_tot) / (double)(steps > 0 ? steps : 1);
    if (run.conv_tol > 0.0 && steps > 0 && rel <= run.conv_tol) {
      run.converged = 1;
      run.active = 0;
    }
  }
  const double e[8] = {e_tot, e_kin, e_pot, e_int, e_cpl, mu, atoms, rel};
  for (int i = 0; i < 8; ++i) {
    run.energy[i] = e[i];
    if (out) out[8 * inst + i] = e[i];
  }
  run.e_prev = e_tot;
  run.steps_prev = run.step;
  run.n_checks += 1;
}

// ------------------------------------------------- standalone primitives
// Device forms of tensor_tools.fft_2d / ifft_2d (tensor_tools.py:201-258):
//   fft_2d : fftshift(FFTN(psi)) * c  = DFT_y DFT_x [(-1)^(x+y) c psi]
//   ifft_2d: IFFTN(ifftshift(psik))/c = (-1)^(x+y) / (Nx Ny c) IDFT_x IDFT_y psik
// Both kernels transform up to two arrays per launch (the two spinor
// components), array blockIdx.z of the pair.
// Rows: optional pre- (forward) or post- (inverse) factor (-1)^(x+y) * scale.
template <typename R, int N, int E, int ROWS>
__global__ __launch_bounds__(ROWS*(N / E)) void fft_rows_kernel(FftPair io, int inv, int shift, double scale,
                                                                const typename Cx<R>::T* __restrict__ tws) {
  typedef typename Cx<R>::T C;
  typedef LdsView<1, 1 << 30, 0> V;
  constexpr int T = N / E;
  extern __shared__ double smem[];
  const int row = threadIdx.x / T;
  const int t = threadIdx.x % T;
  const int y = blockIdx.x * ROWS + row;
  const C* src = reinterpret_cast<const C*>(io.in[blockIdx.z]) + (size_t)y * N;
  C* dst = reinterpret_cast<C*>(io.out[blockIdx.z]) + (size_t)y * N;
  C* lds = reinterpret_cast<C*>(smem) + (size_t)row * N;
  constexpr size_t kData = (size_t)ROWS * N * sizeof(C);
  const C* tw = stage_twiddles<tw_in_lds(kData, N, sizeof(C))>(tws, N, reinterpret_cast<C*>(smem) + ROWS * N);
  C a[E];
#pragma unroll
  for (int j = 0; j < E; ++j) a[j] = ld_global(src + t + j * T);
  // (-1)^(x+y) with x = t + j*T: T is even, so the sign depends on t and y only.
  const R f = (R)(shift && ((t + y) & 1) ? -scale : scale);
  if (!inv) {
#pragma unroll
    for (int j = 0; j < E; ++j) a[j] = cscale(a[j], f);
    fft_run<N, E, false, V>(a, t, tw, lds);
  } else {
    fft_run<N, E, true, V>(a, t, tw, lds);
#pragma unroll
    for (int j = 0; j < E; ++j) a[j] = cscale(a[j], f);
  }
  // Store addresses from an opaque copy of t: recomputed here rather than
  // held in registers since the loads (as in the pass kernels).
  int tf = t;
  asm volatile("" : "+v"(tf));
#pragma unroll
  for (int j = 0; j < E; ++j) st_global(dst + tf + j * T, a[j]);
}

template <typename R, int N, int E, int W>
__global__ __launch_bounds__(W*(N / E)) void fft_cols_kernel(FftPair io, int nx, int inv,
                                                             const typename Cx<R>::T* __restrict__ tws) {
  typedef typename Cx<R>::T C;
  typedef LdsView<W, 1 << 30, 0> V;
  constexpr int T = N / E;
  extern __shared__ double smem[];
  const C* in = reinterpret_cast<const C*>(io.in[blockIdx.z]);
  C* out = reinterpret_cast<C*>(io.out[blockIdx.z]);
  const int c = threadIdx.x % W;
  const int t = threadIdx.x / W;
  const int jx = xcd_tile(blockIdx.x, gridDim.x) * W + c;  // neighbouring tiles share an L2
  C* lds = reinterpret_cast<C*>(smem) + c;
  constexpr size_t kData = (size_t)W * N * sizeof(C);
  const C* tw = stage_twiddles<tw_in_lds(kData, N, sizeof(C))>(tws, N, reinterpret_cast<C*>(smem) + W * N);
  C a[E];
#pragma unroll
  for (int j = 0; j < E; ++j) a[j] = ld_global(in + (size_t)(t + j * T) * nx + jx);
  if (inv)
    fft_run<N, E, true, V>(a, t, tw, lds);
  else
    fft_run<N, E, false, V>(a, t, tw, lds);
  // Store addresses from an opaque copy of t (complex64 2048-point columns:
  // 84 -> 60 VGPRs, room for a second 1024-thread workgroup per CU).
  int tf = t;
  asm volatile("" : "+v"(tf));
#pragma unroll
  for (int j = 0; j < E; ++j) st_global(out + (size_t)(tf + j * T) * nx + jx, a[j]);
}

// Per-block sums of |psi0|^2 and |psi1|^2 (float64), written to part[2*block + s].
template <typename R>
__global__ __launch_bounds__(256) void density_sums_kernel(const typename Cx<R>::T* __restrict__ p0,
                                                           const typename Cx<R>::T* __restrict__ p1,
                                                           size_t n, double* __restrict__ part) {
  __shared__ double scratch[16];
  double v[2] = {0.0, 0.0};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    v[0] += norm2d<R>(p0[i]);
    v[1] += norm2d<R>(p1[i]);
  }
  block_sum<2>(v, scratch);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x + 0] = v[0];
    part[2 * blockIdx.x + 1] = v[1];
  }
}

// norm() (tensor_tools.py:261-311): nf = sum(dens0 + dens1) * vol / atom_num;
// psi /= sqrt(nf); dens = |psi_before|^2 / nf. Every block reduces the same
// partials in the same order, so all blocks use the same nf.
template <typename R>
__global__ __launch_bounds__(256) void norm_apply_kernel(typename Cx<R>::T* __restrict__ p0,
                                                         typename Cx<R>::T* __restrict__ p1, size_t n,
                                                         const double* __restrict__ part, int nblk, double vol,
                                                         double atom_num, R* __restrict__ d0,
                                                         R* __restrict__ d1, double* __restrict__ nf_out) {
  __shared__ double scratch[16];
  double v[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) v[0] += part[2 * i] + part[2 * i + 1];
  block_sum<2>(v, scratch);
  const double nf = v[0] * vol / atom_num;
  if (nf_out && blockIdx.x == 0 && threadIdx.x == 0) *nf_out = nf;
  const R inv_sqrt = (R)(1.0 / sqrt(nf));
  const R inv_nf = (R)(1.0 / nf);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const typename Cx<R>::T a = p0[i], b = p1[i];
    if (d0) d0[i] = (R)((a.x * a.x + a.y * a.y) * inv_nf);
    if (d1) d1[i] = (R)((b.x * b.x + b.y * b.y) * inv_nf);
    p0[i] = cscale(a, inv_sqrt);
    p1[i] = cscale(b, inv_sqrt);
  }
}

// Final reduction of density_sums partials into out2[s] = vol * sum |psi_s|^2.
static __global__ __launch_bounds__(256) void pops_final_kernel(const double* __restrict__ part, int nblk, double vol,
                                                         double* __restrict__ out2) {
  __shared__ double scratch[16];
  double v[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
    v[0] += part[2 * i];
    v[1] += part[2 * i + 1];
  }
  block_sum<2>(v, scratch);
  if (threadIdx.x == 0) {
    out2[0] = v[0] * vol;
    out2[1] = v[1] * vol;
  }
}

}  // namespace sgpe
