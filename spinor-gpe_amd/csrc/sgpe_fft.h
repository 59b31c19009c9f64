// In-register mixed-radix FFT for power-of-two lengths 64..4096.
//
// Schedule: Stockham auto-sort, natural order in and out. A transform of
// length N is done by T = N/E threads; thread t holds a[j] = X[t + j*T]
// (j < E). Each stage of radix R <= E and span L turns the virtual thread
// v = t + u*T (u < E/R) into one radix-R butterfly on X[v + r*N/R], twiddled
// by w_{L*R}^{r*(v mod L)}, and scatters the outputs to Y[(v - k)*R + k + q*L].
// The first stage needs no twiddles, the last stage's outputs already sit at
// t + j*T, so a transform touches LDS only between stages and its global
// loads/stores are unit-stride across threads (coalesced).
//
// Unnormalized: forward uses exp(-2 pi i/N), inverse exp(+2 pi i/N), no 1/N.
// Twiddles come from a host-built stage-major table (rounded from float64
// values computed with exact octant reduction): the stage of span L and
// radix R keeps w_{L R}^{r k} for k < L, 1 <= r < R at (L - 1) + k (R - 1) +
// (r - 1), so a butterfly reads R - 1 consecutive entries and the table has
// N - 1 entries in all. Kernels copy it to LDS when it fits next to their
// data (tw_in_lds) and read it from global memory otherwise.
#pragma once

#include "sgpe_common.h"

namespace sgpe {

constexpr size_t kLdsBytes = 160 * 1024;

// Whether a kernel with data_bytes of LDS exchange buffers also holds the
// n-entry twiddle table in LDS (same predicate on host and device): when
// both fit in less than half the LDS (two workgroups per CU with room to
// spare), or when the data alone already takes more than half (one
// workgroup per CU either way) and the table still fits. In between, the
// table stays in global memory so that two workgroups share a CU with
// room: the 2048-point complex128 row pass (64 KiB of data, 32 KiB of
// twiddles) and the 2048-point complex64 column pass with 4 columns (64 KiB
// and 16 KiB, exactly half) both run faster that way (measured).
constexpr bool tw_in_lds(size_t data_bytes, int n, size_t elem_bytes) {
  const size_t total = data_bytes + (size_t)n * elem_bytes;
  return total < kLdsBytes / 2 || (data_bytes > kLdsBytes / 2 && total <= kLdsBytes);
}

// Multiply by -i (forward) or +i (inverse).
template <bool INV, typename C> __device__ __forceinline__ C rot90(C a) {
  return INV ? C{-a.y, a.x} : C{a.y, -a.x};
}

template <bool INV, typename C> __device__ __forceinline__ void dft2(C& a, C& b) {
  C s = cadd(a, b), d = csub(a, b);
  a = s;
  b = d;
}

template <bool INV, typename C> __device__ __forceinline__ void dft4(C& b0, C& b1, C& b2, C& b3) {
  C t0 = cadd(b0, b2), t1 = csub(b0, b2);
  C t2 = cadd(b1, b3), t3 = rot90<INV>(csub(b1, b3));
  b0 = cadd(t0, t2);
  b2 = csub(t0, t2);
  b1 = cadd(t1, t3);
  b3 = csub(t1, t3);
}

// w8^1 * a and w8^3 * a with w8 = exp(-/+ i pi/4).
template <bool INV, typename C, typename R> __device__ __forceinline__ C w8_1(C a, R h) {
  return INV ? C{(a.x - a.y) * h, (a.x + a.y) * h} : C{(a.x + a.y) * h, (a.y - a.x) * h};
}
template <bool INV, typename C, typename R> __device__ __forceinline__ C w8_3(C a, R h) {
  return INV ? C{-(a.x + a.y) * h, (a.x - a.y) * h} : C{(a.y - a.x) * h, -(a.x + a.y) * h};
}

template <bool INV, typename C> __device__ __forceinline__ void dft8(C* b) {
  typedef decltype(b[0].x) R;
  const R h = (R)0.70710678118654752440;  // 1/sqrt(2)
  C e0 = b[0], e1 = b[2], e2 = b[4], e3 = b[6];
  C o0 = b[1], o1 = b[3], o2 = b[5], o3 = b[7];
  dft4<INV>(e0, e1, e2, e3);
  dft4<INV>(o0, o1, o2, o3);
  o1 = w8_1<INV>(o1, h);
  o2 = rot90<INV>(o2);
  o3 = w8_3<INV>(o3, h);
  b[0] = cadd(e0, o0);
  b[4] = csub(e0, o0);
  b[1] = cadd(e1, o1);
  b[5] = csub(e1, o1);
  b[2] = cadd(e2, o2);
  b[6] = csub(e2, o2);
  b[3] = cadd(e3, o3);
  b[7] = csub(e3, o3);
}

template <bool INV, typename C> __device__ __forceinline__ void dft16(C* b) {
  typedef decltype(b[0].x) R;
  const R h = (R)0.70710678118654752440;   // cos(pi/4)
  const R c1 = (R)0.92387953251128675613;  // cos(pi/8)
  const R s1 = (R)0.38268343236508977173;  // sin(pi/8)
  const R sg = INV ? (R)1 : (R)-1;         // sign of the twiddle's imaginary part
  C e[8], o[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    e[i] = b[2 * i];
    o[i] = b[2 * i + 1];
  }
  dft8<INV>(e);
  dft8<INV>(o);
  // o[q] *= w16^q
  o[1] = cmul(o[1], C{c1, sg * s1});
  o[2] = w8_1<INV>(o[2], h);
  o[3] = cmul(o[3], C{s1, sg * c1});
  o[4] = rot90<INV>(o[4]);
  o[5] = cmul(o[5], C{-s1, sg * c1});
  o[6] = w8_3<INV>(o[6], h);
  o[7] = cmul(o[7], C{-c1, sg * s1});
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    b[q] = cadd(e[q], o[q]);
    b[q + 8] = csub(e[q], o[q]);
  }
}

template <int R, bool INV, typename C> __device__ __forceinline__ void dft(C* b) {
  if constexpr (R == 2) {
    dft2<INV>(b[0], b[1]);
  } else if constexpr (R == 4) {
    dft4<INV>(b[0], b[1], b[2], b[3]);
  } else if constexpr (R == 8) {
    dft8<INV>(b);
  } else {
    static_assert(R == 16, "radix must be 2, 4, 8 or 16");
    dft16<INV>(b);
  }
}

// LDS view of one transform: element i lives at base[i*STRIDE + (i/PADB)*STRIDE*PADW].
// STRIDE interleaves several transforms (columns); PADB/PADW insert padding
// to break power-of-two bank strides (PADW = 0: none).
template <int STRIDE, int PADB, int PADW> struct LdsView {
  static __device__ __forceinline__ int at(int i) { return (i + (PADW ? (i / PADB) * PADW : 0)) * STRIDE; }
  static constexpr int span(int n) { return (n + (PADW ? (n / PADB) * PADW : 0)) * STRIDE; }
};

// LDS view of one transform with an XOR swizzle and no padding: element i
// sits at i ^ f(i / G), where G elements of ELEM_BYTES fill one 128-byte
// window of LDS banks (8 of 16 bytes, 16 of 8 bytes) and f permutes the
// slots of a window. A Stockham stage of radix R and span L scatters
// consecutive lanes R elements apart (L = 1) or in runs of L, which without
// the swizzle lands several lanes of one window on the same banks; f spreads
// them (chosen by counting conflicts of every compiled shape's stages:
// 2.0x / 2.7x the conflict-free cycle count on average without it, 1.02x /
// 1.05x with it, for 16- / 8-byte elements). Unit-stride reads stay
// conflict-free (f is constant within a window) and the footprint is
// unchanged.
template <int ELEM_BYTES> struct SwzView {
  static constexpr int G = 128 / ELEM_BYTES;
  static __device__ __forceinline__ int at(int i) {
    const int w = i / G;
    int f;
    if constexpr (G == 8)
      f = (((w & 7) * 3) ^ (w >> 3)) & 7;
    else
      f = (w * 13) & 15;
    return i ^ f;
  }
  static constexpr int span(int n) { return n; }
};

// One Stockham stage. Non-final stages write their outputs to LDS; the final
// stage leaves output Y[t + j*T] in a[j].
template <int N, int E, int R, int L, bool INV, bool LAST, class V, typename C>
__device__ __forceinline__ void fft_stage(C (&a)[E], int t, const C* tw, C* lds) {
  constexpr int T = N / E;
  constexpr int U = E / R;
  C out[LAST ? E : 1];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int v = t + u * T;
    const int k = v & (L - 1);
    C b[R];
#pragma unroll
    for (int r = 0; r < R; ++r) b[r] = a[u + r * U];
    if constexpr (L > 1) {
      const C* wk = tw + (L - 1) + k * (R - 1) - 1;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        C w = wk[r];
        if (INV) w.y = -w.y;
        b[r] = cmul(b[r], w);
      }
    }
    dft<R, INV>(b);
    if constexpr (LAST) {
#pragma unroll
      for (int q = 0; q < R; ++q) out[u + q * U] = b[q];
    } else {
      const int base = (v - k) * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) lds[V::at(base + q * L)] = b[q];
    }
  }
  if constexpr (LAST) {
#pragma unroll
    for (int j = 0; j < E; ++j) a[j] = out[j];
  }
}

// Orders the LDS writes of one exchange before its reads, and those reads
// before the next exchange's writes. WAVE: every transform lies within one
// wave (its threads are consecutive lanes of one wave), whose LDS
// instructions execute in order, so waiting for the wave's own LDS traffic
// (and keeping the compiler from moving LDS accesses across) suffices;
// otherwise the whole workgroup synchronizes.
template <bool WAVE> __device__ __forceinline__ void exchange_sync() {
  if constexpr (WAVE)
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  else
    __syncthreads();
}

// Full transform: stages of radix E while N/L >= E, then the remainder radix.
// Every thread of the block must call this (it contains barriers unless
// WAVE: see exchange_sync; tables staged in LDS by other waves must then be
// published by a barrier before the call).
template <int N, int E, bool INV, class V, bool WAVE = false, int L = 1, typename C>
__device__ __forceinline__ void fft_run(C (&a)[E], int t, const C* tw, C* lds) {
  constexpr int R = (N / L >= E) ? E : N / L;
  constexpr bool LAST = (L * R == N);
  constexpr int T = N / E;
  fft_stage<N, E, R, L, INV, LAST, V>(a, t, tw, lds);
  if constexpr (!LAST) {
    exchange_sync<WAVE>();
#pragma unroll
    for (int j = 0; j < E; ++j) a[j] = lds[V::at(t + j * T)];
    exchange_sync<WAVE>();
    fft_run<N, E, INV, V, WAVE, L * R>(a, t, tw, lds);
  }
}

// Two independent transforms (e.g. both spinor components of one row) that
// share barriers: lds0/lds1 are disjoint views.
template <int N, int E, bool INV, class V, bool WAVE = false, int L = 1, typename C>
__device__ __forceinline__ void fft_run2(C (&a)[E], C (&b)[E], int t, const C* tw, C* lds0, C* lds1) {
  constexpr int R = (N / L >= E) ? E : N / L;
  constexpr bool LAST = (L * R == N);
  constexpr int T = N / E;
  fft_stage<N, E, R, L, INV, LAST, V>(a, t, tw, lds0);
  fft_stage<N, E, R, L, INV, LAST, V>(b, t, tw, lds1);
  if constexpr (!LAST) {
    exchange_sync<WAVE>();
#pragma unroll
    for (int j = 0; j < E; ++j) {
      a[j] = lds0[V::at(t + j * T)];
      b[j] = lds1[V::at(t + j * T)];
    }
    exchange_sync<WAVE>();
    fft_run2<N, E, INV, V, WAVE, L * R>(a, b, t, tw, lds0, lds1);
  }
}


// Double-buffered variant of fft_run2: stage s writes the buffer pair
// (l0, l1) or (m0, m1) alternately, so a buffer is rewritten only after the
// next exchange's barrier, which every thread reaches after its reads of
// that buffer: one barrier per exchange instead of two.
template <int N, int E, bool INV, class V, bool WAVE = false, int L = 1, typename C>
__device__ __forceinline__ void fft_run2_db(C (&a)[E], C (&b)[E], int t, const C* tw, C* l0, C* l1, C* m0, C* m1) {
  constexpr int R = (N / L >= E) ? E : N / L;
  constexpr bool LAST = (L * R == N);
  constexpr int T = N / E;
  fft_stage<N, E, R, L, INV, LAST, V>(a, t, tw, l0);
  fft_stage<N, E, R, L, INV, LAST, V>(b, t, tw, l1);
  if constexpr (!LAST) {
    exchange_sync<WAVE>();
#pragma unroll
    for (int j = 0; j < E; ++j) {
      a[j] = l0[V::at(t + j * T)];
      b[j] = l1[V::at(t + j * T)];
    }
    fft_run2_db<N, E, INV, V, WAVE, L * R>(a, b, t, tw, m0, m1, l0, l1);
  }
}

// Number of LDS exchanges of a transform (stages - 1).
template <int N, int E, int L = 1> constexpr int fft_exchanges() {
  constexpr int R = (N / L >= E) ? E : N / L;
  if constexpr (L * R == N)
    return 0;
  else
    return 1 + fft_exchanges<N, E, L * R>();
}

// Whether a row pass double-buffers its exchanges: for radix 2 only (many
// stages, where the saved barriers matter) and only while the second buffer
// keeps the kernel within 64 KiB of LDS. Radix-4 row passes run batched
// plans, where the second buffer's LDS would cap the workgroups per CU
// below what their registers allow (complex128: 6 instead of 4). Same
// predicate on host and device.
constexpr size_t kDbLdsBytes = 64 * 1024;
constexpr bool row_double_buffer(int e, size_t data_bytes, size_t other_bytes) {
  return e <= 2 && 2 * data_bytes + other_bytes <= kDbLdsBytes;
}

// Twiddle table for the FFT stages: copies the n - 1 stage-major entries to
// LDS at `dst` when IN_LDS (the first exchange barrier of fft_run orders the
// copy before any read) and returns the table to read from.
template <bool IN_LDS, typename C>
__device__ __forceinline__ const C* stage_twiddles(const C* __restrict__ tw, int n, C* dst) {
  if constexpr (IN_LDS) {
    for (int i = threadIdx.x; i < n - 1; i += blockDim.x) dst[i] = tw[i];
    return dst;
  } else {
    (void)n;
    (void)dst;
    return tw;
  }
}

// Registers of one thread's share of a table: K entries, each held as two
// scalar members (member access to scalars keeps them in VGPRs, where an
// indexed array of vector-typed entries may be put in scratch).
template <typename C> struct CxScalar;
template <> struct CxScalar<double2> { typedef double T; };
template <> struct CxScalar<float2> { typedef float T; };
template <int K, typename C> struct TableRegs {
  typename CxScalar<C>::T re, im;
  TableRegs<K - 1, C> rest;
  template <int N, int BLOCK> __device__ __forceinline__ void load(const C* __restrict__ s, int i) {
    if (N % BLOCK == 0 || i < N) {
      const C v = ld_global(s + i);
      re = v.x;
      im = v.y;
    }
    rest.template load<N, BLOCK>(s, i + BLOCK);
  }
  template <int N, int BLOCK> __device__ __forceinline__ void store(C* dst, int i) const {
    if (N % BLOCK == 0 || i < N) dst[i] = C{re, im};
    rest.template store<N, BLOCK>(dst, i + BLOCK);
  }
};
template <typename C> struct TableRegs<0, C> {
  template <int N, int BLOCK> __device__ __forceinline__ void load(const C*, int) {}
  template <int N, int BLOCK> __device__ __forceinline__ void store(C*, int) const {}
};
// Two-phase copy of an N-entry table into LDS by a block of BLOCK threads:
// load() issues the global loads into registers, store() writes them to LDS
// later, so that the loads of a kernel's prologue are all in flight together.
// With IN_LDS false it only forwards the global pointer. The caller makes the
// stores visible with a barrier before the first read.
template <bool IN_LDS, int N, int BLOCK, typename C> struct TableStage {
  static constexpr int kPer = IN_LDS ? (N + BLOCK - 1) / BLOCK : 0;
  TableRegs<kPer, C> regs;
  const C* src = nullptr;
  __device__ __forceinline__ void load(const C* __restrict__ s) {
    src = s;
    regs.template load<N, BLOCK>(s, threadIdx.x);
  }
  __device__ __forceinline__ const C* store(C* dst) {
    if constexpr (IN_LDS) {
      regs.template store<N, BLOCK>(dst, threadIdx.x);
      return dst;
    } else {
      (void)dst;
      return src;
    }
  }
};

// Deterministic block-wide sum of NV doubles per thread (fixed tree order).
// `scratch` needs NV * (blockDim.x / 64) doubles. All threads get the result.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* scratch) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwave = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_xor(v[i], off, 64);
  }
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) scratch[i * nwave + wave] = v[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = 0.0;
    for (int w = 0; w < nwave; ++w) s += scratch[i * nwave + w];
    v[i] = s;
  }
  __syncthreads();
}

// Block-wide sum for the end of a kernel: the same fixed tree order as
// block_sum, but only thread 0 gets the result. `scratch` (NV * waves
// doubles) must be LDS that no other code of the kernel touches, so that
// one barrier suffices (block_sum needs three: its scratch may overlap
// buffers still being read, and it is reused afterwards).
template <int NV>
__device__ __forceinline__ void block_sum_thread0(double (&v)[NV], double* scratch) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwave = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_xor(v[i], off, 64);
  }
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) scratch[i * nwave + wave] = v[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double s = 0.0;
      for (int w = 0; w < nwave; ++w) s += scratch[i * nwave + w];
      v[i] = s;
    }
  }
}

// LDS for the column pass's end-of-pass sums (2 doubles per wave, up to 16
// waves), placed after its exchange buffer and tables when that still fits
// in LDS (same predicate on host and device); otherwise the sums reuse the
// exchange buffer (block_sum).
constexpr size_t kColSumBytes = 256;
constexpr bool col_sum_separate(size_t base_bytes) { return base_bytes + kColSumBytes <= kLdsBytes; }

}  // namespace sgpe
