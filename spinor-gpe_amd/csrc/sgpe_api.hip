// Host side of libsgpe: contexts, plans, kernel registry, launch sequences,
// hipGraph capture and the extern "C" entry points declared in sgpe.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/sgpe.h"
#include "sgpe_generic.h"
#include "sgpe_persist.h"
#include "sgpe_registry.h"

using namespace sgpe;

// ------------------------------------------------------------------ errors

static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
  try {
    g_err = msg;
  } catch (...) {
  }
  return code;
}

// Runs an entry point's body so that no C++ exception crosses the C ABI: a
// failed host allocation (std::bad_alloc from containers or new) becomes
// SGPE_E_NOMEM, anything else SGPE_E_HIP, each with a message.
template <class F> static int guarded(F&& body) noexcept {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return fail(SGPE_E_NOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(SGPE_E_HIP, std::string("internal error: ") + e.what());
  } catch (...) {
    return fail(SGPE_E_HIP, "internal error");
  }
}

#define HIP_TRY(expr)                                                                             \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      return fail(SGPE_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));                  \
  } while (0)

// -------------------------------------------------------------- twiddles
// tw[m] = exp(-2 pi i m / n), from libm cos/sin of angles reduced exactly to
// [0, pi/4] by the symmetries of the unit circle (n is a power of two, so
// m/n and the reductions are exact).
static void twiddles_f64(int n, std::vector<double>& re, std::vector<double>& im) {
  re.resize(n);
  im.resize(n);
  const int q = n / 4, o = n / 8;
  for (int m = 0; m < n; ++m) {
    const int quad = m / q, j = m % q;
    double c, s;  // cos, sin of 2 pi j / n, j in [0, n/4)
    if (j <= o) {
      const double a = (2.0 * M_PI / n) * j;
      c = std::cos(a);
      s = std::sin(a);
    } else {
      const double a = (2.0 * M_PI / n) * (q - j);
      c = std::sin(a);
      s = std::cos(a);
    }
    // rotate by quad * pi/2: (c, s) -> (-s, c)
    double cc = c, ss = s;
    for (int r = 0; r < quad; ++r) {
      const double t = cc;
      cc = -ss;
      ss = t;
    }
    re[m] = cc;
    im[m] = -ss;  // exp(-i angle)
  }
}

// ------------------------------------------------------- kernel registry

static const KEntry* find_entry(int kind, int prec, int n, int e, int p) {
  const KTable t = prec == SGPE_C128 ? registry_f64(kind) : registry_f32(kind);
  for (int i = 0; i < t.count; ++i)
    if (t.entries[i].n == n && t.entries[i].e == e && t.entries[i].p == p) return &t.entries[i];
  return nullptr;
}

static const size_t kMaxLds = kLdsBytes;

// Device allocations of the library go through dev_alloc, so that a test can
// make the next n of them fail (sgpe_debug_fail_allocs) and exercise the
// failure paths of the entry points.
static std::atomic<int> g_fail_allocs{0};
static hipError_t dev_alloc(void** ptr, size_t bytes) {
  int n = g_fail_allocs.load();
  while (n > 0) {
    if (g_fail_allocs.compare_exchange_weak(n, n - 1)) {
      *ptr = nullptr;
      return hipErrorOutOfMemory;
    }
  }
  return hipMalloc(ptr, bytes);
}

static size_t csize(int prec) { return prec == SGPE_C128 ? 16 : 8; }

// Default shapes per length (before any tuning).
static void default_row(int prec, int n, int* e, int* rows) {
  (void)prec;
  switch (n) {
    case 64: *e = 8; *rows = 8; break;
    case 128: *e = 8; *rows = 4; break;
    case 256: *e = 8; *rows = 2; break;
    case 512: *e = 8; *rows = 1; break;
    case 1024: *e = 8; *rows = 1; break;
    case 2048: *e = 8; *rows = 1; break;
    default: *e = 16; *rows = 1; break;
  }
}
static void default_col(int prec, int n, int* e, int* w) {
  switch (n) {
    case 64: *e = 8; *w = 8; break;
    case 128: *e = 8; *w = 8; break;
    case 256: *e = 8; *w = 4; break;
    case 512: *e = 8; *w = 4; break;
    case 1024: *e = 8; *w = 4; break;
    case 2048: *e = prec == SGPE_C128 ? 8 : 16; *w = 4; break;
    default: *e = prec == SGPE_C128 ? 8 : 16; *w = prec == SGPE_C128 ? 2 : 4; break;
  }
}

static int set_lds_attr(const void* kernel, size_t bytes) {
  if (bytes > 64 * 1024) HIP_TRY(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return SGPE_OK;
}

static bool pow2_in_range(int n) { return n >= 64 && n <= 4096 && (n & (n - 1)) == 0; }

// Lengths of the runtime-length kernels (sgpe_generic.h): every even length
// up to kGenMaxLen. Powers of two from 64 to 4096 have compiled kernels; a
// power-of-two axis also takes the runtime-length kernels when no compiled
// tile of its pass divides the other axis.
static bool gen_length_ok(int n) { return n >= 2 && n <= kGenMaxLen && n % 2 == 0; }
static bool length_ok(int n) { return pow2_in_range(n) || gen_length_ok(n); }
static const char* const kLengthMsg = "grid sizes must be even and at most 2048, or powers of two up to 4096";

// Stage radices of an n-point runtime-length transform: 8s, then one 4 or
// 2, then the odd prime factors in increasing order.
static GenFft gen_factor(int n) {
  GenFft f;
  std::memset(&f, 0, sizeof(f));
  f.n = n;
  int m = n;
  auto push = [&](int r) {
    f.radix[f.nst++] = r;
    m /= r;
  };
  while (m % 8 == 0) push(8);
  if (m % 4 == 0) push(4);
  if (m % 2 == 0) push(2);
  for (int q = 3; m > 1; q += 2)
    while (m % q == 0) push(q);
  return f;
}

// LDS of the runtime-length kernels: two buffers of the workgroup's
// transforms plus 256 bytes for block sums.
static size_t gen_cols_lds(int w, int n, size_t cs) { return 2 * (size_t)w * n * cs + 256; }
static size_t gen_rows_lds(int rows, int n, size_t cs) { return 2 * (size_t)rows * n * cs + 256; }

// Column tile of the runtime-length column kernels: the widest of 8, 4, 2
// dividing nx whose buffers stay within 64 KiB, else 1.
static int gen_col_width(int n, int nx, size_t cs) {
  for (int w = 8; w >= 2; w /= 2)
    if (nx % w == 0 && gen_cols_lds(w, n, cs) <= 64 * 1024) return w;
  return 1;
}

// A compiled shape of pass `kind` for length n whose tile divides `other`:
// the given default if it does, else the entry with the same radix and the
// widest tile that does, else any radix. False: none.
static const KEntry* find_entry(int kind, int prec, int n, int e, int p);
static bool fit_shape(int kind, int prec, int n, int other, int* e, int* p) {
  if (!pow2_in_range(n)) return false;
  if (find_entry(kind, prec, n, *e, *p) && other % *p == 0) return true;
  const KTable t = prec == SGPE_C128 ? registry_f64(kind) : registry_f32(kind);
  const KEntry* best = nullptr;
  for (int pass = 0; pass < 2 && !best; ++pass)
    for (int i = 0; i < t.count; ++i) {
      const KEntry& k = t.entries[i];
      if (k.n != n || other % k.p || (pass == 0 && k.e != *e)) continue;
      if (!best || k.p > best->p) best = &k;
    }
  if (!best) return false;
  *e = best->e;
  *p = best->p;
  return true;
}

// Bound of one wait of the persistent kernel: 2 s of the 100 MHz wall clock
// (a pass takes microseconds to a few milliseconds).
static const long long kDefaultWaitTicks = 200000000LL;

static const PEntry* find_persist(int prec, int n, int re, int rr, int ce, int cw) {
  const PTable t = prec == SGPE_C128 ? persist_registry_f64() : persist_registry_f32();
  for (int i = 0; i < t.count; ++i) {
    const PEntry& e = t.entries[i];
    if (e.n == n && e.row_e == re && e.row_rows == rr && e.col_e == ce && e.col_w == cw) return &e;
  }
  return nullptr;
}

// Scratch of sgpe_norm_noalloc, one copy per device for the life of the
// process (16.6 KiB): the reduction partials (two per block, at most
// kNormBlocks blocks: the kernels stride over larger arrays, so the
// partials of any grid fit) and the norm factor. Host code never uses the
// symbol itself, only the device address hipGetSymbolAddress returns for
// each device (scratch_addr).
static constexpr int kNormBlocks = 1024;
static constexpr int kNormScratchDoubles = 2 * kNormBlocks + 32;
static __device__ __align__(256) double scratch[kNormScratchDoubles];

// Device address of `scratch` on `device`, looked up once per device (with
// that device current) and kept for the process.
static int scratch_addr(int device, double** out) {
  static std::mutex mu;
  static std::map<int, double*> addr;
  std::lock_guard<std::mutex> lock(mu);
  auto it = addr.find(device);
  if (it == addr.end()) {
    HIP_TRY(hipSetDevice(device));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(SGPE_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    it = addr.emplace(device, (double*)ptr).first;
  }
  *out = it->second;
  return SGPE_OK;
}

// ----------------------------------------------------------------- context

struct sgpe_ctx {
  int device = 0;
  int n_cu = 0;                      // compute units (residency check of persistent launches)
  hipStream_t stream = nullptr;
  hipStream_t cap_stream = nullptr;  // private stream used only for graph capture
  std::map<int, void*> tw;           // stage-major twiddles, key (prec*8192 + n)*32 + e
  double* red = nullptr;             // reduction scratch for primitives
  size_t red_count = 0;
  double* red_host = nullptr;        // pinned
  double* scratch_dev = nullptr;     // this device's address of `scratch` (sgpe_norm_noalloc)
};

// Stage-major twiddle table of an n-point transform done with e elements per
// thread (radices e, e, ..., remainder): the stage of span L and radix R keeps
// w_{L R}^{r k} (k < L, 1 <= r < R) at (L - 1) + k (R - 1) + (r - 1).
static int ctx_twiddles(sgpe_ctx* ctx, int prec, int n, int e, const void** out) {
  const int key = (prec * 8192 + n) * 32 + e;
  auto it = ctx->tw.find(key);
  if (it != ctx->tw.end()) {
    *out = it->second;
    return SGPE_OK;
  }
  std::vector<double> re, im;
  twiddles_f64(n, re, im);
  std::vector<double> sre(n, 0.0), sim(n, 0.0);
  for (int L = 1; L < n;) {
    const int R = (n / L >= e) ? e : n / L;
    for (int k = 0; k < L; ++k)
      for (int r = 1; r < R; ++r) {
        const int m = r * k * (n / (L * R));
        sre[(L - 1) + k * (R - 1) + (r - 1)] = re[m];
        sim[(L - 1) + k * (R - 1) + (r - 1)] = im[m];
      }
    L *= R;
  }
  std::vector<char> host(n * csize(prec));
  if (prec == SGPE_C128) {
    double* h = (double*)host.data();
    for (int i = 0; i < n; ++i) {
      h[2 * i] = sre[i];
      h[2 * i + 1] = sim[i];
    }
  } else {
    float* h = (float*)host.data();
    for (int i = 0; i < n; ++i) {
      h[2 * i] = (float)sre[i];
      h[2 * i + 1] = (float)sim[i];
    }
  }
  void* d = nullptr;
  HIP_TRY(hipSetDevice(ctx->device));
  if (dev_alloc(&d, host.size()) != hipSuccess) return fail(SGPE_E_NOMEM, "twiddle table allocation failed");
  HIP_TRY(hipMemcpy(d, host.data(), host.size(), hipMemcpyHostToDevice));
  ctx->tw[key] = d;
  *out = d;
  return SGPE_OK;
}

// Roots of unity root[m] = exp(-2 pi i m / n), m < n, of the runtime-length
// kernels (any even n): angles reduced exactly to [0, pi/4] (4m mod n and
// its complement are integers), long double cos/sin there, then the
// quarter-turn rotation; quarter turns come out exact.
static int ctx_roots(sgpe_ctx* ctx, int prec, int n, const void** out) {
  const int key = (prec * 8192 + n) * 32;  // e = 0: no compiled shape uses it
  auto it = ctx->tw.find(key);
  if (it != ctx->tw.end()) {
    *out = it->second;
    return SGPE_OK;
  }
  std::vector<char> host(n * csize(prec));
  const long double half_pi = 1.5707963267948966192313216916397514L;
  for (int m = 0; m < n; ++m) {
    const long long u = 4LL * m;      // angle = (pi / 2) u / n
    const int quad = (int)(u / n);    // whole quarter turns
    const long long rem = u - (long long)quad * n;
    long double c, sn;                // cos, sin of (pi / 2) rem / n
    if (2 * rem <= n) {
      const long double a = half_pi * (long double)rem / (long double)n;
      c = cosl(a);
      sn = sinl(a);
    } else {
      const long double a = half_pi * (long double)(n - rem) / (long double)n;
      c = sinl(a);
      sn = cosl(a);
    }
    for (int r = 0; r < quad; ++r) {  // rotate by pi / 2: (c, s) -> (-s, c)
      const long double t = c;
      c = -sn;
      sn = t;
    }
    const double re = (double)c, im = (double)-sn;  // exp(-i angle)
    if (prec == SGPE_C128) {
      ((double*)host.data())[2 * m] = re;
      ((double*)host.data())[2 * m + 1] = im;
    } else {
      ((float*)host.data())[2 * m] = (float)re;
      ((float*)host.data())[2 * m + 1] = (float)im;
    }
  }
  void* d = nullptr;
  HIP_TRY(hipSetDevice(ctx->device));
  if (dev_alloc(&d, host.size()) != hipSuccess) return fail(SGPE_E_NOMEM, "root table allocation failed");
  HIP_TRY(hipMemcpy(d, host.data(), host.size(), hipMemcpyHostToDevice));
  ctx->tw[key] = d;
  *out = d;
  return SGPE_OK;
}

// Dynamic LDS of a kernel with data_bytes of exchange buffers for n-point
// transforms: the twiddle table joins it when it fits (tw_in_lds), and the
// column pass adds its `tables` (the 2 kinetic y-tables) when those fit
// too; the same predicates as in the kernels.
static size_t lds_bytes(size_t data_bytes, int n, size_t cs, int tables = 0) {
  size_t total = data_bytes;
  if (tw_in_lds(data_bytes, n, cs)) {
    total += (size_t)n * cs;
    if (tables && data_bytes + (size_t)(1 + tables) * n * cs <= kLdsBytes) total += (size_t)tables * n * cs;
  }
  return std::max<size_t>(total, 256);
}
static const int kColTables = 2;

// Dynamic LDS of a column pass: exchange buffer, tables, and the separate
// slot for its end-of-pass sums when it fits (col_sum_separate).
static size_t col_lds_bytes(int w, int n, size_t cs, bool global_tables = false) {
  const size_t base = lds_bytes((size_t)w * n * cs, n, cs, global_tables ? 0 : kColTables);
  return base + (col_sum_separate(base) ? kColSumBytes : 0);
}

// Dynamic LDS of a row pass: exchange buffers and twiddles, plus a second
// exchange buffer when the kernel double-buffers (row_double_buffer).
static size_t row_lds_bytes(int e, size_t data_bytes, int n, size_t cs) {
  const size_t tw = tw_in_lds(data_bytes, n, cs) ? (size_t)n * cs : 0;
  return lds_bytes(data_bytes, n, cs) + (row_double_buffer(e, data_bytes, tw) ? data_bytes : 0);
}

// A row shape's rows per workgroup is negative for the split layout
// (row_split_kernel: two waves per row, one per component). Its threads per
// workgroup and dynamic LDS for either layout:
static int row_block_threads(int e, int rows, int n) { return rows < 0 ? -rows * 2 * (n / e) : rows * (n / e); }
static size_t row_pass_lds(int e, int rows, int n, size_t cs) {
  if (rows < 0) return lds_bytes((size_t)(-rows) * 2 * n * cs, n, cs);
  return row_lds_bytes(e, (size_t)rows * 2 * n * cs, n, cs);
}

static int ctx_reduction(sgpe_ctx* ctx, size_t count) {
  if (ctx->red_count >= count) return SGPE_OK;
  if (ctx->red) HIP_TRY(hipFree(ctx->red));
  ctx->red = nullptr;
  if (dev_alloc((void**)&ctx->red, count * sizeof(double)) != hipSuccess) return fail(SGPE_E_NOMEM, "scratch allocation failed");
  ctx->red_count = count;
  return SGPE_OK;
}

// -------------------------------------------------------------------- plan

struct Slot {
  void* buf = nullptr;  // tables, f64 copies, partials of this slot
  size_t bytes = 0;
  bool loaded = false;
};

struct sgpe_plan {
  sgpe_ctx* ctx = nullptr;
  int prec = 0, ny = 0, nx = 0, nslots = 0;
  int nlaunch = 0;  // slots the kernels are launched for: the first nlaunch (sgpe_plan_compact)
  int row_e = 8, row_rows = 1, col_e = 8, col_w = 4;
  int row_stream = 0, col_stream = 0;  // non-temporal state accesses per pass
  unsigned epoch = 0;                  // bumped whenever captured graphs go stale
  bool dense_kin = false;              // some slot has a dense kinetic array
  bool general_row = false;            // some slot has a dense potential or coupling array
  const KEntry* row = nullptr;
  const KEntry* col = nullptr;
  const void* tw_x = nullptr;  // twiddles for nx (rows)
  const void* tw_y = nullptr;  // twiddles for ny (columns)
  // Runtime-length kernels (sgpe_generic.h) of an axis without a compiled
  // pass for its length or tiles: gen_x for the row pass (x transforms, row_e
  // = 0, row_rows = 1), gen_y for the column pass (col_e = 0, col_w from
  // gen_col_width); their stage radices and root tables.
  bool gen_x = false, gen_y = false;
  // Column pass variant with its y phase tables in global memory (listed
  // with -W columns per block; col_w keeps the positive tile width).
  bool col_gky = false;
  GenFft gfx{}, gfy{};
  const void* root_x = nullptr;
  const void* root_y = nullptr;
  InstDev* d_inst = nullptr;
  RunDev* d_run = nullptr;
  std::vector<InstDev> h_inst;
  std::vector<Slot> slots;
  int last_dur = DUR_OUT;  // duration of the last sub-step issued (closing K pending)
  bool pending = false;    // a sub-step was issued since the last sync (active states are hybrid)
  std::map<int, hipGraphExec_t> graphs;
  std::map<int, hipGraph_t> graph_defs;
  bool energy_ready = false;  // energy kernels selected, their LDS set
  const KEntry* ecol = nullptr;
  const KEntry* erow = nullptr;
  int e_col_e = 8, e_col_w = 4, e_row_e = 8, e_row_rows = 1;
  const void* etw_x = nullptr;  // twiddles of the energy kernels' shapes
  const void* etw_y = nullptr;
  bool e_gen_x = false, e_gen_y = false;  // energy passes on the runtime-length kernels
  // Persistent step kernel (sgpe_persist.h): preference, the entry for the
  // current shapes (null: none compiled or not resident), its grid and LDS.
  bool persist = false;
  const PEntry* pk = nullptr;
  int pgrid = 0;
  size_t plds = 0;
  long long wait_ticks = kDefaultWaitTicks;
  PersistSync* d_sync = nullptr;  // [nslots]
  unsigned* d_pstatus = nullptr;  // [PST_WORDS]
};

static void drop_graphs(sgpe_plan* p) {
  for (auto& g : p->graphs) (void)hipGraphExecDestroy(g.second);
  for (auto& g : p->graph_defs) (void)hipGraphDestroy(g.second);
  p->graphs.clear();
  p->graph_defs.clear();
  ++p->epoch;
}

// Pass shapes of a plan. row_e = 0 (col_e = 0): the row (column) pass runs
// on the runtime-length kernels, row_rows (col_w) then ignored.
static int plan_select(sgpe_plan* p, int row_e, int row_rows, int col_e, int col_w) {
  const bool gx = row_e == 0, gy = col_e == 0;
  if ((gx && !gen_length_ok(p->nx)) || (gy && !gen_length_ok(p->ny)))
    return fail(SGPE_E_UNSUPPORTED, "no compiled FFT tile of this length divides the other grid axis");
  const size_t cs = csize(p->prec);
  if (gx) row_rows = 1;
  if (gy) col_w = gen_col_width(p->ny, p->nx, cs);
  const KEntry* r = gx ? nullptr : find_entry(K_ROW, p->prec, p->nx, row_e, row_rows);
  const KEntry* c = gy ? nullptr : find_entry(K_COL, p->prec, p->ny, col_e, col_w);
  if ((!gx && !r) || (!gy && !c)) return fail(SGPE_E_UNSUPPORTED, "FFT shape not compiled for this grid/precision");
  const bool gky = !gy && col_w < 0;
  col_w = std::abs(col_w);
  if (p->ny % std::abs(row_rows) || p->nx % col_w) return fail(SGPE_E_UNSUPPORTED, "tile does not divide the grid");
  const size_t row_data = (size_t)std::abs(row_rows) * 2 * p->nx * cs;
  const size_t col_data = (size_t)col_w * p->ny * cs;
  if (row_data > kMaxLds || col_data > kMaxLds) return fail(SGPE_E_UNSUPPORTED, "FFT shape exceeds LDS");
  const size_t rl = gx ? gen_rows_lds(2, p->nx, cs) : row_pass_lds(row_e, row_rows, p->nx, cs);
  const size_t cl = gy ? gen_cols_lds(col_w, p->ny, cs) : col_lds_bytes(col_w, p->ny, cs, gky);
  if (rl > kMaxLds || cl > kMaxLds) return fail(SGPE_E_UNSUPPORTED, "FFT shape exceeds LDS");
  int rc = SGPE_OK;
  const void *twx = nullptr, *twy = nullptr;
  if (gx) {
    rc = set_lds_attr(p->prec == SGPE_C128 ? (const void*)gen_row_pass_kernel<double>
                                           : (const void*)gen_row_pass_kernel<float>, rl);
    if (!rc) rc = ctx_roots(p->ctx, p->prec, p->nx, &twx);
  } else {
    rc = set_lds_attr(r->kernel, rl);
    if (!rc) rc = set_lds_attr(r->kernel_g, rl);
    if (!rc) rc = ctx_twiddles(p->ctx, p->prec, p->nx, row_e, &twx);
  }
  if (rc) return rc;
  if (gy) {
    rc = set_lds_attr(p->prec == SGPE_C128 ? (const void*)gen_col_pass_kernel<double>
                                           : (const void*)gen_col_pass_kernel<float>, cl);
    if (!rc) rc = ctx_roots(p->ctx, p->prec, p->ny, &twy);
  } else {
    rc = set_lds_attr(c->kernel, cl);
    if (!rc) rc = set_lds_attr(c->kernel_g, cl);
    if (!rc) rc = ctx_twiddles(p->ctx, p->prec, p->ny, col_e, &twy);
  }
  if (rc) return rc;
  p->gen_x = gx;
  p->gen_y = gy;
  p->gfx = gen_factor(p->nx);
  p->gfy = gen_factor(p->ny);
  p->tw_x = gx ? nullptr : twx;
  p->tw_y = gy ? nullptr : twy;
  p->root_x = gx ? twx : nullptr;
  p->root_y = gy ? twy : nullptr;
  p->row = r;
  p->col = c;
  p->row_e = gx ? 0 : row_e;
  p->row_rows = row_rows;
  p->col_e = gy ? 0 : col_e;
  p->col_w = col_w;
  p->col_gky = gky;
  // Persistent kernel of these shapes, if compiled and if every workgroup of
  // its grid can be resident at once (occupancy of the kernel x CUs).
  p->pk = nullptr;
  p->pgrid = 0;
  const PEntry* pe = !gx && !gy && !gky && row_rows > 0 && p->nx == p->ny && p->nslots <= kPersistMaxSlots
                         ? find_persist(p->prec, p->nx, row_e, row_rows, col_e, col_w)
                         : nullptr;
  if (pe) {
    const size_t lds = std::max(rl, cl);
    const int block = row_rows * (p->nx / row_e);
    int per_cu = 0;
    if ((rc = set_lds_attr(pe->kernel, lds))) return rc;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pe->kernel, block, lds));
    const int items = std::max(2 * (p->nx / col_w), p->ny / row_rows) * p->nslots;
    if (per_cu > 0 && p->ctx->n_cu > 0) {
      p->pk = pe;
      p->plds = lds;
      p->pgrid = std::min(per_cu * p->ctx->n_cu, items);
    }
  }
  drop_graphs(p);
  return SGPE_OK;
}

// State address handed to the pass kernels as an argument: for a plan of
// one loaded slot (the kernels then issue their state loads before reading
// the slot's run flags). Captured graphs hold it, so rebinding the state
// drops them.
static void* direct_state(const sgpe_plan* p) {
  return p->nslots == 1 && p->slots[0].loaded ? p->h_inst[0].psi : nullptr;
}

// Launch with start/stop events when given (profiling), else a plain launch
// (the form graph capture records).
#define SGPE_LAUNCH_EV(K, g, b, l, s, e0, e1, ...)                        \
  do {                                                                    \
    if (e0)                                                               \
      hipExtLaunchKernelGGL(K, g, b, (uint32_t)(l), s, e0, e1, 0, __VA_ARGS__); \
    else                                                                  \
      hipLaunchKernelGGL(K, g, b, l, s, __VA_ARGS__);                     \
  } while (0)

static void launch_col(sgpe_plan* p, hipStream_t s, int mode, int dprev, int dcur, hipEvent_t e0 = nullptr,
                       hipEvent_t e1 = nullptr) {
  if (p->gen_y) {
    const size_t lds = gen_cols_lds(p->col_w, p->ny, csize(p->prec));
    const dim3 grid(p->nx / p->col_w, 2, p->nlaunch), block(kGenThreads);
    if (p->prec == SGPE_C128)
      SGPE_LAUNCH_EV(gen_col_pass_kernel<double>, grid, block, lds, s, e0, e1, p->d_inst, p->d_run, p->gfy,
                     (const double2*)p->root_y, p->col_w, mode, dprev, dcur);
    else
      SGPE_LAUNCH_EV(gen_col_pass_kernel<float>, grid, block, lds, s, e0, e1, p->d_inst, p->d_run, p->gfy,
                     (const float2*)p->root_y, p->col_w, mode, dprev, dcur);
    return;
  }
  const size_t lds = col_lds_bytes(p->col_w, p->ny, csize(p->prec), p->col_gky);
  dim3 grid(p->nx / p->col_w, 2, p->nlaunch), block(p->col_w * (p->ny / p->col_e));
  ((ColFn)(p->dense_kin ? p->col->launch_g : p->col->launch))(grid, block, lds, s, e0, e1, p->d_inst, p->d_run,
                                                              direct_state(p), p->nx, p->tw_y, mode, dprev, dcur,
                                                              p->col_stream);
}

static void launch_row(sgpe_plan* p, hipStream_t s, int dur, int first, hipEvent_t e0 = nullptr,
                       hipEvent_t e1 = nullptr) {
  if (p->gen_x) {
    const size_t lds = gen_rows_lds(2, p->nx, csize(p->prec));
    const dim3 grid(p->ny, 1, p->nlaunch), block(kGenThreads);
    if (p->prec == SGPE_C128)
      SGPE_LAUNCH_EV(gen_row_pass_kernel<double>, grid, block, lds, s, e0, e1, p->d_inst, p->d_run, p->gfx,
                     (const double2*)p->root_x, dur, first, p->col_w);
    else
      SGPE_LAUNCH_EV(gen_row_pass_kernel<float>, grid, block, lds, s, e0, e1, p->d_inst, p->d_run, p->gfx,
                     (const float2*)p->root_x, dur, first, p->col_w);
    return;
  }
  const size_t lds = row_pass_lds(p->row_e, p->row_rows, p->nx, csize(p->prec));
  dim3 grid(p->ny / std::abs(p->row_rows), 1, p->nlaunch), block(row_block_threads(p->row_e, p->row_rows, p->nx));
  ((RowFn)(p->general_row ? p->row->launch_g : p->row->launch))(grid, block, lds, s, e0, e1, p->d_inst, p->d_run,
                                                                direct_state(p), p->ny, p->tw_x, dur, first, p->col_w,
                                                                p->row_stream);
}

static void issue_sub_step(sgpe_plan* p, hipStream_t s, int which, int first) {
  launch_col(p, s, COL_MID, p->last_dur, which);
  launch_row(p, s, which, first);
  p->last_dur = which;
  p->pending = true;
}

static void issue_full_step(sgpe_plan* p, hipStream_t s) {
  issue_sub_step(p, s, DUR_OUT, 1);
  issue_sub_step(p, s, DUR_IN, 0);
  issue_sub_step(p, s, DUR_OUT, 0);
}

// Whether n full steps run as one launch of the persistent kernel: asked
// for, compiled and resident for the current shapes, and every loaded slot
// on the fast operator path.
static bool use_persist(const sgpe_plan* p) { return p->persist && p->pk && !p->dense_kin && !p->general_row; }

// n full steps as one persistent launch, after a reset of its counters.
static void issue_persist(sgpe_plan* p, hipStream_t s, int n, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  hipLaunchKernelGGL(persist_reset_kernel, dim3(1), dim3(256), 0, s, p->d_sync, p->nslots, p->d_pstatus, 0);
  PersistArgs a;
  a.insts = p->d_inst;
  a.runs = p->d_run;
  a.direct_psi = direct_state(p);
  a.nslots = p->nlaunch;
  a.n_steps = n;
  a.d_first_prev = p->last_dur;
  a.col_stream = p->col_stream;
  a.row_stream = p->row_stream;
  a.tw_x = p->tw_x;
  a.tw_y = p->tw_y;
  a.sync = p->d_sync;
  a.status = p->d_pstatus;
  a.wait_ticks = p->wait_ticks;
  const int block = p->row_rows * (p->nx / p->row_e);
  ((PFn)p->pk->launch)(dim3(p->pgrid), dim3(block), p->plds, s, e0, e1, a);
  p->last_dur = DUR_OUT;
  p->pending = true;
}

static const char* const kAbortMsg =
    "a persistent step launch timed out waiting for another workgroup and stopped; the plan's states are "
    "invalid until they are set again (sgpe_plan_set_slot / sgpe_plan_set_state)";

// Clears the abort flag of the persistent kernel (new states bound).
static void clear_persist_abort(sgpe_plan* p, hipStream_t s) {
  hipLaunchKernelGGL(persist_reset_kernel, dim3(1), dim3(256), 0, s, p->d_sync, p->nslots, p->d_pstatus, 1);
}

// n full steps: one persistent launch, or six pass launches per step.
static void issue_steps(sgpe_plan* p, hipStream_t s, int n) {
  if (n <= 0) return;
  if (use_persist(p)) {
    issue_persist(p, s, n);
    return;
  }
  for (int i = 0; i < n; ++i) issue_full_step(p, s);
}

static void issue_sync(sgpe_plan* p, hipStream_t s) {
  launch_col(p, s, COL_EXIT, p->last_dur, p->last_dur);
  const size_t n = (size_t)2 * p->ny * p->nx;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 1024);
  dim3 grid(blocks, 1, p->nlaunch);
  if (p->prec == SGPE_C128)
    hipLaunchKernelGGL(normalize_kernel<double>, grid, dim3(256), 0, s, p->d_inst, p->d_run, p->col_w);
  else
    hipLaunchKernelGGL(normalize_kernel<float>, grid, dim3(256), 0, s, p->d_inst, p->d_run, p->col_w);
  hipLaunchKernelGGL(set_form_kernel, dim3((p->nslots + 63) / 64), dim3(64), 0, s, p->d_run, p->nslots,
                     (int)FORM_KSPACE);
  p->last_dur = DUR_OUT;
  p->pending = false;
}

struct sgpe_group {
  sgpe_ctx* ctx = nullptr;
  std::vector<sgpe_plan*> plans;
  std::vector<hipStream_t> side;  // one capture branch per plan
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr;
  std::map<long, hipGraphExec_t> graphs;
  std::map<long, hipGraph_t> defs;
  std::vector<unsigned> epochs;  // plan epochs the cached graphs were captured at
};

// Energy kernel shapes of a plan (once; must not run during graph capture).
// The evaluation works in place on the states: no work buffer.
static int plan_energy_prepare(sgpe_plan* p) {
  if (p->energy_ready) return SGPE_OK;
  const size_t cs = csize(p->prec);
  int rc;
  const bool f64 = p->prec == SGPE_C128;
  // Compiled default shapes where one divides the other axis, else the
  // runtime-length kernels.
  default_col(p->prec, p->ny, &p->e_col_e, &p->e_col_w);
  default_row(p->prec, p->nx, &p->e_row_e, &p->e_row_rows);
  p->e_gen_y = !fit_shape(K_ECOL, p->prec, p->ny, p->nx, &p->e_col_e, &p->e_col_w);
  p->e_gen_x = !fit_shape(K_EROW, p->prec, p->nx, p->ny, &p->e_row_e, &p->e_row_rows);
  if ((p->e_gen_y && !gen_length_ok(p->ny)) || (p->e_gen_x && !gen_length_ok(p->nx)))
    return fail(SGPE_E_UNSUPPORTED, "energy kernels not compiled for this grid");
  if (p->e_gen_y) {
    p->e_col_e = 0;
    p->e_col_w = gen_col_width(p->ny, p->nx, cs);
    const size_t cl = gen_cols_lds(p->e_col_w, p->ny, cs);
    if ((rc = set_lds_attr(f64 ? (const void*)gen_energy_col_kernel<double, false>
                               : (const void*)gen_energy_col_kernel<float, false>, cl)))
      return rc;
    if ((rc = set_lds_attr(f64 ? (const void*)gen_energy_col_kernel<double, true>
                               : (const void*)gen_energy_col_kernel<float, true>, cl)))
      return rc;
    if ((rc = ctx_roots(p->ctx, p->prec, p->ny, &p->etw_y))) return rc;
  } else {
    p->ecol = find_entry(K_ECOL, p->prec, p->ny, p->e_col_e, p->e_col_w);
    const size_t cl = lds_bytes((size_t)p->e_col_w * p->ny * cs, p->ny, cs);
    if ((rc = set_lds_attr(p->ecol->kernel, cl))) return rc;
    if ((rc = set_lds_attr(p->ecol->kernel_g, cl))) return rc;
    if ((rc = ctx_twiddles(p->ctx, p->prec, p->ny, p->e_col_e, &p->etw_y))) return rc;
  }
  if (p->e_gen_x) {
    p->e_row_e = 0;
    p->e_row_rows = 1;
    if ((rc = set_lds_attr(f64 ? (const void*)gen_energy_row_kernel<double> : (const void*)gen_energy_row_kernel<float>,
                           gen_rows_lds(2, p->nx, cs))))
      return rc;
    if ((rc = ctx_roots(p->ctx, p->prec, p->nx, &p->etw_x))) return rc;
  } else {
    p->erow = find_entry(K_EROW, p->prec, p->nx, p->e_row_e, p->e_row_rows);
    if ((rc = set_lds_attr(p->erow->kernel, lds_bytes((size_t)p->e_row_rows * 2 * p->nx * cs, p->nx, cs)))) return rc;
    if ((rc = ctx_twiddles(p->ctx, p->prec, p->nx, p->e_row_e, &p->etw_x))) return rc;
  }
  p->energy_ready = true;
  return SGPE_OK;
}

// Energy column kernel (FWD: the closing DFT_y that restores psik).
template <bool FWD> static void launch_energy_col(sgpe_plan* p, hipStream_t s) {
  const size_t cs = csize(p->prec);
  const dim3 cg(p->nx / p->e_col_w, 2, p->nlaunch);
  if (p->e_gen_y) {
    const size_t cl = gen_cols_lds(p->e_col_w, p->ny, cs);
    if (p->prec == SGPE_C128)
      hipLaunchKernelGGL((gen_energy_col_kernel<double, FWD>), cg, dim3(kGenThreads), cl, s, p->d_inst, p->d_run,
                         p->gfy, (const double2*)p->etw_y, p->e_col_w);
    else
      hipLaunchKernelGGL((gen_energy_col_kernel<float, FWD>), cg, dim3(kGenThreads), cl, s, p->d_inst, p->d_run,
                         p->gfy, (const float2*)p->etw_y, p->e_col_w);
    return;
  }
  const size_t cl = lds_bytes((size_t)p->e_col_w * p->ny * cs, p->ny, cs);
  const dim3 cb(p->e_col_w * (p->ny / p->e_col_e));
  ((EFn)(FWD ? p->ecol->launch_g : p->ecol->launch))(cg, cb, cl, s, p->d_inst, p->d_run, p->etw_y);
}

// Normalized k-space state, then the energy passes and the per-slot
// reduction / convergence test (all stream-ordered kernels, graph-safe).
static void issue_energy(sgpe_plan* p, hipStream_t s, double* out) {
  issue_sync(p, s);
  const size_t cs = csize(p->prec);
  launch_energy_col<false>(p, s);
  const dim3 rg(p->ny / p->e_row_rows, 1, p->nlaunch);
  if (p->e_gen_x) {
    const size_t rl = gen_rows_lds(2, p->nx, cs);
    if (p->prec == SGPE_C128)
      hipLaunchKernelGGL(gen_energy_row_kernel<double>, rg, dim3(kGenThreads), rl, s, p->d_inst, p->d_run, p->gfx,
                         (const double2*)p->etw_x);
    else
      hipLaunchKernelGGL(gen_energy_row_kernel<float>, rg, dim3(kGenThreads), rl, s, p->d_inst, p->d_run, p->gfx,
                         (const float2*)p->etw_x);
  } else {
    const size_t rl = lds_bytes((size_t)p->e_row_rows * 2 * p->nx * cs, p->nx, cs);
    const dim3 rb(p->e_row_rows * (p->nx / p->e_row_e));
    ((EFn)p->erow->launch)(rg, rb, rl, s, p->d_inst, p->d_run, p->etw_x);
  }
  launch_energy_col<true>(p, s);  // restores psik
  hipLaunchKernelGGL(energy_final_kernel, dim3(p->nlaunch), dim3(256), 0, s, p->d_inst, p->d_run,
                     p->nx / p->e_col_w, p->ny / p->e_row_rows, out);
}

// Host-side construction of one slot's InstDev and device buffer.
struct Staging {
  // Offset 0 marks an absent table, so real tables start after a header.
  std::vector<char> bytes = std::vector<char>(256);
  size_t add(size_t n, size_t align = 256) {
    size_t off = (bytes.size() + align - 1) / align * align;
    bytes.resize(off + n);
    return off;
  }
};

static void put_cplx(Staging& st, size_t off, int prec, int i, double re, double im) {
  if (prec == SGPE_C128) {
    double* d = (double*)(st.bytes.data() + off);
    d[2 * i] = re;
    d[2 * i + 1] = im;
  } else {
    float* d = (float*)(st.bytes.data() + off);
    d[2 * i] = (float)re;
    d[2 * i + 1] = (float)im;
  }
}

// exp(-i e Delta): real time Delta = dt, imaginary time Delta = -i dt.
static void phase_host(double e_dt, bool imag, double* re, double* im) {
  if (imag) {
    *re = std::exp(-e_dt);
    *im = 0.0;
  } else {
    *re = std::cos(e_dt);
    *im = -std::sin(e_dt);
  }
}

static int plan_load_slot(sgpe_plan* p, int slot, const sgpe_inst_desc* d, void* psik) {
  const int ny = p->ny, nx = p->nx, prec = p->prec;
  const size_t cs = csize(prec);
  if (d->ny != ny || d->nx != nx) return fail(SGPE_E_ARG, "descriptor grid does not match the plan");
  if (!(d->t_step > 0) || !(d->atom_num > 0) || !(d->dx > 0) || !(d->dy > 0))
    return fail(SGPE_E_ARG, "t_step, atom_num, dx, dy must be positive");
  if (d->time_mode != SGPE_TIME_IMAG && d->time_mode != SGPE_TIME_REAL) return fail(SGPE_E_ARG, "bad time_mode");
  if (d->coupling < 0 || d->coupling > 2) return fail(SGPE_E_ARG, "bad coupling mode");
  if (d->coupling == SGPE_COUPLING_ARRAY && !d->omega_dense) return fail(SGPE_E_ARG, "coupling array missing");
  for (int s = 0; s < 2; ++s) {
    if (!d->t_dense[s] && !(d->t_x[s] && d->t_y[s])) return fail(SGPE_E_ARG, "kinetic energy missing");
    if (!d->v_dense[s] && !(d->v_x[s] && d->v_y[s])) return fail(SGPE_E_ARG, "potential energy missing");
  }
  if (!psik) return fail(SGPE_E_ARG, "state pointer is NULL");

  const bool imag = d->time_mode == SGPE_TIME_IMAG;
  const double gamma = 1.0 / (2.0 + std::pow(2.0, 1.0 / 3.0));
  const double dts[2] = {d->t_step * gamma, d->t_step * (1.0 - 2.0 * gamma)};

  InstDev h;
  std::memset(&h, 0, sizeof(h));
  h.ny = ny;
  h.nx = nx;
  h.imag = imag ? 1 : 0;
  h.coupling = d->coupling;
  h.atom_num = d->atom_num;
  h.g_uu = d->g_uu;
  h.g_dd = d->g_dd;
  h.g_ud = d->g_ud;
  h.dv_r = d->dx * d->dy;
  h.dv_k = (M_PI / (0.5 * nx * d->dx)) * (M_PI / (0.5 * ny * d->dy));
  h.fft_c = d->dx * d->dy / (2.0 * M_PI);
  h.dt_sub[0] = dts[0];
  h.dt_sub[1] = dts[1];
  h.omega = d->coupling == SGPE_COUPLING_UNIFORM ? d->omega : 0.0;
  h.estride = 2 * nx + ny;

  Staging st;
  size_t o_kx[2][2], o_ky[2][2], o_px[2][2], o_py[2][2], o_eth = 0;
  size_t o_tx[2], o_ty[2], o_vx[2], o_vy[2], o_thx = 0, o_td[2], o_vd[2], o_om = 0;
  const size_t dense = (size_t)ny * nx * sizeof(double);
  for (int s = 0; s < 2; ++s) {
    o_tx[s] = o_ty[s] = o_vx[s] = o_vy[s] = o_td[s] = o_vd[s] = 0;
    if (d->t_dense[s]) {
      o_td[s] = st.add(dense);
      std::memcpy(st.bytes.data() + o_td[s], d->t_dense[s], dense);
    } else {
      o_tx[s] = st.add(nx * sizeof(double));
      std::memcpy(st.bytes.data() + o_tx[s], d->t_x[s], nx * sizeof(double));
      o_ty[s] = st.add(ny * sizeof(double));
      std::memcpy(st.bytes.data() + o_ty[s], d->t_y[s], ny * sizeof(double));
    }
    if (d->v_dense[s]) {
      o_vd[s] = st.add(dense);
      std::memcpy(st.bytes.data() + o_vd[s], d->v_dense[s], dense);
    } else {
      o_vx[s] = st.add(nx * sizeof(double));
      std::memcpy(st.bytes.data() + o_vx[s], d->v_x[s], nx * sizeof(double));
      o_vy[s] = st.add(ny * sizeof(double));
      std::memcpy(st.bytes.data() + o_vy[s], d->v_y[s], ny * sizeof(double));
    }
  }
  if (d->coupling == SGPE_COUPLING_ARRAY) {
    o_om = st.add(dense);
    std::memcpy(st.bytes.data() + o_om, d->omega_dense, dense);
  }
  if (d->theta_x) {
    o_thx = st.add(nx * sizeof(double));
    std::memcpy(st.bytes.data() + o_thx, d->theta_x, nx * sizeof(double));
  }
  // e^{-i theta(x)} for the coupling's off-diagonal; all ones without theta
  // (rotated frame), so the kernels apply it unconditionally.
  o_eth = st.add(nx * cs);
  for (int dd = 0; dd < 2; ++dd)
    for (int s = 0; s < 2; ++s) {
      o_kx[dd][s] = o_ky[dd][s] = o_px[dd][s] = o_py[dd][s] = 0;
      if (!d->t_dense[s]) {
        o_kx[dd][s] = st.add(nx * cs);
        o_ky[dd][s] = st.add(ny * cs);
      }
      if (!d->v_dense[s]) {
        o_px[dd][s] = st.add(nx * cs);
        o_py[dd][s] = st.add(ny * cs);
      }
    }
  const size_t o_part = st.add((size_t)PART_SLOTS * 2 * nx * sizeof(double));
  const size_t o_epart = st.add((size_t)EPART_SLOTS * (2 * nx + ny) * sizeof(double));

  // Phase tables in float64, rounded once to the plan precision.
  for (int dd = 0; dd < 2; ++dd) {
    const double dt = dts[dd];
    for (int s = 0; s < 2; ++s) {
      double re, im;
      if (!d->t_dense[s]) {
        for (int i = 0; i < nx; ++i) {
          phase_host(d->t_x[s][i] * (0.5 * dt), imag, &re, &im);
          put_cplx(st, o_kx[dd][s], prec, i, re, im);
        }
        for (int i = 0; i < ny; ++i) {
          phase_host(d->t_y[s][i] * (0.5 * dt), imag, &re, &im);
          put_cplx(st, o_ky[dd][s], prec, i, re, im);
        }
      }
      if (!d->v_dense[s]) {
        for (int i = 0; i < nx; ++i) {
          phase_host(d->v_x[s][i] * dt, imag, &re, &im);
          put_cplx(st, o_px[dd][s], prec, i, re, im);
        }
        for (int i = 0; i < ny; ++i) {
          phase_host(d->v_y[s][i] * dt, imag, &re, &im);
          put_cplx(st, o_py[dd][s], prec, i, re, im);
        }
      }
    }
    // Uniform coupling: angle a = Omega Delta/4 for both durations
    // (tensor_propagator.py:142-149 with tensor_tools.py:586).
    const double a = (d->coupling == SGPE_COUPLING_UNIFORM ? d->omega : 0.0) * dt * 0.25;
    if (imag) {
      h.c_alpha[dd] = std::cosh(a);
      h.c_beta_re[dd] = -std::sinh(a);
      h.c_beta_im[dd] = 0.0;
    } else {
      h.c_alpha[dd] = std::cos(a);
      h.c_beta_re[dd] = 0.0;
      h.c_beta_im[dd] = -std::sin(a);
    }
  }
  for (int i = 0; i < nx; ++i) {
    if (d->theta_x)
      put_cplx(st, o_eth, prec, i, std::cos(d->theta_x[i]), -std::sin(d->theta_x[i]));
    else
      put_cplx(st, o_eth, prec, i, 1.0, 0.0);
  }

  // A slot whose tables grow gets a new buffer; the old one is released
  // only once the new tables are in place, so a failed allocation leaves
  // the slot's previous physics, state and run untouched.
  Slot& sl = p->slots[slot];
  void* fresh = nullptr;
  if (sl.bytes < st.bytes.size() && dev_alloc(&fresh, st.bytes.size()) != hipSuccess)
    return fail(SGPE_E_NOMEM, "slot table allocation failed");
  struct FreshGuard {
    void*& ptr;
    ~FreshGuard() {
      if (ptr) (void)hipFree(ptr);
    }
  } guard{fresh};
  char* base = (char*)(fresh ? fresh : sl.buf);
  auto P = [&](size_t off) -> char* { return off ? base + off : nullptr; };
  for (int s = 0; s < 2; ++s) {
    h.t_dense[s] = (const double*)P(o_td[s]);
    h.v_dense[s] = (const double*)P(o_vd[s]);
    h.tx[s] = (const double*)P(o_tx[s]);
    h.ty[s] = (const double*)P(o_ty[s]);
    h.vx[s] = (const double*)P(o_vx[s]);
    h.vy[s] = (const double*)P(o_vy[s]);
    for (int dd = 0; dd < 2; ++dd) {
      h.kx[dd][s] = P(o_kx[dd][s]);
      h.ky[dd][s] = P(o_ky[dd][s]);
      h.px[dd][s] = P(o_px[dd][s]);
      h.py[dd][s] = P(o_py[dd][s]);
    }
  }
  h.omega_dense = (const double*)P(o_om);
  h.theta_x = (const double*)P(o_thx);
  h.etheta = P(o_eth);
  h.part = (double*)(base + o_part);
  h.epart = (double*)(base + o_epart);
  h.psi = psik;

  hipStream_t s = p->ctx->stream;
  HIP_TRY(hipMemcpyAsync(base, st.bytes.data(), st.bytes.size(), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(p->d_inst + slot, &h, sizeof(h), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(config_run_kernel, dim3(1), dim3(64), 0, s, p->d_run, slot, (double*)nullptr, 0, 1, 1,
                     (int)FORM_KSPACE, 0.0);
  clear_persist_abort(p, s);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s));  // staging buffer goes out of scope; old tables unused from here on
  if (fresh) {
    if (sl.buf) (void)hipFree(sl.buf);
    sl.buf = fresh;
    sl.bytes = st.bytes.size();
    fresh = nullptr;  // kept by the slot
  }
  p->h_inst[slot] = h;
  sl.loaded = true;
  // Kernel variants: the dense-operator ones once any loaded slot needs them.
  bool dk = false, gr = false;
  for (int i = 0; i < p->nslots; ++i) {
    if (!p->slots[i].loaded) continue;
    const InstDev& q = p->h_inst[i];
    dk = dk || q.t_dense[0] || q.t_dense[1];
    gr = gr || q.v_dense[0] || q.v_dense[1] || q.coupling == SGPE_COUPLING_ARRAY;
  }
  if (dk != p->dense_kin || gr != p->general_row || p->nslots == 1) {  // one slot: graphs hold its state address
    p->dense_kin = dk;
    p->general_row = gr;
    drop_graphs(p);
  }
  return SGPE_OK;
}

// A plan steps once any slot is set; slots never set keep active = 0 (the
// run state starts zeroed and only a set slot can be activated), so no
// kernel reads their empty descriptors.
static int plan_ready(const sgpe_plan* p) {
  if (!p) return fail(SGPE_E_ARG, "plan is NULL");
  for (int i = 0; i < p->nslots; ++i)
    if (p->slots[i].loaded) return SGPE_OK;
  return fail(SGPE_E_ARG, "no slot of the plan has been set");
}

// ------------------------------------------------------------------- C ABI

extern "C" {

int sgpe_abi_version(void) { return SGPE_ABI_VERSION; }

int sgpe_debug_fail_allocs(int n) {
  if (n < 0) return fail(SGPE_E_ARG, "n < 0");
  g_fail_allocs.store(n);
  return SGPE_OK;
}

const char* sgpe_last_error(void) { return g_err.c_str(); }

int sgpe_ctx_create(int device, void* stream, sgpe_ctx** out) {
  return guarded([&]() -> int {
    if (!out) return fail(SGPE_E_ARG, "out is NULL");
    *out = nullptr;
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(SGPE_E_ARG, "no such device");
    HIP_TRY(hipSetDevice(device));
    sgpe_ctx* c = new sgpe_ctx();
    c->device = device;
    if (hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) c->n_cu = 0;
    c->stream = (hipStream_t)stream;
    if (hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking) != hipSuccess) {
      delete c;
      return fail(SGPE_E_HIP, "capture stream creation failed");
    }
    if (hipHostMalloc((void**)&c->red_host, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
      (void)hipStreamDestroy(c->cap_stream);
      delete c;
      return fail(SGPE_E_HIP, "pinned buffer allocation failed");
    }
    const int rs = scratch_addr(device, &c->scratch_dev);
    if (rs) {
      (void)hipHostFree(c->red_host);
      (void)hipStreamDestroy(c->cap_stream);
      delete c;
      return rs;
    }
    *out = c;
    return SGPE_OK;
  });
}

int sgpe_ctx_set_stream(sgpe_ctx* ctx, void* stream) {
  return guarded([&]() -> int {
    if (!ctx) return fail(SGPE_E_ARG, "ctx is NULL");
    ctx->stream = (hipStream_t)stream;
    return SGPE_OK;
  });
}

int sgpe_ctx_synchronize(sgpe_ctx* ctx) {
  return guarded([&]() -> int {
    if (!ctx) return fail(SGPE_E_ARG, "ctx is NULL");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SGPE_OK;
  });
}

void sgpe_ctx_destroy(sgpe_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& t : ctx->tw) (void)hipFree(t.second);
  if (ctx->red) (void)hipFree(ctx->red);
  if (ctx->red_host) (void)hipHostFree(ctx->red_host);
  if (ctx->cap_stream) (void)hipStreamDestroy(ctx->cap_stream);
  delete ctx;
}

int sgpe_fft2(sgpe_ctx* ctx, int precision, int ny, int nx, int n_arrays, void* const* in, void* const* out,
              int inverse, double dx, double dy) {
  return guarded([&]() -> int {
    if (!ctx || !in || !out || n_arrays < 0) return fail(SGPE_E_ARG, "bad arguments");
    if (precision != SGPE_C128 && precision != SGPE_C64) return fail(SGPE_E_ARG, "bad precision");
    if (!length_ok(nx) || !length_ok(ny)) return fail(SGPE_E_UNSUPPORTED, kLengthMsg);
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t cs = csize(precision);
    const bool f64 = precision == SGPE_C128;
    int re, rr, ce, cw;
    default_row(precision, nx, &re, &rr);
    default_col(precision, ny, &ce, &cw);
    // Compiled shapes where one fits, else the runtime-length kernels.
    const bool gx = !fit_shape(K_FROW, precision, nx, ny, &re, &rr);
    const bool gy = !fit_shape(K_FCOL, precision, ny, nx, &ce, &cw);
    if ((gx && !gen_length_ok(nx)) || (gy && !gen_length_ok(ny)))
      return fail(SGPE_E_UNSUPPORTED, "no compiled FFT tile of this length divides the other grid axis");
    if (gy) cw = gen_col_width(ny, nx, cs);
    const KEntry* fr = gx ? nullptr : find_entry(K_FROW, precision, nx, re, rr);
    const KEntry* fc = gy ? nullptr : find_entry(K_FCOL, precision, ny, ce, cw);
    const size_t rl = gx ? gen_rows_lds(1, nx, cs) : lds_bytes((size_t)rr * nx * cs, nx, cs);
    const size_t cl = gy ? gen_cols_lds(cw, ny, cs) : lds_bytes((size_t)cw * ny * cs, ny, cs);
    const void* grk = f64 ? (const void*)gen_fft_rows_kernel<double> : (const void*)gen_fft_rows_kernel<float>;
    const void* gck = f64 ? (const void*)gen_fft_cols_kernel<double> : (const void*)gen_fft_cols_kernel<float>;
    int rc = set_lds_attr(gx ? grk : fr->kernel, rl);
    if (rc) return rc;
    if ((rc = set_lds_attr(gy ? gck : fc->kernel, cl))) return rc;
    const void *twx, *twy;
    if ((rc = gx ? ctx_roots(ctx, precision, nx, &twx) : ctx_twiddles(ctx, precision, nx, re, &twx))) return rc;
    if ((rc = gy ? ctx_roots(ctx, precision, ny, &twy) : ctx_twiddles(ctx, precision, ny, ce, &twy))) return rc;
    const GenFft gfx = gen_factor(nx), gfy = gen_factor(ny);
    const double c = dx * dy / (2.0 * M_PI);
    hipStream_t st = ctx->stream;
    // k arrays (1 or 2) of the pair io per launch, array blockIdx.z.
    auto rows = [&](const FftPair& io, int k, int inv, double scale) {
      if (!gx) {
        ((FRowFn)fr->launch)(dim3(ny / rr, 1, k), dim3(rr * (nx / re)), rl, st, io, inv, 1, scale, twx);
      } else if (f64) {
        hipLaunchKernelGGL(gen_fft_rows_kernel<double>, dim3(ny, 1, k), dim3(kGenThreads), rl, st, io, gfx,
                           (const double2*)twx, inv, scale);
      } else {
        hipLaunchKernelGGL(gen_fft_rows_kernel<float>, dim3(ny, 1, k), dim3(kGenThreads), rl, st, io, gfx,
                           (const float2*)twx, inv, scale);
      }
    };
    auto cols = [&](const FftPair& io, int k, int inv) {
      if (!gy) {
        ((FColFn)fc->launch)(dim3(nx / cw, 1, k), dim3(cw * (ny / ce)), cl, st, io, nx, inv, twy);
      } else if (f64) {
        hipLaunchKernelGGL(gen_fft_cols_kernel<double>, dim3(nx / cw, 1, k), dim3(kGenThreads), cl, st, io, gfy,
                           (const double2*)twy, nx, cw, inv);
      } else {
        hipLaunchKernelGGL(gen_fft_cols_kernel<float>, dim3(nx / cw, 1, k), dim3(kGenThreads), cl, st, io, gfy,
                           (const float2*)twy, nx, cw, inv);
      }
    };
    for (int i = 0; i < n_arrays; ++i)
      if (!in[i] || !out[i]) return fail(SGPE_E_ARG, "NULL array");
    // Byte ranges [p, p + bytes) of a and b overlap (compared as integers:
    // the arrays may belong to unrelated allocations).
    const uintptr_t bytes = (uintptr_t)ny * nx * cs;
    auto overlap = [bytes](const void* a, const void* b) {
      const uintptr_t pa = (uintptr_t)a, pb = (uintptr_t)b;
      return pa < pb + bytes && pb < pa + bytes;
    };
    for (int i = 0; i < n_arrays; ++i) {
      if (in[i] != out[i] && overlap(in[i], out[i]))
        return fail(SGPE_E_ARG, "an output partly overlaps its own input");
      // Arrays 2m and 2m + 1 are transformed in the same launch.
      const int j = i ^ 1;
      if (j < n_arrays && (overlap(out[i], out[j]) || overlap(out[i], in[j])))
        return fail(SGPE_E_ARG, "an output overlaps another array of its launch pair");
    }
    for (int i = 0; i < n_arrays; i += 2) {
      const int k = std::min(2, n_arrays - i);
      FftPair src{}, dst{};
      for (int j = 0; j < k; ++j) {
        src.in[j] = in[i + j];
        src.out[j] = out[i + j];
        dst.in[j] = out[i + j];
        dst.out[j] = out[i + j];
      }
      if (!inverse) {
        rows(src, k, 0, c);
        cols(dst, k, 0);
      } else {
        cols(src, k, 1);
        rows(dst, k, 1, 1.0 / ((double)nx * ny * c));
      }
    }
    HIP_TRY(hipGetLastError());
    return SGPE_OK;
  });
}

static int reduce_blocks(size_t n) { return (int)std::min<size_t>((n + 255) / 256, kNormBlocks); }

// norm() on the context's stream with the partials and norm factor in
// `part` (2 nb + 1 doubles, nb = reduce_blocks(ny nx)).
static int launch_norm(sgpe_ctx* ctx, int precision, int ny, int nx, void* psi0, void* psi1, double vol_elem,
                       double atom_num, void* dens0, void* dens1, double* nf_out, double* part) {
  const size_t n = (size_t)ny * nx;
  const int nb = reduce_blocks(n);
  double* nf_dev = part + 2 * nb;
  if (precision == SGPE_C128) {
    hipLaunchKernelGGL(density_sums_kernel<double>, dim3(nb), dim3(256), 0, ctx->stream, (const double2*)psi0,
                       (const double2*)psi1, n, part);
    hipLaunchKernelGGL(norm_apply_kernel<double>, dim3(nb), dim3(256), 0, ctx->stream, (double2*)psi0,
                       (double2*)psi1, n, (const double*)part, nb, vol_elem, atom_num, (double*)dens0,
                       (double*)dens1, nf_dev);
  } else {
    hipLaunchKernelGGL(density_sums_kernel<float>, dim3(nb), dim3(256), 0, ctx->stream, (const float2*)psi0,
                       (const float2*)psi1, n, part);
    hipLaunchKernelGGL(norm_apply_kernel<float>, dim3(nb), dim3(256), 0, ctx->stream, (float2*)psi0,
                       (float2*)psi1, n, (const double*)part, nb, vol_elem, atom_num, (float*)dens0,
                       (float*)dens1, nf_dev);
  }
  HIP_TRY(hipGetLastError());
  if (nf_out) {
    HIP_TRY(hipMemcpyAsync(ctx->red_host, nf_dev, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    *nf_out = ctx->red_host[0];
  }
  return SGPE_OK;
}

int sgpe_norm(sgpe_ctx* ctx, int precision, int ny, int nx, void* psi0, void* psi1, double vol_elem,
              double atom_num, void* dens0, void* dens1, double* nf_out) {
  return guarded([&]() -> int {
    if (!ctx || !psi0 || !psi1 || ny <= 0 || nx <= 0) return fail(SGPE_E_ARG, "bad arguments");
    if (precision != SGPE_C128 && precision != SGPE_C64) return fail(SGPE_E_ARG, "bad precision");
    HIP_TRY(hipSetDevice(ctx->device));
    const int nb = reduce_blocks((size_t)ny * nx);
    int rc = ctx_reduction(ctx, 2 * (size_t)nb + 1);
    if (rc) return rc;
    return launch_norm(ctx, precision, ny, nx, psi0, psi1, vol_elem, atom_num, dens0, dens1, nf_out, ctx->red);
  });
}

int sgpe_norm_noalloc(sgpe_ctx* ctx, int precision, int ny, int nx, void* psi0, void* psi1, double vol_elem,
                      double atom_num, void* dens0, void* dens1, double* nf_out) {
  return guarded([&]() -> int {
    if (!ctx || !psi0 || !psi1 || ny <= 0 || nx <= 0) return fail(SGPE_E_ARG, "bad arguments");
    if (precision != SGPE_C128 && precision != SGPE_C64) return fail(SGPE_E_ARG, "bad precision");
    const int nb = reduce_blocks((size_t)ny * nx);
    if (2 * (size_t)nb + 1 > (size_t)kNormScratchDoubles)
      return fail(SGPE_E_ARG, "sgpe_norm_noalloc: the reduction needs more than the 16.6 KiB scratch");
    if (!ctx->scratch_dev) return fail(SGPE_E_HIP, "sgpe_norm_noalloc: no scratch address for this device");
    HIP_TRY(hipSetDevice(ctx->device));
    return launch_norm(ctx, precision, ny, nx, psi0, psi1, vol_elem, atom_num, dens0, dens1, nf_out,
                       ctx->scratch_dev);
  });
}

int sgpe_pops(sgpe_ctx* ctx, int precision, int ny, int nx, const void* psi0, const void* psi1, double vol_elem,
              double* out2) {
  return guarded([&]() -> int {
    if (!ctx || !psi0 || !psi1 || !out2 || ny <= 0 || nx <= 0) return fail(SGPE_E_ARG, "bad arguments");
    if (precision != SGPE_C128 && precision != SGPE_C64) return fail(SGPE_E_ARG, "bad precision");
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t n = (size_t)ny * nx;
    const int nb = reduce_blocks(n);
    int rc = ctx_reduction(ctx, 2 * (size_t)nb + 2);
    if (rc) return rc;
    double* res = ctx->red + 2 * nb;
    if (precision == SGPE_C128)
      hipLaunchKernelGGL(density_sums_kernel<double>, dim3(nb), dim3(256), 0, ctx->stream, (const double2*)psi0,
                         (const double2*)psi1, n, ctx->red);
    else
      hipLaunchKernelGGL(density_sums_kernel<float>, dim3(nb), dim3(256), 0, ctx->stream, (const float2*)psi0,
                         (const float2*)psi1, n, ctx->red);
    hipLaunchKernelGGL(pops_final_kernel, dim3(1), dim3(256), 0, ctx->stream, (const double*)ctx->red, nb, vol_elem,
                       res);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(ctx->red_host, res, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    out2[0] = ctx->red_host[0];
    out2[1] = ctx->red_host[1];
    return SGPE_OK;
  });
}

int sgpe_plan_create(sgpe_ctx* ctx, int precision, int ny, int nx, int n_slots, sgpe_plan** out) {
  return guarded([&]() -> int {
    if (!ctx || !out || n_slots <= 0) return fail(SGPE_E_ARG, "bad arguments");
    *out = nullptr;
    if (precision != SGPE_C128 && precision != SGPE_C64) return fail(SGPE_E_ARG, "bad precision");
    if (!length_ok(nx) || !length_ok(ny)) return fail(SGPE_E_UNSUPPORTED, kLengthMsg);
    if (n_slots > 65535) return fail(SGPE_E_ARG, "too many slots");
    HIP_TRY(hipSetDevice(ctx->device));
    sgpe_plan* p = new sgpe_plan();
    p->ctx = ctx;
    p->prec = precision;
    p->ny = ny;
    p->nx = nx;
    p->nslots = n_slots;
    p->nlaunch = n_slots;
    p->row_stream = p->col_stream = n_slots == 1 ? 1 : 0;
    p->h_inst.resize(n_slots);
    p->slots.resize(n_slots);
    // Compiled default shapes where they fit the grid, else the runtime-length
    // kernels for that axis (e = 0).
    int re, rr, ce, cw;
    default_row(precision, nx, &re, &rr);
    default_col(precision, ny, &ce, &cw);
    if (!fit_shape(K_ROW, precision, nx, ny, &re, &rr)) re = 0;
    if (!fit_shape(K_COL, precision, ny, nx, &ce, &cw)) ce = 0;
    int rc = plan_select(p, re, rr, ce, cw);
    if (!rc && dev_alloc((void**)&p->d_inst, n_slots * sizeof(InstDev)) != hipSuccess)
      rc = fail(SGPE_E_NOMEM, "slot table allocation failed");
    if (!rc && dev_alloc((void**)&p->d_run, n_slots * sizeof(RunDev)) != hipSuccess)
      rc = fail(SGPE_E_NOMEM, "run state allocation failed");
    if (!rc && hipMemset(p->d_run, 0, n_slots * sizeof(RunDev)) != hipSuccess)
      rc = fail(SGPE_E_HIP, "run state initialization failed");
    if (!rc && hipMemset(p->d_inst, 0, n_slots * sizeof(InstDev)) != hipSuccess)
      rc = fail(SGPE_E_HIP, "slot table initialization failed");
    if (!rc && dev_alloc((void**)&p->d_sync, n_slots * sizeof(PersistSync)) != hipSuccess)
      rc = fail(SGPE_E_NOMEM, "persistent counters allocation failed");
    if (!rc && dev_alloc((void**)&p->d_pstatus, PST_WORDS * sizeof(unsigned)) != hipSuccess)
      rc = fail(SGPE_E_NOMEM, "persistent status allocation failed");
    if (!rc && hipMemset(p->d_pstatus, 0, PST_WORDS * sizeof(unsigned)) != hipSuccess)
      rc = fail(SGPE_E_HIP, "persistent status initialization failed");
    if (rc) {
      sgpe_plan_destroy(p);
      return rc;
    }
    *out = p;
    return SGPE_OK;
  });
}

void sgpe_plan_destroy(sgpe_plan* p) {
  if (!p) return;
  (void)hipSetDevice(p->ctx->device);
  (void)hipStreamSynchronize(p->ctx->stream);
  drop_graphs(p);
  for (auto& s : p->slots)
    if (s.buf) (void)hipFree(s.buf);
  if (p->d_inst) (void)hipFree(p->d_inst);
  if (p->d_run) (void)hipFree(p->d_run);
  if (p->d_sync) (void)hipFree(p->d_sync);
  if (p->d_pstatus) (void)hipFree(p->d_pstatus);
  delete p;
}

int sgpe_plan_set_slot(sgpe_plan* p, int slot, const sgpe_inst_desc* d, void* psik) {
  return guarded([&]() -> int {
    if (!p || !d) return fail(SGPE_E_ARG, "bad arguments");
    if (slot < 0 || slot >= p->nslots) return fail(SGPE_E_ARG, "slot out of range");
    if (slot >= p->nlaunch) return fail(SGPE_E_ARG, "slot is beyond the launched slots (sgpe_plan_compact)");
    HIP_TRY(hipSetDevice(p->ctx->device));
    return plan_load_slot(p, slot, d, psik);
  });
}

int sgpe_plan_set_state(sgpe_plan* p, int slot, void* psik) {
  return guarded([&]() -> int {
    if (!p || !psik || slot < 0 || slot >= p->nslots) return fail(SGPE_E_ARG, "bad arguments");
    if (!p->slots[slot].loaded) return fail(SGPE_E_ARG, "slot has not been set");
    HIP_TRY(hipSetDevice(p->ctx->device));
    InstDev& h = p->h_inst[slot];
    h.psi = psik;
    if (p->nslots == 1) drop_graphs(p);  // captured single-slot passes hold the state address
    HIP_TRY(hipMemcpyAsync(&p->d_inst[slot].psi, &h.psi, sizeof(void*), hipMemcpyHostToDevice, p->ctx->stream));
    hipLaunchKernelGGL(config_run_kernel, dim3(1), dim3(64), 0, p->ctx->stream, p->d_run, slot, (double*)nullptr, 0, 1, 1,
                       (int)FORM_KSPACE, 0.0);
    clear_persist_abort(p, p->ctx->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(p->ctx->stream));
    return SGPE_OK;
  });
}

int sgpe_plan_set_run(sgpe_plan* p, int slot, double* pops, int cap, double conv_tol) {
  return guarded([&]() -> int {
    if (!p || slot < 0 || slot >= p->nslots || cap < 0) return fail(SGPE_E_ARG, "bad arguments");
    if (!p->slots[slot].loaded) return fail(SGPE_E_ARG, "slot has not been set");
    HIP_TRY(hipSetDevice(p->ctx->device));
    hipLaunchKernelGGL(config_run_kernel, dim3(1), dim3(64), 0, p->ctx->stream, p->d_run, slot, pops, pops ? cap : 0, 1,
                       1, -1, conv_tol);
    HIP_TRY(hipGetLastError());
    return SGPE_OK;
  });
}

int sgpe_plan_status(sgpe_plan* p, int* out3, double* energy8) {
  return guarded([&]() -> int {
    if (!p || !out3) return fail(SGPE_E_ARG, "bad arguments");
    HIP_TRY(hipSetDevice(p->ctx->device));
    std::vector<RunDev> h(p->nslots);
    unsigned pst[PST_WORDS] = {0, 0};
    HIP_TRY(hipMemcpyAsync(h.data(), p->d_run, p->nslots * sizeof(RunDev), hipMemcpyDeviceToHost, p->ctx->stream));
    HIP_TRY(hipMemcpyAsync(pst, p->d_pstatus, sizeof(pst), hipMemcpyDeviceToHost, p->ctx->stream));
    HIP_TRY(hipStreamSynchronize(p->ctx->stream));
    if (pst[PST_ABORT]) return fail(SGPE_E_HIP, kAbortMsg);
    for (int i = 0; i < p->nslots; ++i) {
      out3[3 * i + 0] = h[i].active;
      out3[3 * i + 1] = h[i].converged;
      out3[3 * i + 2] = h[i].step;
      if (energy8)
        for (int j = 0; j < 8; ++j) energy8[8 * i + j] = h[i].energy[j];
    }
    return SGPE_OK;
  });
}

int sgpe_plan_set_active(sgpe_plan* p, int slot, int active) {
  return guarded([&]() -> int {
    if (!p || slot < 0 || slot >= p->nslots) return fail(SGPE_E_ARG, "bad arguments");
    if (active && !p->slots[slot].loaded) return fail(SGPE_E_ARG, "slot has not been set");
    if (active && slot >= p->nlaunch) return fail(SGPE_E_ARG, "slot is beyond the launched slots (sgpe_plan_compact)");
    HIP_TRY(hipSetDevice(p->ctx->device));
    // A disabled slot keeps its state as it is; finish the pending sub-step
    // first so that it is frozen as normalized k-space (on re-enabling, the
    // next step then starts from k-space, whatever sub-steps ran meanwhile).
    if (!active && p->pending) issue_sync(p, p->ctx->stream);
    hipLaunchKernelGGL(set_active_kernel, dim3(1), dim3(64), 0, p->ctx->stream, p->d_run, slot, active ? 1 : 0);
    HIP_TRY(hipGetLastError());
    return SGPE_OK;
  });
}

int sgpe_plan_compact(sgpe_plan* p, const int* order, int n_launch) {
  return guarded([&]() -> int {
    if (!p || !order) return fail(SGPE_E_ARG, "bad arguments");
    if (n_launch < 1 || n_launch > p->nslots) return fail(SGPE_E_ARG, "n_launch out of range");
    std::vector<char> seen(p->nslots, 0);
    for (int i = 0; i < p->nslots; ++i) {
      if (order[i] < 0 || order[i] >= p->nslots || seen[order[i]])
        return fail(SGPE_E_ARG, "order is not a permutation of the slots");
      seen[order[i]] = 1;
    }
    HIP_TRY(hipSetDevice(p->ctx->device));
    hipStream_t s = p->ctx->stream;
    std::vector<RunDev> run(p->nslots), moved(p->nslots);
    HIP_TRY(hipMemcpyAsync(run.data(), p->d_run, p->nslots * sizeof(RunDev), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int i = n_launch; i < p->nslots; ++i)
      if (run[order[i]].active) return fail(SGPE_E_ARG, "a slot left out of the launches is still active");
    std::vector<InstDev> inst(p->nslots);
    std::vector<Slot> slots(p->nslots);
    for (int i = 0; i < p->nslots; ++i) {
      moved[i] = run[order[i]];
      inst[i] = p->h_inst[order[i]];
      slots[i] = p->slots[order[i]];
    }
    HIP_TRY(hipMemcpyAsync(p->d_run, moved.data(), p->nslots * sizeof(RunDev), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(p->d_inst, inst.data(), p->nslots * sizeof(InstDev), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    p->h_inst.swap(inst);
    p->slots.swap(slots);
    p->nlaunch = n_launch;
    drop_graphs(p);
    return SGPE_OK;
  });
}

int sgpe_list_shapes(int pass, int precision, int n, int* out_pairs, int cap) {
  return guarded([&]() -> int {
    if ((pass != 0 && pass != 1) || (precision != SGPE_C128 && precision != SGPE_C64) || cap < 0)
      return fail(SGPE_E_ARG, "bad arguments");
    const KTable t = precision == SGPE_C128 ? registry_f64(pass == 0 ? K_ROW : K_COL)
                                            : registry_f32(pass == 0 ? K_ROW : K_COL);
    int count = 0;
    for (int i = 0; i < t.count; ++i) {
      if (t.entries[i].n != n) continue;
      if (out_pairs && count < cap) {
        out_pairs[2 * count] = t.entries[i].e;
        out_pairs[2 * count + 1] = t.entries[i].p;
      }
      ++count;
    }
    return count;
  });
}

int sgpe_plan_tune(sgpe_plan* p, int row_e, int row_rows, int col_e, int col_w) {
  return guarded([&]() -> int {
    if (!p) return fail(SGPE_E_ARG, "plan is NULL");
    HIP_TRY(hipSetDevice(p->ctx->device));
    return plan_select(p, row_e ? row_e : p->row_e, row_rows ? row_rows : p->row_rows, col_e ? col_e : p->col_e,
                       col_w ? col_w : (p->col_gky ? -p->col_w : p->col_w));
  });
}

int sgpe_plan_set_streaming(sgpe_plan* p, int row_stream, int col_stream) {
  return guarded([&]() -> int {
    if (!p) return fail(SGPE_E_ARG, "plan is NULL");
    p->row_stream = row_stream ? 1 : 0;
    p->col_stream = col_stream ? 1 : 0;
    drop_graphs(p);
    return SGPE_OK;
  });
}

int sgpe_plan_get_streaming(const sgpe_plan* p, int* out2) {
  return guarded([&]() -> int {
    if (!p || !out2) return fail(SGPE_E_ARG, "bad arguments");
    out2[0] = p->row_stream;
    out2[1] = p->col_stream;
    return SGPE_OK;
  });
}

int sgpe_list_persistent(int precision, int n, int* out4, int cap) {
  return guarded([&]() -> int {
    if ((precision != SGPE_C128 && precision != SGPE_C64) || cap < 0) return fail(SGPE_E_ARG, "bad arguments");
    const PTable t = precision == SGPE_C128 ? persist_registry_f64() : persist_registry_f32();
    int count = 0;
    for (int i = 0; i < t.count; ++i) {
      const PEntry& e = t.entries[i];
      if (e.n != n) continue;
      if (out4 && count < cap) {
        out4[4 * count + 0] = e.row_e;
        out4[4 * count + 1] = e.row_rows;
        out4[4 * count + 2] = e.col_e;
        out4[4 * count + 3] = e.col_w;
      }
      ++count;
    }
    return count;
  });
}

int sgpe_plan_set_persistent(sgpe_plan* p, int on, long long wait_ticks) {
  return guarded([&]() -> int {
    if (!p) return fail(SGPE_E_ARG, "plan is NULL");
    p->persist = on != 0;
    p->wait_ticks = wait_ticks > 0 ? wait_ticks : kDefaultWaitTicks;
    drop_graphs(p);
    return SGPE_OK;
  });
}

int sgpe_plan_get_persistent(const sgpe_plan* p, int* out3) {
  return guarded([&]() -> int {
    if (!p || !out3) return fail(SGPE_E_ARG, "bad arguments");
    out3[0] = p->persist ? 1 : 0;
    out3[1] = use_persist(p) ? 1 : 0;
    out3[2] = p->pk ? p->pgrid : 0;
    return SGPE_OK;
  });
}

int sgpe_plan_check(sgpe_plan* p) {
  return guarded([&]() -> int {
    if (!p) return fail(SGPE_E_ARG, "plan is NULL");
    HIP_TRY(hipSetDevice(p->ctx->device));
    unsigned pst[PST_WORDS] = {0, 0};
    HIP_TRY(hipMemcpyAsync(pst, p->d_pstatus, sizeof(pst), hipMemcpyDeviceToHost, p->ctx->stream));
    HIP_TRY(hipStreamSynchronize(p->ctx->stream));
    if (pst[PST_ABORT]) return fail(SGPE_E_HIP, kAbortMsg);
    return SGPE_OK;
  });
}

int sgpe_profile_persistent(sgpe_plan* p, int n_steps, double* ms_out) {
  return guarded([&]() -> int {
    int rc = plan_ready(p);
    if (rc) return rc;
    if (n_steps <= 0 || !ms_out) return fail(SGPE_E_ARG, "bad arguments");
    if (!use_persist(p)) return fail(SGPE_E_UNSUPPORTED, "the persistent path is off or not available for this plan");
    HIP_TRY(hipSetDevice(p->ctx->device));
    hipStream_t s = p->ctx->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
      if (e0) (void)hipEventDestroy(e0);
      return fail(SGPE_E_HIP, "event creation failed");
    }
    issue_persist(p, s, n_steps, e0, e1);
    hipError_t le = hipGetLastError();
    hipError_t se = hipStreamSynchronize(s);
    float ms = 0.f;
    hipError_t te = (le == hipSuccess && se == hipSuccess) ? hipEventElapsedTime(&ms, e0, e1) : hipSuccess;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (le != hipSuccess) return fail(SGPE_E_HIP, std::string("launch failed: ") + hipGetErrorString(le));
    if (se != hipSuccess) return fail(SGPE_E_HIP, std::string("synchronize failed: ") + hipGetErrorString(se));
    if (te != hipSuccess) return fail(SGPE_E_HIP, "event timing failed");
    *ms_out = ms;
    return SGPE_OK;
  });
}

int sgpe_plan_get_tuning(const sgpe_plan* p, int* out4) {
  return guarded([&]() -> int {
    if (!p || !out4) return fail(SGPE_E_ARG, "bad arguments");
    out4[0] = p->row_e;
    out4[1] = p->row_rows;
    out4[2] = p->col_e;
    out4[3] = p->col_gky ? -p->col_w : p->col_w;
    return SGPE_OK;
  });
}

int sgpe_single_step(sgpe_plan* p, int which) {
  return guarded([&]() -> int {
    int rc = plan_ready(p);
    if (rc) return rc;
    if (which != SGPE_SUB_OUTER && which != SGPE_SUB_INNER) return fail(SGPE_E_ARG, "bad sub-step");
    HIP_TRY(hipSetDevice(p->ctx->device));
    issue_sub_step(p, p->ctx->stream, which, 0);
    HIP_TRY(hipGetLastError());
    return SGPE_OK;
  });
}

int sgpe_full_steps(sgpe_plan* p, int n_steps) {
  return guarded([&]() -> int {
    int rc = plan_ready(p);
    if (rc) return rc;
    if (n_steps < 0) return fail(SGPE_E_ARG, "n_steps < 0");
    HIP_TRY(hipSetDevice(p->ctx->device));
    issue_steps(p, p->ctx->stream, n_steps);
    HIP_TRY(hipGetLastError());
    return SGPE_OK;
  });
}

int sgpe_graph_steps(sgpe_plan* p, int k, int n_graphs) {
  return guarded([&]() -> int {
    int rc = plan_ready(p);
    if (rc) return rc;
    if (k <= 0 || n_graphs < 0) return fail(SGPE_E_ARG, "bad step counts");
    HIP_TRY(hipSetDevice(p->ctx->device));
    if (n_graphs == 0) return SGPE_OK;
    // Captured blocks assume the pending closing half-step is an outer one.
    if (p->last_dur != DUR_OUT) issue_sync(p, p->ctx->stream);
    auto it = p->graphs.find(k);
    if (it == p->graphs.end()) {
      hipGraph_t g = nullptr;
      hipGraphExec_t ge = nullptr;
      hipStream_t cs = p->ctx->cap_stream;
      HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      issue_steps(p, cs, k);
      hipError_t le = hipGetLastError();
      hipError_t ce = hipStreamEndCapture(cs, &g);
      if (le != hipSuccess || ce != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(SGPE_E_HIP, std::string("graph capture failed: ") + hipGetErrorString(le != hipSuccess ? le : ce));
      }
      if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(SGPE_E_HIP, "graph instantiation failed");
      }
      p->graph_defs[k] = g;
      it = p->graphs.emplace(k, ge).first;
    }
    for (int i = 0; i < n_graphs; ++i) HIP_TRY(hipGraphLaunch(it->second, p->ctx->stream));
    p->last_dur = DUR_OUT;
    p->pending = true;
    return SGPE_OK;
  });
}

int sgpe_profile_full_steps(sgpe_plan* p, int n_steps, double* out4) {
  return guarded([&]() -> int {
    int rc = plan_ready(p);
    if (rc) return rc;
    if (n_steps <= 0 || !out4) return fail(SGPE_E_ARG, "bad arguments");
    HIP_TRY(hipSetDevice(p->ctx->device));
    hipStream_t s = p->ctx->stream;
    const int n_launch = 6 * n_steps;
    std::vector<hipEvent_t> ev(2 * n_launch, nullptr);
    for (auto& e : ev)
      if (hipEventCreate(&e) != hipSuccess) {
        for (auto& f : ev)
          if (f) (void)hipEventDestroy(f);
        return fail(SGPE_E_HIP, "event creation failed");
      }
    // The launches of sgpe_full_steps, each carrying a start/stop event pair
    // that timestamps the kernel itself (no extra packets between kernels).
    int k = 0;
    for (int i = 0; i < n_steps; ++i) {
      const int durs[3] = {DUR_OUT, DUR_IN, DUR_OUT};
      for (int j = 0; j < 3; ++j) {
        launch_col(p, s, COL_MID, p->last_dur, durs[j], ev[2 * k], ev[2 * k + 1]);
        ++k;
        launch_row(p, s, durs[j], j == 0, ev[2 * k], ev[2 * k + 1]);
        ++k;
        p->last_dur = durs[j];
        p->pending = true;
      }
    }
    hipError_t le = hipGetLastError();
    hipError_t se = hipStreamSynchronize(s);
    double col = 0.0, row = 0.0;
    for (int i = 0; i < n_launch && le == hipSuccess && se == hipSuccess; ++i) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]) != hipSuccess) break;
      if (i % 2 == 0)
        col += ms;
      else
        row += ms;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (le != hipSuccess) return fail(SGPE_E_HIP, std::string("launch failed: ") + hipGetErrorString(le));
    if (se != hipSuccess) return fail(SGPE_E_HIP, std::string("synchronize failed: ") + hipGetErrorString(se));
    out4[0] = col;
    out4[1] = row;
    out4[2] = 3.0 * n_steps;
    out4[3] = 3.0 * n_steps;
    return SGPE_OK;
  });
}

int sgpe_plan_sync(sgpe_plan* p) {
  return guarded([&]() -> int {
    int rc = plan_ready(p);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(p->ctx->device));
    issue_sync(p, p->ctx->stream);
    HIP_TRY(hipGetLastError());
    return SGPE_OK;
  });
}

int sgpe_energy(sgpe_plan* p, double* out) {
  return guarded([&]() -> int {
    int rc = plan_ready(p);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(p->ctx->device));
    if ((rc = plan_energy_prepare(p))) return rc;
    issue_energy(p, p->ctx->stream, out);
    HIP_TRY(hipGetLastError());
    return SGPE_OK;
  });
}

int sgpe_group_create(sgpe_ctx* ctx, int n_plans, sgpe_plan* const* plans, sgpe_group** out) {
  return guarded([&]() -> int {
    if (!ctx || !out || n_plans <= 0 || !plans) return fail(SGPE_E_ARG, "bad arguments");
    *out = nullptr;
    for (int i = 0; i < n_plans; ++i) {
      if (!plans[i] || plans[i]->ctx != ctx) return fail(SGPE_E_ARG, "plans must be non-NULL and share the context");
      for (int j = 0; j < i; ++j)
        if (plans[j] == plans[i]) return fail(SGPE_E_ARG, "a plan appears twice");
    }
    HIP_TRY(hipSetDevice(ctx->device));
    sgpe_group* g = new sgpe_group();
    g->ctx = ctx;
    g->plans.assign(plans, plans + n_plans);
    bool ok = hipEventCreateWithFlags(&g->fork, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; i < n_plans && ok; ++i) {
      hipStream_t st = nullptr;
      hipEvent_t ev = nullptr;
      ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
      if (ok) g->side.push_back(st);
      ok = ok && hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
      if (ok) g->join.push_back(ev);
    }
    if (!ok) {
      sgpe_group_destroy(g);
      return fail(SGPE_E_HIP, "group stream/event creation failed");
    }
    *out = g;
    return SGPE_OK;
  });
}

void sgpe_group_destroy(sgpe_group* g) {
  if (!g) return;
  (void)hipSetDevice(g->ctx->device);
  (void)hipStreamSynchronize(g->ctx->stream);
  for (auto& e : g->graphs) (void)hipGraphExecDestroy(e.second);
  for (auto& e : g->defs) (void)hipGraphDestroy(e.second);
  for (auto s : g->side) (void)hipStreamDestroy(s);
  for (auto e : g->join) (void)hipEventDestroy(e);
  if (g->fork) (void)hipEventDestroy(g->fork);
  delete g;
}

int sgpe_group_run(sgpe_group* g, int k, int check, int n_graphs) {
  return guarded([&]() -> int {
    if (!g || k <= 0 || n_graphs < 0) return fail(SGPE_E_ARG, "bad arguments");
    HIP_TRY(hipSetDevice(g->ctx->device));
    int rc;
    for (auto* p : g->plans) {
      if ((rc = plan_ready(p))) return rc;
      if (check && (rc = plan_energy_prepare(p))) return rc;
      if (p->last_dur != DUR_OUT) issue_sync(p, g->ctx->stream);
    }
    if (n_graphs == 0) return SGPE_OK;
    // A plan re-tuned, re-streamed or switched kernel variants since the
    // capture: its kernels in the cached graphs are stale.
    std::vector<unsigned> now;
    for (auto* p : g->plans) now.push_back(p->epoch);
    if (now != g->epochs) {
      for (auto& x : g->graphs) (void)hipGraphExecDestroy(x.second);
      for (auto& x : g->defs) (void)hipGraphDestroy(x.second);
      g->graphs.clear();
      g->defs.clear();
      g->epochs = now;
    }
    const long key = 2L * k + (check ? 1 : 0);
    auto it = g->graphs.find(key);
    if (it == g->graphs.end()) {
      // One capture: the context stream forks into one branch per plan, each
      // branch runs k full steps (and the energy/convergence check), and the
      // branches join back, so different grid sizes run concurrently.
      hipStream_t cs = g->ctx->cap_stream;
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      // Persistent launches each take the whole GPU (their workgroups must
      // all be resident), so the plans that use one run one after another
      // (a chain of branches); the others run concurrently.
      HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      hipError_t e = hipEventRecord(g->fork, cs);
      hipEvent_t chain = g->fork;
      for (size_t i = 0; i < g->plans.size() && e == hipSuccess; ++i) {
        const bool persistent = use_persist(g->plans[i]);
        e = hipStreamWaitEvent(g->side[i], persistent ? chain : g->fork, 0);
        if (e != hipSuccess) break;
        issue_steps(g->plans[i], g->side[i], k);
        if (check) issue_energy(g->plans[i], g->side[i], nullptr);
        e = hipEventRecord(g->join[i], g->side[i]);
        if (persistent) chain = g->join[i];
        if (e == hipSuccess) e = hipStreamWaitEvent(cs, g->join[i], 0);
      }
      hipError_t le = hipGetLastError();
      hipError_t ce = hipStreamEndCapture(cs, &graph);
      if (e != hipSuccess || le != hipSuccess || ce != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        hipError_t bad = e != hipSuccess ? e : (le != hipSuccess ? le : ce);
        return fail(SGPE_E_HIP, std::string("group graph capture failed: ") + hipGetErrorString(bad));
      }
      if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return fail(SGPE_E_HIP, "group graph instantiation failed");
      }
      g->defs[key] = graph;
      it = g->graphs.emplace(key, exec).first;
    }
    for (int i = 0; i < n_graphs; ++i) HIP_TRY(hipGraphLaunch(it->second, g->ctx->stream));
    // With the energy check the block ends on a sync (normalized k-space).
    for (auto* p : g->plans) {
      p->last_dur = DUR_OUT;
      p->pending = !check;
    }
    return SGPE_OK;
  });
}

}  // extern "C"
