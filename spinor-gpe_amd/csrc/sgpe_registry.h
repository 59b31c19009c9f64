// Registry of compiled kernel shapes. Each precision's instantiations live in
// their own translation unit (sgpe_inst_f64.hip, sgpe_inst_f32.hip) so that
// they compile in parallel; sgpe_api.hip looks shapes up here.
#pragma once

#include <hip/hip_ext.h>

#include "sgpe_kernels.h"

namespace sgpe {

// Split-step pass launchers; with non-null events the launch carries them
// (hipExtLaunchKernelGGL), which timestamps the kernel's own execution.
// The void*/int pair is the direct state address and length (null: read
// from the slot table).
typedef void (*ColFn)(dim3, dim3, size_t, hipStream_t, hipEvent_t, hipEvent_t, const InstDev*, RunDev*, void*, int,
                      const void*, int, int, int, int);
typedef void (*RowFn)(dim3, dim3, size_t, hipStream_t, hipEvent_t, hipEvent_t, const InstDev*, RunDev*, void*, int,
                      const void*, int, int, int, int);
typedef void (*FRowFn)(dim3, dim3, size_t, hipStream_t, FftPair, int, int, double, const void*);
typedef void (*FColFn)(dim3, dim3, size_t, hipStream_t, FftPair, int, int, const void*);
typedef void (*EFn)(dim3, dim3, size_t, hipStream_t, const InstDev*, const RunDev*, const void*);

struct KEntry {
  int prec, n, e, p;  // p = W (columns per block) or ROWS (rows per block; -ROWS: split row layout)
  void* launch;
  const void* kernel;
  void* launch_g;  // split-step passes: the variant for dense operators
  const void* kernel_g;
};

struct KTable {
  const KEntry* entries;
  int count;
};

// One table per kernel kind and precision: [kind][prec].
enum { K_COL = 0, K_ROW = 1, K_FROW = 2, K_FCOL = 3, K_ECOL = 4, K_EROW = 5, K_KINDS = 6 };
KTable registry_f64(int kind);
KTable registry_f32(int kind);

template <typename R, int N, int E, int W, bool G, bool GKY = false>
void col_launch(dim3 g, dim3 b, size_t l, hipStream_t s, hipEvent_t e0, hipEvent_t e1, const InstDev* i, RunDev* r,
                void* dpsi, int dn, const void* tw, int m, int dp, int dc, int st) {
  typedef typename Cx<R>::T C;
  if (e0)
    hipExtLaunchKernelGGL((col_pass_kernel<R, N, E, W, G, GKY>), g, b, (uint32_t)l, s, e0, e1, 0, i, r, dpsi, dn,
                          (const C*)tw, m, dp, dc, st);
  else
    hipLaunchKernelGGL((col_pass_kernel<R, N, E, W, G, GKY>), g, b, l, s, i, r, dpsi, dn, (const C*)tw, m, dp, dc,
                       st);
}
template <typename R, int N, int E, int ROWS, bool G>
void row_launch(dim3 g, dim3 b, size_t l, hipStream_t s, hipEvent_t e0, hipEvent_t e1, const InstDev* i, RunDev* r,
                void* dpsi, int dn, const void* tw, int dur, int first, int colw, int st) {
  typedef typename Cx<R>::T C;
  if (e0)
    hipExtLaunchKernelGGL((row_pass_kernel<R, N, E, ROWS, G>), g, b, (uint32_t)l, s, e0, e1, 0, i, r, dpsi, dn,
                          (const C*)tw, dur, first, colw, st);
  else
    hipLaunchKernelGGL((row_pass_kernel<R, N, E, ROWS, G>), g, b, l, s, i, r, dpsi, dn, (const C*)tw, dur, first, colw,
                       st);
}
template <typename R, int N, int E, int ROWS, bool G>
void rowsplit_launch(dim3 g, dim3 b, size_t l, hipStream_t s, hipEvent_t e0, hipEvent_t e1, const InstDev* i,
                     RunDev* r, void* dpsi, int dn, const void* tw, int dur, int first, int colw, int st) {
  typedef typename Cx<R>::T C;
  if (e0)
    hipExtLaunchKernelGGL((row_split_kernel<R, N, E, ROWS, G>), g, b, (uint32_t)l, s, e0, e1, 0, i, r, dpsi, dn,
                          (const C*)tw, dur, first, colw, st);
  else
    hipLaunchKernelGGL((row_split_kernel<R, N, E, ROWS, G>), g, b, l, s, i, r, dpsi, dn, (const C*)tw, dur, first,
                       colw, st);
}
template <typename R, int N, int E, int ROWS>
void frow_launch(dim3 g, dim3 b, size_t l, hipStream_t s, FftPair io, int inv, int shift, double scale,
                 const void* tw) {
  typedef typename Cx<R>::T C;
  hipLaunchKernelGGL((fft_rows_kernel<R, N, E, ROWS>), g, b, l, s, io, inv, shift, scale, (const C*)tw);
}
template <typename R, int N, int E, int W>
void fcol_launch(dim3 g, dim3 b, size_t l, hipStream_t s, FftPair io, int nx, int inv, const void* tw) {
  typedef typename Cx<R>::T C;
  hipLaunchKernelGGL((fft_cols_kernel<R, N, E, W>), g, b, l, s, io, nx, inv, (const C*)tw);
}
template <typename R, int N, int E, int W, bool FWD>
void ecol_launch(dim3 g, dim3 b, size_t l, hipStream_t s, const InstDev* i, const RunDev* r, const void* tw) {
  hipLaunchKernelGGL((energy_col_kernel<R, N, E, W, FWD>), g, b, l, s, i, r, (const typename Cx<R>::T*)tw);
}
template <typename R, int N, int E, int ROWS>
void erow_launch(dim3 g, dim3 b, size_t l, hipStream_t s, const InstDev* i, const RunDev* r, const void* tw) {
  hipLaunchKernelGGL((energy_row_kernel<R, N, E, ROWS>), g, b, l, s, i, r, (const typename Cx<R>::T*)tw);
}

}  // namespace sgpe

#define SGPE_COL_E(R, PR, N, E, W)                                                                   \
  {PR,                                                                                               \
   N,                                                                                                \
   E,                                                                                                \
   W,                                                                                                \
   (void*)&sgpe::col_launch<R, N, E, W, false>,                                                      \
   (const void*)&sgpe::col_pass_kernel<R, N, E, W, false>,                                           \
   (void*)&sgpe::col_launch<R, N, E, W, true>,                                                       \
   (const void*)&sgpe::col_pass_kernel<R, N, E, W, true>},
#define SGPE_ROW_E(R, PR, N, E, RW)                                                                  \
  {PR,                                                                                               \
   N,                                                                                                \
   E,                                                                                                \
   RW,                                                                                               \
   (void*)&sgpe::row_launch<R, N, E, RW, false>,                                                     \
   (const void*)&sgpe::row_pass_kernel<R, N, E, RW, false>,                                          \
   (void*)&sgpe::row_launch<R, N, E, RW, true>,                                                      \
   (const void*)&sgpe::row_pass_kernel<R, N, E, RW, true>},
// Column variant with the y phase tables in global memory, listed with
// columns per block -W.
#define SGPE_COLG_E(R, PR, N, E, W)                                                                  \
  {PR,                                                                                               \
   N,                                                                                                \
   E,                                                                                                \
   -(W),                                                                                             \
   (void*)&sgpe::col_launch<R, N, E, W, false, true>,                                                \
   (const void*)&sgpe::col_pass_kernel<R, N, E, W, false, true>,                                     \
   (void*)&sgpe::col_launch<R, N, E, W, true, true>,                                                 \
   (const void*)&sgpe::col_pass_kernel<R, N, E, W, true, true>},
#define SGPE_ROWSPLIT_E(R, PR, N, E, RW)                                                             \
  {PR,                                                                                               \
   N,                                                                                                \
   E,                                                                                                \
   -(RW),                                                                                            \
   (void*)&sgpe::rowsplit_launch<R, N, E, RW, false>,                                                \
   (const void*)&sgpe::row_split_kernel<R, N, E, RW, false>,                                         \
   (void*)&sgpe::rowsplit_launch<R, N, E, RW, true>,                                                 \
   (const void*)&sgpe::row_split_kernel<R, N, E, RW, true>},
#define SGPE_FROW_E(R, PR, N, E, RW)                                                                          \
  {PR, N, E, RW, (void*)&sgpe::frow_launch<R, N, E, RW>, (const void*)&sgpe::fft_rows_kernel<R, N, E, RW>, \
   nullptr, nullptr},
#define SGPE_FCOL_E(R, PR, N, E, W)                                                                          \
  {PR, N, E, W, (void*)&sgpe::fcol_launch<R, N, E, W>, (const void*)&sgpe::fft_cols_kernel<R, N, E, W>, \
   nullptr, nullptr},
// Energy column kernels: the summing inverse pass, and (launch_g/kernel_g)
// the restoring forward pass.
#define SGPE_ECOL_E(R, PR, N, E, W)                                                                              \
  {PR,                                                                                                           \
   N,                                                                                                            \
   E,                                                                                                            \
   W,                                                                                                            \
   (void*)&sgpe::ecol_launch<R, N, E, W, false>,                                                                 \
   (const void*)&sgpe::energy_col_kernel<R, N, E, W, false>,                                                     \
   (void*)&sgpe::ecol_launch<R, N, E, W, true>,                                                                  \
   (const void*)&sgpe::energy_col_kernel<R, N, E, W, true>},
#define SGPE_EROW_E(R, PR, N, E, RW)                                                                             \
  {PR, N, E, RW, (void*)&sgpe::erow_launch<R, N, E, RW>, (const void*)&sgpe::energy_row_kernel<R, N, E, RW>, \
   nullptr, nullptr},

// Tunable shapes of the two split-step passes: columns (N, E, W), rows (N, E, ROWS).
#define SGPE_COL_SHAPES(X, R, PR)                                                                          \
  X(R, PR, 64, 2, 4) X(R, PR, 64, 4, 4) X(R, PR, 64, 4, 8) X(R, PR, 64, 8, 8) X(R, PR, 64, 8, 16)           \
  X(R, PR, 128, 2, 4) X(R, PR, 128, 4, 4) X(R, PR, 128, 4, 8) X(R, PR, 128, 8, 4) X(R, PR, 128, 8, 8)       \
  X(R, PR, 128, 8, 16) X(R, PR, 256, 2, 2) X(R, PR, 256, 2, 4) X(R, PR, 256, 4, 4) X(R, PR, 256, 4, 8)      \
  X(R, PR, 256, 8, 4) X(R, PR, 256, 8, 8) X(R, PR, 256, 8, 16) X(R, PR, 256, 16, 8) X(R, PR, 512, 2, 2)      \
  X(R, PR, 512, 2, 4)                                                                                      \
  X(R, PR, 512, 4, 2) X(R, PR, 512, 4, 4) X(R, PR, 512, 8, 2) X(R, PR, 512, 8, 4) X(R, PR, 512, 8, 8)       \
  X(R, PR, 512, 8, 16)                                                                                     \
  X(R, PR, 512, 16, 4) X(R, PR, 512, 16, 8) X(R, PR, 1024, 4, 2) X(R, PR, 1024, 4, 4) X(R, PR, 1024, 8, 2)  \
  X(R, PR, 1024, 8, 4) X(R, PR, 1024, 16, 4) X(R, PR, 1024, 16, 8) X(R, PR, 2048, 4, 2) X(R, PR, 2048, 8, 2) \
  X(R, PR, 2048, 8, 4) X(R, PR, 2048, 16, 2) X(R, PR, 2048, 16, 4) X(R, PR, 4096, 8, 2) X(R, PR, 4096, 16, 2)
// Column shapes whose exchange buffer fits the LDS only in complex64
// (4 columns of 4096 complex128 points would need 256 KiB).
#define SGPE_COL_SHAPES_F32_ONLY(X, R, PR) X(R, PR, 4096, 16, 4)
#define SGPE_COL_SHAPES_F64_ONLY(X, R, PR)
#define SGPE_ROW_SHAPES(X, R, PR)                                                                          \
  X(R, PR, 64, 2, 2) X(R, PR, 64, 4, 4) X(R, PR, 64, 8, 8) X(R, PR, 64, 8, 16) X(R, PR, 128, 2, 1)          \
  X(R, PR, 128, 2, 2) X(R, PR, 128, 4, 2) X(R, PR, 128, 4, 4) X(R, PR, 128, 8, 4) X(R, PR, 128, 8, 8)       \
  X(R, PR, 256, 2, 1) X(R, PR, 256, 4, 1) X(R, PR, 256, 4, 2) X(R, PR, 256, 8, 2) X(R, PR, 256, 8, 4)       \
  X(R, PR, 256, 16, 4) X(R, PR, 512, 2, 1) X(R, PR, 512, 4, 1) X(R, PR, 512, 4, 2) X(R, PR, 512, 8, 1)      \
  X(R, PR, 512, 8, 2) X(R, PR, 512, 16, 2) X(R, PR, 512, 16, 4) X(R, PR, 1024, 4, 1) X(R, PR, 1024, 8, 1)   \
  X(R, PR, 1024, 16, 1) X(R, PR, 1024, 16, 2) X(R, PR, 2048, 4, 1) X(R, PR, 2048, 8, 1) X(R, PR, 2048, 16, 1) \
  X(R, PR, 4096, 8, 1) X(R, PR, 4096, 16, 1)

// Column shapes that also have the global-table variant (listed with -W):
// the batched shapes where staging the two y tables in LDS can cost a
// workgroup per CU.
#define SGPE_COLG_SHAPES(X, R, PR)                                                                     \
  X(R, PR, 128, 4, 8) X(R, PR, 128, 8, 8) X(R, PR, 128, 8, 16) X(R, PR, 256, 4, 8) X(R, PR, 256, 8, 8)   \
  X(R, PR, 256, 8, 16) X(R, PR, 512, 8, 4) X(R, PR, 512, 8, 8) X(R, PR, 512, 16, 8) X(R, PR, 1024, 8, 2) \
  X(R, PR, 1024, 8, 4) X(R, PR, 1024, 16, 4)

// Split row layout (row_split_kernel: one wave of N/E = 64 threads per
// component) with ROWS rows per workgroup; listed with rows = -ROWS.
#define SGPE_ROW_SPLIT_SHAPES(X, R, PR)                                                                \
  X(R, PR, 128, 2, 2) X(R, PR, 128, 2, 4) X(R, PR, 256, 4, 1) X(R, PR, 256, 4, 2) X(R, PR, 512, 8, 1)  \
  X(R, PR, 512, 8, 2) X(R, PR, 1024, 16, 1) X(R, PR, 1024, 16, 2)

// Default shapes (standalone FFTs and energy evaluation use only these);
// they must match default_row / default_col in sgpe_api.hip.
#define SGPE_DEF_ROWS(X, R, PR)                                                                     \
  X(R, PR, 64, 8, 8) X(R, PR, 128, 8, 4) X(R, PR, 256, 8, 2) X(R, PR, 512, 8, 1) X(R, PR, 1024, 8, 1) \
  X(R, PR, 2048, 8, 1) X(R, PR, 4096, 16, 1)
#define SGPE_DEF_COLS_F64(X, R, PR)                                                                 \
  X(R, PR, 64, 8, 8) X(R, PR, 128, 8, 8) X(R, PR, 256, 8, 4) X(R, PR, 512, 8, 4) X(R, PR, 1024, 8, 4) \
  X(R, PR, 2048, 8, 4) X(R, PR, 4096, 8, 2)
#define SGPE_DEF_COLS_F32(X, R, PR)                                                                 \
  X(R, PR, 64, 8, 8) X(R, PR, 128, 8, 8) X(R, PR, 256, 8, 4) X(R, PR, 512, 8, 4) X(R, PR, 1024, 8, 4) \
  X(R, PR, 2048, 16, 4) X(R, PR, 4096, 16, 4)

#define SGPE_DEFINE_REGISTRY(FN, R, PR, DEFCOLS, COLONLY)                                            \
  namespace sgpe {                                                                                   \
  static const KEntry k_col[] = {SGPE_COL_SHAPES(SGPE_COL_E, R, PR) COLONLY(SGPE_COL_E, R, PR)      \
                                 SGPE_COLG_SHAPES(SGPE_COLG_E, R, PR)};                              \
  static const KEntry k_row[] = {SGPE_ROW_SHAPES(SGPE_ROW_E, R, PR)                                  \
                                 SGPE_ROW_SPLIT_SHAPES(SGPE_ROWSPLIT_E, R, PR)};                     \
  static const KEntry k_frow[] = {SGPE_DEF_ROWS(SGPE_FROW_E, R, PR)};                                \
  static const KEntry k_fcol[] = {DEFCOLS(SGPE_FCOL_E, R, PR)};                                      \
  static const KEntry k_ecol[] = {DEFCOLS(SGPE_ECOL_E, R, PR)};                                      \
  static const KEntry k_erow[] = {SGPE_DEF_ROWS(SGPE_EROW_E, R, PR)};                                \
  KTable FN(int kind) {                                                                              \
    switch (kind) {                                                                                  \
      case K_COL: return {k_col, (int)(sizeof(k_col) / sizeof(KEntry))};                            \
      case K_ROW: return {k_row, (int)(sizeof(k_row) / sizeof(KEntry))};                            \
      case K_FROW: return {k_frow, (int)(sizeof(k_frow) / sizeof(KEntry))};                         \
      case K_FCOL: return {k_fcol, (int)(sizeof(k_fcol) / sizeof(KEntry))};                         \
      case K_ECOL: return {k_ecol, (int)(sizeof(k_ecol) / sizeof(KEntry))};                         \
      default: return {k_erow, (int)(sizeof(k_erow) / sizeof(KEntry))};                            \
    }                                                                                                \
  }                                                                                                  \
  }
