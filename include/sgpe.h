/*
 * sgpe.h - C ABI of libsgpe, the MI355X (gfx950) split-step engine for the
 * quasi-2D two-component Gross-Pitaevskii equation.
 *
 * The reference (ultracoldYEG/spinor-gpe) is pure Python on PyTorch; this
 * library replaces the torch calls below its TensorPropagator and
 * tensor_tools layers. Each entry point names the reference code it stands
 * in for (paths relative to the reference repository root). The Python
 * package spinor_gpe binds these symbols with ctypes.
 *
 * Conventions
 *  - Every int-returning function returns SGPE_OK (0) or a negative error
 *    code; sgpe_last_error() then describes the failure (thread-local).
 *    No C++ exception crosses this boundary.
 *  - Device pointers are HIP device addresses (e.g. torch.Tensor.data_ptr()
 *    of a tensor on device 'cuda'). Host pointers are only read during the
 *    call.
 *  - A state ("psik") is one contiguous [2][ny][nx] complex array
 *    (complex128 or complex64, interleaved re/im), component 0 first, in the
 *    reference's fft-shifted k-space layout (x along the fast axis).
 *  - All work is enqueued on the context's stream and is asynchronous unless
 *    stated otherwise.
 */
#ifndef SGPE_H
#define SGPE_H

#ifdef __cplusplus
extern "C" {
#endif

#define SGPE_ABI_VERSION 1

enum {
  SGPE_OK = 0,
  SGPE_E_ARG = -1,         /* invalid argument */
  SGPE_E_HIP = -2,         /* HIP runtime error */
  SGPE_E_NOMEM = -3,       /* device or host allocation failed */
  SGPE_E_UNSUPPORTED = -4  /* grid size / precision not supported */
};

enum { SGPE_C128 = 0, SGPE_C64 = 1 };            /* state precision */
enum { SGPE_TIME_IMAG = 0, SGPE_TIME_REAL = 1 };  /* tensor_propagator.py:96-99 */
enum { SGPE_COUPLING_OFF = 0, SGPE_COUPLING_UNIFORM = 1, SGPE_COUPLING_ARRAY = 2 };
enum { SGPE_SUB_OUTER = 0, SGPE_SUB_INNER = 1 };  /* tensor_propagator.py:101-103 */

typedef struct sgpe_ctx sgpe_ctx;
typedef struct sgpe_plan sgpe_plan;
typedef struct sgpe_group sgpe_group;

/*
 * Physical description of one instance: the inputs TensorPropagator.__init__
 * reads from a PSpinor (tensor_propagator.py:91-129). All pointers are HOST
 * float64 arrays, copied into the plan by sgpe_plan_set_slot.
 *
 * Kinetic energy T_s (kin_eng_spin) and potential V_s (pot_eng_spin) are each
 * given either separably, T_s[jy][jx] = t_x[s][jx] + t_y[s][jy], or densely
 * (t_dense[s], ny*nx row-major); a non-NULL dense pointer wins.
 * theta_x: the coupling phase expon = 2 kL x (:126-129), a function of x only;
 * NULL in the rotated frame (expon = 0).
 */
typedef struct sgpe_inst_desc {
  int ny, nx;
  int time_mode;           /* SGPE_TIME_* */
  double t_step;           /* full-step duration dt > 0 */
  double atom_num;
  double g_uu, g_dd, g_ud; /* scaled interaction strengths (spin.g_sc) */
  double dx, dy;           /* space['dr'] */
  const double* t_x[2];
  const double* t_y[2];
  const double* t_dense[2];
  const double* v_x[2];
  const double* v_y[2];
  const double* v_dense[2];
  int coupling;            /* SGPE_COUPLING_*; spin.is_coupling */
  double omega;            /* uniform coupling strength (SGPE_COUPLING_UNIFORM) */
  const double* omega_dense; /* ny*nx coupling array (SGPE_COUPLING_ARRAY) */
  const double* theta_x;   /* nx values or NULL */
} sgpe_inst_desc;

/* ABI version of the loaded library (compare with SGPE_ABI_VERSION). */
int sgpe_abi_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* sgpe_last_error(void);

/* Test hook (no reference counterpart): the next n device allocations of the
 * library fail as if the device were out of memory, so that tests can reach
 * the SGPE_E_NOMEM paths of the entry points. 0 restores normal behaviour. */
int sgpe_debug_fail_allocs(int n);

/* ---------------------------------------------------------------- context */

/* Bind a device and a stream (hipStream_t, NULL = default stream). Replaces
 * the reference's implicit torch device/stream (tensor_propagator.py:92). */
int sgpe_ctx_create(int device, void* stream, sgpe_ctx** out);
int sgpe_ctx_set_stream(sgpe_ctx* ctx, void* stream);
int sgpe_ctx_synchronize(sgpe_ctx* ctx);
void sgpe_ctx_destroy(sgpe_ctx* ctx);

/* ------------------------------------------------- tensor_tools primitives */

/* fft_2d / ifft_2d (spinor_gpe/pspinor/tensor_tools.py:201-228, 231-258) on
 * n_arrays device arrays of shape (ny, nx):
 *   inverse = 0: out = fftshift(fftn(in)) * dx*dy/(2 pi)
 *   inverse = 1: out = ifftn(ifftshift(in)) / (dx*dy/(2 pi))
 * in[i] may equal out[i]; otherwise out[i] must not overlap in[i], and
 * neither output of a launch pair (arrays 2m, 2m + 1 are transformed in one
 * launch) may overlap the other array's input or output: SGPE_E_ARG if so.
 * Grid lengths as for sgpe_plan_create. */
int sgpe_fft2(sgpe_ctx* ctx, int precision, int ny, int nx, int n_arrays, void* const* in, void* const* out,
              int inverse, double dx, double dy);

/* norm (tensor_tools.py:261-311, pop_frac=None): normalize the two
 * components to atom_num in place; optional dens0/dens1 (float64 for
 * SGPE_C128, float32 for SGPE_C64) receive the normalized densities;
 * nf_out (host, optional) receives the norm factor. Synchronizes the stream
 * when nf_out is given. */
int sgpe_norm(sgpe_ctx* ctx, int precision, int ny, int nx, void* psi0, void* psi1, double vol_elem,
              double atom_num, void* dens0, void* dens1, double* nf_out);

/* sgpe_norm without any allocation. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments and results as sgpe_norm; its reduction partials and norm factor
 * live in a static 16.6 KiB device array of the library, one per device (the
 * partials of any grid fit: the reduction strides over the grid with at most
 * 1024 blocks), so it calls no hipMalloc/hipFree. Every call on a device
 * shares that array: calls that overlap on one device (different streams or
 * threads) race and must be serialized by the caller. */
int sgpe_norm_noalloc(sgpe_ctx* ctx, int precision, int ny, int nx, void* psi0, void* psi1, double vol_elem,
                      double atom_num, void* dens0, void* dens1, double* nf_out);

/* calc_pops (tensor_tools.py:466-484): out2[s] = vol_elem * sum |psi_s|^2,
 * written to HOST memory (synchronizes the stream). */
int sgpe_pops(sgpe_ctx* ctx, int precision, int ny, int nx, const void* psi0, const void* psi1, double vol_elem,
              double* out2);

/* --------------------------------------------------------- propagation */

/* A plan of n_slots instances sharing one grid (ny, nx) and precision.
 * Stands in for TensorPropagator's operator set-up (tensor_propagator.py:
 * 67-149), batched. ny, nx: even lengths up to 2048 (the reference's
 * requirement, pspinor.py:331-332), and powers of two up to 4096. Powers of
 * two from 64 up run on compiled register-FFT passes; other lengths (and a
 * power-of-two axis no compiled tile of which divides the other axis) on
 * runtime-length LDS mixed-radix passes, slower but with the same results
 * to rounding. Other sizes: SGPE_E_UNSUPPORTED. Slots start empty
 * and inactive; a plan can step once at least one slot is set (empty slots
 * are skipped), so a server can fill slots as requests arrive. */
int sgpe_plan_create(sgpe_ctx* ctx, int precision, int ny, int nx, int n_slots, sgpe_plan** out);
void sgpe_plan_destroy(sgpe_plan* plan);

/* Load one slot's physics (host arrays are copied) and bind its device state
 * psik ([2][ny][nx]); the slot becomes active with a k-space state and a
 * fresh run (tensor_propagator.py:91-149). Synchronizes the stream. */
int sgpe_plan_set_slot(sgpe_plan* plan, int slot, const sgpe_inst_desc* desc, void* psik);

/* Rebind a slot to a (new) normalized k-space state psik and start a fresh
 * run; physics tables are kept. Synchronizes the stream. On a one-slot plan
 * the pass kernels take the state address as an argument, so its captured
 * step graphs are dropped and captured again on the next graph call. */
int sgpe_plan_set_state(sgpe_plan* plan, int slot, void* psik);

/* Start a run for one (set) slot and activate it: device buffer pops (2 doubles
 * per full step, capacity cap steps; NULL for none) receives the per-step
 * populations of prop_loop (tensor_propagator.py:173-194); conv_tol > 0
 * enables the convergence test of sgpe_energy, which deactivates the slot
 * once |E_n - E_prev| / |E_n| per full step <= conv_tol. */
int sgpe_plan_set_run(sgpe_plan* plan, int slot, double* pops, int cap, double conv_tol);

/* Per-slot run state, read back to HOST memory (synchronizes the stream):
 * out3[3*slot + {0,1,2}] = {active, converged, full steps in this run};
 * energy8 (optional) = the last sgpe_energy results per slot. Returns an
 * error if a persistent launch stopped (see sgpe_plan_set_persistent). */
int sgpe_plan_status(sgpe_plan* plan, int* out3, double* energy8);

/* Enable (1) or disable (0) a slot; disabled slots are skipped by every
 * kernel (their blocks exit at once). */
int sgpe_plan_set_active(sgpe_plan* plan, int slot, int active);

/* Reorder the slots (new slot i is old slot order[i], with its state binding,
 * tables and run state: step counts, populations buffer, energies,
 * convergence test) and launch the kernels for the first n_launch slots only;
 * every slot from n_launch on must be inactive, and keeps its state and
 * status. No counterpart in the reference (one instance per propagator): a
 * batched sweep whose slots converge one by one stops launching workgroups
 * for them, without rebuilding the plan. Slot indices change: callers map
 * their instances through `order`. Synchronizes the stream. */
int sgpe_plan_compact(sgpe_plan* plan, const int* order, int n_launch);

/* Select FFT shapes: elements per thread and rows per block for the row
 * pass, elements per thread and columns per block for the column pass.
 * 0 keeps the current value. A negative row_rows selects the split row
 * layout (two waves per row, one per component) with -row_rows rows per
 * block; a negative col_w the column pass that reads its y phase tables from
 * global memory instead of LDS, with -col_w columns per block. Returns
 * SGPE_E_UNSUPPORTED if not compiled. */
int sgpe_plan_tune(sgpe_plan* plan, int row_e, int row_rows, int col_e, int col_w);

/* Compiled shapes for one pass (0 = rows: {E, rows per block}, 1 = columns:
 * {E, columns per block}) of length n; writes up to cap pairs and returns
 * the number available (>= 0) or a negative error. Feeds the autotuner. */
int sgpe_list_shapes(int pass, int precision, int n, int* out_pairs, int cap);

/* Current shapes: out4 = {row_e, row_rows, col_e, col_w}, negative row_rows
 * and col_w as in sgpe_plan_tune; row_e = 0 (col_e = 0): that pass runs on
 * the runtime-length kernels. */
int sgpe_plan_get_tuning(const sgpe_plan* plan, int* out4);

/* Run blocks of full steps (sgpe_full_steps, sgpe_graph_steps, sgpe_group_run)
 * as ONE launch of a persistent kernel per block instead of six pass
 * launches per step (on = 1), for the loop of prop_loop (tensor_propagator.py:
 * 185-194). Used where a persistent kernel is compiled for the plan's
 * shapes (square grids), all its workgroups can be resident at once, and
 * every loaded slot has separable operators with no or uniform coupling;
 * elsewhere the pass launches are used. Results are those of the pass
 * launches. wait_ticks (<= 0: default, 2 s) bounds each wait of one
 * workgroup for another, in ticks of the 100 MHz device wall clock: past it
 * the launch stops (every workgroup returns) and sgpe_plan_status /
 * sgpe_plan_check report the error. */
int sgpe_plan_set_persistent(sgpe_plan* plan, int on, long long wait_ticks);

/* Shapes with a compiled persistent kernel for n x n grids: writes up to cap
 * quadruples {row_e, row_rows, col_e, col_w} and returns the number
 * available (>= 0) or a negative error. */
int sgpe_list_persistent(int precision, int n, int* out4, int cap);

/* out3 = {requested, used for the next block, grid size (0: no persistent
 * kernel for the current shapes)}. */
int sgpe_plan_get_persistent(const sgpe_plan* plan, int* out3);

/* Synchronizes the stream; returns an error if a persistent launch of the
 * plan stopped on its wait bound since the states were last set. */
int sgpe_plan_check(sgpe_plan* plan);

/* Measurement: n_steps full steps as one persistent launch whose execution
 * is timestamped by a hipEvent pair; *ms = its duration. Synchronizes. */
int sgpe_profile_persistent(sgpe_plan* plan, int n_steps, double* ms);

/* Cache policy of the two passes' state reads and writes: 1 = non-temporal
 * (streaming) accesses, 0 = normal. Default: streaming for single-slot plans,
 * normal for batched ones. No reference counterpart (launch configuration). */
int sgpe_plan_set_streaming(sgpe_plan* plan, int row_stream, int col_stream);
int sgpe_plan_get_streaming(const sgpe_plan* plan, int* out2);

/* single_step (tensor_propagator.py:224-271) for every active slot with
 * the outer (SGPE_SUB_OUTER) or inner (SGPE_SUB_INNER) sub-step duration and
 * operators. The closing kinetic half-step and renormalization are applied
 * lazily (by the next step or by sgpe_plan_sync). */
int sgpe_single_step(sgpe_plan* plan, int which);

/* n_steps x full_step (tensor_propagator.py:214-222), including the per-step
 * populations of the prop_loop body (:191-194) when a pops buffer is set. */
int sgpe_full_steps(sgpe_plan* plan, int n_steps);

/* Same as sgpe_full_steps(plan, steps_per_graph * n_graphs), issued as
 * n_graphs launches of one captured hipGraph of steps_per_graph steps. */
int sgpe_graph_steps(sgpe_plan* plan, int steps_per_graph, int n_graphs);

/* Instrumented form of sgpe_full_steps for measurement: the same launches,
 * issued eagerly on the context stream, each carrying a start/stop hipEvent
 * pair (hipExtLaunchKernelGGL) that timestamps the kernel's execution; then
 * synchronized. out4 (host) = {column-pass ms, row-pass ms, column launches,
 * row launches}. It serves the per-step timing methodology
 * of the reference's benchmark (spinor_gpe/benchmarks/benchmark_prop.py:
 * 69-79), which times prop.full_step() calls. */
int sgpe_profile_full_steps(sgpe_plan* plan, int n_steps, double* out4);

/* Bring every active slot's state to normalized k-space psik (closing
 * kinetic half-step + tensor_propagator.py:271) and record the last step's
 * populations. */
int sgpe_plan_sync(sgpe_plan* plan);

/* GP energy functional of each active slot's current state, in float64:
 * out[8*slot + i] (device buffer) = {E_tot, E_kin, E_pot, E_int, E_cpl, mu,
 * N, rel_change} with E the total energy (not per atom) and mu the chemical
 * potential. The reference computes its (different) end-of-run energy on the
 * host in NumPy (tensor_propagator.py:273-324); this is the engine's own
 * spectral evaluation, used for convergence and chemical potential. Brings
 * the state to normalized k-space first (as sgpe_plan_sync). */
int sgpe_energy(sgpe_plan* plan, double* out);

/* ---------------------------------------------------- batched sweeps */

/* A group of plans (e.g. one per grid size of a parameter sweep) advanced
 * together. No counterpart in the reference, which propagates one instance
 * per TensorPropagator (tensor_propagator.py:185-194); this batches many. */
int sgpe_group_create(sgpe_ctx* ctx, int n_plans, sgpe_plan* const* plans, sgpe_group** out);
void sgpe_group_destroy(sgpe_group* group);

/* n_graphs launches of one captured hipGraph in which every plan, on its own
 * concurrent branch, runs k full steps and, if check != 0, the energy
 * evaluation and convergence test of sgpe_energy (results kept in the run
 * state, see sgpe_plan_status). */
int sgpe_group_run(sgpe_group* group, int k, int check, int n_graphs);

/* ------------------------------------------------------------- host tools */

/* 2D phase unwrapping by reliability-sorted path following (Herraez et al.,
 * Appl. Opt. 41, 7437 (2002)), the published algorithm behind
 * skimage.restoration.unwrap_phase used by tensor_tools.phase_comp
 * (tensor_tools.py:528-531). Host arrays, ny*nx row-major, no wrap-around. */
int sgpe_unwrap_phase_2d(const double* wrapped, double* out, int ny, int nx);

#ifdef __cplusplus
}
#endif

#endif /* SGPE_H */
