/*
 * vpfp.h - C ABI of the MI355X (gfx950) Vlasov-Poisson-Fokker-Planck step.
 *
 * One library, libvpfp.so, built with hipcc for gfx950.  Every entry point
 * takes plain device pointers and sizes, enqueues HIP kernels on the given
 * stream (a hipStream_t passed as void*, NULL = default stream) and returns
 * an int status; nothing is synchronised unless stated.  The reference
 * (VlaPy, pure NumPy) binds these through ctypes; the Python bindings live in
 * vlapy_amd/_lib.py.
 *
 * Array conventions follow vlapy.core: f is (nx, nv) row-major with v
 * fastest, stored as double (VPFP_F64) or float (VPFP_F32); fields, moments,
 * grids and collision diagonals are always double.
 *
 * Status codes map to the reference's Python exceptions:
 *   VPFP_OK          0
 *   VPFP_EBADARG     1  -> ValueError
 *   VPFP_EUNSUPPORTED 2 -> NotImplementedError (name or size the GPU path
 *                          does not implement)
 *   VPFP_EHIP        3  -> RuntimeError (HIP runtime error)
 * vpfp_last_error() returns a message for the last failure on this thread.
 */
#ifndef VPFP_H
#define VPFP_H

#ifdef __cplusplus
extern "C" {
#endif

#define VPFP_OK 0
#define VPFP_EBADARG 1
#define VPFP_EUNSUPPORTED 2
#define VPFP_EHIP 3

#define VPFP_F64 0
#define VPFP_F32 1

#define VPFP_OP_LB 0 /* Lenard-Bernstein, collisions.py:26-83 */
#define VPFP_OP_DG 1 /* Dougherty, collisions.py:86-160 */

#define VPFP_NMOMENTS 8 /* n, j, T, q, fv4, vN, f^2, f log f (step.py:164-171,219-224) */

typedef struct vpfp_plan_s* vpfp_plan_t;

/* Grid and operator description; host arrays, copied at plan creation.
 * Replaces the captured `stuff_for_time_loop` of the reference factories
 * (outer_loop.py:125-144): kx, one_over_kx (nx), v, kv (nv), dv, nu, dt and
 * the collision operator name all_params["fokker-planck"]["type"]. */
typedef struct {
  int nx;
  int nv;
  int precision;              /* VPFP_F64 or VPFP_F32 (f storage and advection) */
  const double* kx;           /* [nx] fftfreq(nx, dx) * 2 pi */
  const double* one_over_kx;  /* [nx] 1/kx, 0 at DC */
  const double* v;            /* [nv] cell-centred velocities */
  const double* kv;           /* [nv] fftfreq(nv, dv) * 2 pi */
  double dv;
  double nu;                  /* collision frequency (>= 0) */
  double dt;                  /* time step used by vpfp_collide / vpfp_step_n */
  int op;                     /* VPFP_OP_LB or VPFP_OP_DG */
} vpfp_grid_desc;

/* Plan creation: copies the grid to the device, builds twiddle tables and
 * scratch.  Replaces the factory calls get_vdfdx / get_edfdv /
 * get_field_solver / get_collision_step (vlasov.py:213-260, field.py:91-107,
 * step.py:70-113).  Checks that kx and kv are fftfreq tables. */
int vpfp_plan_create(const vpfp_grid_desc* desc, vpfp_plan_t* out);
int vpfp_plan_destroy(vpfp_plan_t plan);

/* Delta-f storage for a plan (no counterpart in the reference, which stores
 * f itself; the batch form and its algorithm are described at
 * vpfp_batch_set_delta).  With on = 1 the f arrays of vpfp_step_n,
 * vpfp_edfdv, vpfp_vdfdx, vpfp_collide, vpfp_moments and vpfp_field hold
 * f - M(v); their results are those of the plain plan on f (moments are of f
 * itself).  Entry points whose output would change meaning
 * (vpfp_compute_charges, vpfp_collision_diagonals, vpfp_tridiag_solve,
 * vpfp_x_modes) return VPFP_EUNSUPPORTED while it is on.  Needs a
 * power-of-two nv.  vpfp_plan_background adds sign * M(v) (sign 1 or -1) to
 * every row of an (nx, nv) array. */
int vpfp_plan_set_delta(vpfp_plan_t plan, int on);
int vpfp_plan_background(vpfp_plan_t plan, void* f, int sign, void* stream);

/* x-advection f_out = Re ifft_x(exp(-i kx v dt) fft_x f_in).
 * Replaces step_vdfdx_exponential, vlasov.py:94-108.  f_out may equal f_in. */
int vpfp_vdfdx(vpfp_plan_t plan, const void* f_in, void* f_out, double dt, void* stream);

/* v-advection f_out = Re ifft_v(exp(-i kv e dt) fft_v f_in).
 * Replaces step_edfdv_exponential, vlasov.py:123-138.  e: [nx] device. */
int vpfp_edfdv(vpfp_plan_t plan, const void* f_in, void* f_out, const double* e, double dt, void* stream);

/* n(x) = trapz_v f.  Replaces field.compute_charges, field.py:27-36. */
int vpfp_compute_charges(vpfp_plan_t plan, const void* f, double* n_out, void* stream);

/* E = Re ifft(i one_over_kx fft(1 - n)).  Replaces field.solve_for_field,
 * field.py:52-63.  Any nx >= 2 (direct DFT when nx is not a power of two). */
int vpfp_solve_for_field(vpfp_plan_t plan, const double* charge_density, double* e_out, void* stream);

/* E = driver + solve_for_field(compute_charges(f)).  Replaces
 * solve_total_electric_field, field.py:77-86.  driver may be NULL (zero). */
int vpfp_field(vpfp_plan_t plan, const void* f, const double* driver, double* e_out, void* stream);

/* Allocation-free variant of vpfp_field: E = driver + solve_for_field(
 * compute_charges(f)) for f (nx, nv) with no plan and no device allocation.
 * No counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  one_over_kx [nx] (1/kx, 0 at DC), f
 * (double for VPFP_F64, float for VPFP_F32), driver [nx] (or NULL) and
 * e_out [nx] are device pointers on the current device; stream belongs to
 * that device.  Its scratch (density and twiddle table) is one file-scope
 * device array of VPFP_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends and shared by every call on that GPU: calls that overlap on one GPU
 * race.  nx > VPFP_NOALLOC_MAX_NX is refused with VPFP_EBADARG.  Power-of-two
 * nx >= 4 takes the FFT field kernel (one workgroup), any other nx the direct
 * DFT, as vpfp_field does; the twiddles are formed on the device (sincospi),
 * so results agree with vpfp_field to rounding, not bit for bit. */
#define VPFP_NOALLOC_MAX_NX 4096
#define VPFP_NOALLOC_SCRATCH_BYTES 98304
int vpfp_field_noalloc(int nx, int nv, int precision, double dv, const double* one_over_kx, const void* f,
                       const double* driver, double* e_out, void* stream);

/* Collision diagonals a [nx][nv-1], b [nx][nv], c [nx][nv-1] of f.
 * Replaces make_philharmonic_arrays_for_matrix / make_dougherty_arrays_for_matrix,
 * collisions.py:44-81, 104-158 (plan op, nu, dt, dv). */
int vpfp_collision_diagonals(vpfp_plan_t plan, const void* f, double* a, double* b, double* c, void* stream);

/* Batched tridiagonal solve x = A^-1 d for every row, A = diag(a,-1) +
 * diag(b) + diag(c,1).  Replaces _batched_tridiag_solver_, collisions.py:232-263
 * (and the per-x dense solve of collisions.py:172-217). x may equal d. */
int vpfp_tridiag_solve(vpfp_plan_t plan, const double* a, const double* b, const double* c, const void* d,
                       void* x, void* stream);

/* One implicit collision step with the plan's nu, dt and operator.
 * Replaces take_collision_step, step.py:80-108 (identity when nu == 0). */
int vpfp_collide(vpfp_plan_t plan, const void* f_in, void* f_out, void* stream);

/* n_steps leapfrog Vlasov-Poisson steps each followed by the collision
 * step, the fused inner loop of outer_loop.py:248-274 with step.py:302-326:
 *   kick(e, dt/2); drift(dt); E = driver_i + field(f); kick(E, dt/2);
 *   collide; moments.
 * f [nx][nv] and e [nx] are updated in place (e becomes the last E).
 * driver [n_steps][nx] (field added at step i) may be NULL.
 * e_hist [n_steps][nx], moments [n_steps][VPFP_NMOMENTS][nx] (trapz moments
 * of f after the step) and series [n_steps][2] (mean E^2, mean driver^2) are
 * optional outputs (NULL to skip). */
int vpfp_step_n(vpfp_plan_t plan, void* f, double* e, int n_steps, const double* driver, double* e_hist,
                double* moments, double* series, void* stream);

/* vpfp_step_n with the reference's per-step f store (get_f_update,
 * step.py:116-135, as the inner loop keeps it in temp_storage["stored_f"],
 * outer_loop.py:149-180): after the collision of step i, f_store[i] receives
 *   n_modes == 0: a copy of f, f_store [n_steps][nx][nv] in f's storage type
 *                 (written by the row kernel between the collision and the
 *                 next half kick, no extra launch);
 *   n_modes > 0:  the first n_modes Fourier modes along x of f, f_store
 *                 [n_steps][n_modes][nv] complex128, equal to vpfp_x_modes:
 *                 the row kernel writes f to plan scratch and extra
 *                 workgroups of the next step's field launch sum it over x
 *                 (one more launch after the last step).
 * Three launches per step either way.  Under delta-f storage the store holds
 * f itself. */
int vpfp_step_n_store(vpfp_plan_t plan, void* f, double* e, int n_steps, const double* driver, double* e_hist,
                      double* moments, double* series, void* f_store, int n_modes, void* stream);

/* n_steps of a higher-order splitting, each followed by the collision step,
 * as one device loop: the 4th-order PEFRL (VPFP_SCHEME_PEFRL, replaces
 * full_pefrl_ps_step, vlasov_poisson.py:63-154: five drifts each followed by a
 * field solve, four kicks) or the 6th-order splitting (VPFP_SCHEME_SIXTH,
 * replaces sixth_order_step, vlasov_poisson.py:157-232: kicks around five
 * drifts with their field solves), with the collision step of step.py:321-322
 * after each.  The last kick of a step, the collision, the moments, the f store
 * and the next step's first kick share one row-kernel launch.
 * driver [n_steps][VPFP_SCHEME_FIELDS][nx]: the driver at the step's field
 * solves in order (the reference's driver_function(t + ...) at each; NULL =
 * zero).  e [nx] is the field the 6th-order scheme's first kick uses and
 * becomes the last E.  e_hist [n_steps][nx] (E after each step), moments
 * [n_steps][VPFP_NMOMENTS][nx], series [n_steps][2] (mean E^2 and mean square
 * of the step's last driver) and f_store/n_modes as vpfp_step_n_store (the
 * x-modes by one launch per step) are optional (NULL / 0). */
#define VPFP_SCHEME_PEFRL 1
#define VPFP_SCHEME_SIXTH 2
#define VPFP_SCHEME_FIELDS 5
int vpfp_step_n_scheme(vpfp_plan_t plan, int scheme, void* f, double* e, int n_steps, const double* driver,
                       double* e_hist, double* moments, double* series, void* f_store, int n_modes, void* stream);

/* Trapz moments of f per x after the step: moments [VPFP_NMOMENTS][nx] =
 * n, j, T, q, fv4, vN (integrals of f v^p, p = 0..5) then f^2 and f log f.
 * Replaces update_fields / the f-integrals of update_series, step.py:153-224. */
int vpfp_moments(vpfp_plan_t plan, const void* f, double* moments, void* stream);

/* First n_modes Fourier modes along x of every v column, complex128
 * out [n_modes][nv] (interleaved re, im): sum_x f(x, v) exp(-2 pi i m x / nx).
 * Replaces the "k0..." f-storage rule, step.py:130-135 (the reference keeps
 * them as complex64). */
int vpfp_x_modes(vpfp_plan_t plan, const void* f, int n_modes, double* out, void* stream);

/* ---- Batched parameter scan (no counterpart in the reference, which runs
 * one instance per process).  A batch holds n_inst instances of one size
 * class (nx, nv) with their own box length (kx, one_over_kx per instance)
 * and collision frequency, sharing v, dt and the operator.  f is
 * [n_inst][nx][nv], e [n_inst][nx]; each step is the leapfrog step of
 * vpfp_step_n followed by the collision step, with no storage moments.
 * series [n_steps][n_inst][2] (optional) receives mean E^2 and 0. */
typedef struct vpfp_batch_s* vpfp_batch_t;
typedef struct {
  int n_inst;
  int nx;                     /* power of two in [4, 256] */
  int nv;                     /* power of two in [16, 8192] */
  int precision;              /* VPFP_F64 or VPFP_F32 */
  const double* kx;           /* [n_inst][nx] */
  const double* one_over_kx;  /* [n_inst][nx] */
  const double* v;            /* [nv] */
  const double* kv;           /* [nv] */
  double dv;
  const double* nu;           /* [n_inst] */
  double dt;
  int op;
} vpfp_batch_desc;
int vpfp_batch_create(const vpfp_batch_desc* desc, vpfp_batch_t* out);
int vpfp_batch_step_n(vpfp_batch_t batch, void* f, double* e, int n_steps, double* series, void* stream);
int vpfp_batch_destroy(vpfp_batch_t batch);

/* Several batches stepped together (no counterpart in the reference): the
 * size classes of one ragged scan.  A step is one row launch per distinct
 * nv (kick, collision, kick of every row of every batch with that nv), one
 * drift and one field launch per distinct nx, each launch taking a table of
 * the batches concerned (at most 8 batches per nv and per nx).  The batches
 * must share precision, collision operator, dt and delta-f storage, batches
 * of one nv the v grid, and none may have a closure.  f, e and series are
 * host arrays of device pointers, one per batch in the order given at
 * creation (series entries may be NULL, or series itself); each batch's
 * arrays are as in vpfp_batch_step_n, and the results equal those of
 * vpfp_batch_step_n per batch with the same launch shapes.  The group keeps
 * references to the batches: destroy it before them.
 * vpfp_group_autotune times the row shape of each nv and the drift tile of
 * each nx on scratch state it allocates and keeps the fastest;
 * vpfp_group_get_shapes reports them per batch (row values per thread, drift
 * columns per tile, tiles per drift workgroup, MFMA drift as in vpfp_tuning). */
typedef struct vpfp_group_s* vpfp_group_t;
int vpfp_group_create(vpfp_batch_t const* batches, int n_batches, vpfp_group_t* out);
int vpfp_group_step_n(vpfp_group_t group, void* const* f, double* const* e, int n_steps, double* const* series,
                      void* stream);
int vpfp_group_autotune(vpfp_group_t group, int n_steps, void* stream, double* us_per_step /* or NULL */);
int vpfp_group_get_shapes(vpfp_group_t group, int* row_q, int* drift_tile, int* drift_tiles, int* drift_dft);
/* set the shapes: each nv's row shape and each nx's drift tile are taken
 * from the first batch (in creation order) of that nv / nx */
int vpfp_group_set_shapes(vpfp_group_t group, const int* row_q, const int* drift_tile, const int* drift_tiles,
                          const int* drift_dft);
/* per-kernel timing as vpfp_plan_set_profiling; vpfp_group_launch_times
 * gives one launch of a step: kind 0 row (index = the nv's order of first
 * appearance), 1 drift, 2 field (index = the nx's order) */
int vpfp_group_set_profiling(vpfp_group_t group, int on);
int vpfp_group_kernel_times(vpfp_group_t group, double* ms /*[4]*/, long long* counts /*[4]*/);
int vpfp_group_launch_times(vpfp_group_t group, int kind, int index, double* ms, long long* count);
int vpfp_group_destroy(vpfp_group_t group);

/* Delta-f storage (no counterpart in the reference, which stores f itself).
 * With on = 1, the f arrays of vpfp_batch_step_n hold f - M(v), M the unit
 * Maxwellian exp(-v^2/2)/sqrt(2 pi) formed in FP32 on the device: the kicks
 * add M(v - a) - M(v) = M(v) expm1(a v - a^2/2) to the shifted deviation, the
 * collision and its moments see M + (f - M) in FP64, and the drift and the
 * field (whose k = 0 mode is dropped) see the deviation alone.  For FP32
 * storage this keeps the rounding of f relative to the deviation instead of
 * to M; the algorithm is unchanged.  Not with an attached closure.
 * vpfp_batch_background adds sign * M(v) (sign 1 or -1) to every row of a
 * [n_inst][nx][nv] array: -1 converts f to the stored deviation, 1 back. */
int vpfp_batch_set_delta(vpfp_batch_t batch, int on);
int vpfp_batch_background(vpfp_batch_t batch, void* f, int sign, void* stream);

/* Launch shapes of the step and the autotuner that picks them.  No
 * counterpart in the reference (its NumPy path has no launch shapes); the
 * choices change rounding only, never the algorithm.
 *   row_q       values per thread of the v-direction row kernels: 4
 *               (16 <= nv/2 <= 2048), 8, or 16 (nv/2 >= 256); 0 = by row
 *               count (4 when the grid has few rows, else 8)
 *   drift_tile  x-advection columns per workgroup: 4, 8, 16 or 32 (0 = keep)
 *   fuse_field  1: the last drift workgroup solves the field (nx <= 256);
 *               2: every row workgroup forms the field at its own rows from
 *               the drift's partial densities (direct DFT, nx <= 128);
 *               0 or -1 (default): separate field launch.  Batches always
 *               use the separate launch and report 0.
 *   drift_order 1: consecutive column tiles on the workgroups of one XCD
 *               (workgroup b runs on XCD b % 8); 0: tile = workgroup index
 *   drift_tiles column tiles per drift workgroup (1 = one each; a single
 *               grid never uses more than 1024 workgroups): fewer
 *               workgroups write fewer partial density rows (0 = keep)
 *   us_per_step measured time of the chosen shape (autotune; else 0)
 * vpfp_plan_autotune / vpfp_batch_autotune time n_steps steps of every valid
 * shape on scratch state the call allocates and frees (the caller's f and E
 * are not touched), keep the fastest, and synchronise the stream. */
typedef struct vpfp_tuning {
  int row_q;
  int drift_tile;
  int fuse_field;
  int drift_order;
  int drift_tiles;
  double us_per_step;
  int drift_dft;  /* 1: FP32 x-advection as dense real-DFT products on MFMA
                     (nx 32 or 64, drift_tile 16 or 32, not with fuse_field 1) */
} vpfp_tuning;

int vpfp_plan_get_tuning(vpfp_plan_t plan, vpfp_tuning* out);
int vpfp_plan_set_tuning(vpfp_plan_t plan, const vpfp_tuning* tuning);
int vpfp_plan_autotune(vpfp_plan_t plan, int n_steps, void* stream, vpfp_tuning* chosen /* or NULL */);
int vpfp_batch_get_tuning(vpfp_batch_t batch, vpfp_tuning* out);
int vpfp_batch_set_tuning(vpfp_batch_t batch, const vpfp_tuning* tuning);
int vpfp_batch_autotune(vpfp_batch_t batch, int n_steps, void* stream, vpfp_tuning* chosen /* or NULL */);

/* Instrumentation (no counterpart in the reference, which only logs wall
 * time per step to mlflow, manager.py:105-172).  With profiling on,
 * vpfp_step_n brackets each kernel launch with HIP events on its stream and
 * waits for them before returning; vpfp_plan_kernel_times returns the summed
 * milliseconds and launch counts per kernel class since profiling was
 * switched on: [0] row (kick/collide/moments), [1] drift (x-advection +
 * partial density), [2] field, [3] other. */
int vpfp_plan_set_profiling(vpfp_plan_t plan, int on);
int vpfp_plan_kernel_times(vpfp_plan_t plan, double* ms /*[4]*/, long long* counts /*[4]*/);
/* The same for a batch (vpfp_batch_step_n); [3] is the attached closure. */
int vpfp_batch_set_profiling(vpfp_batch_t batch, int on);
int vpfp_batch_kernel_times(vpfp_batch_t batch, double* ms /*[4]*/, long long* counts /*[4]*/);

const char* vpfp_last_error(void);
int vpfp_version(void);

#ifdef __cplusplus
}
#endif

#endif /* VPFP_H */
