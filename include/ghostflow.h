/*
 * ghostflow.h -- C ABI of the MI355X ghost-cell AB2/CN cylinder step (libghostflow.so).
 *
 * One handle owns the device state of one channel/cylinder case: the staggered
 * fields, the AB2/CN history, the ghost and operator tables, the separable
 * rectangle factors and the boundary-row (Schur) capacitance inverses.  All
 * arguments are plain host pointers and sizes; arrays are FP64 (double) or
 * int32, row-major, in the reference's staggered shapes:
 *     u (ny, nx+1), v (ny+1, nx), p (ny, nx).
 * Velocity arrays handed to the library are the concatenation [u.ravel(), v.ravel()]
 * ("vel layout", length ny*(nx+1) + (ny+1)*nx).
 *
 * Status codes: 0 ok, 1 invalid argument (ValueError), 2 solve/gate failure
 * (RuntimeError), 3 non-finite result (FloatingPointError), 4 HIP error.
 * A handle is not thread-safe; calls on one handle are serialised on its stream.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   mcf_step            GhostCellIntegrator.advance            solver/optimized_ghost_cell.py:236-278
 *                       + per-step body of MAC_CylinderFlow.solve  solver/main.py:271-434
 *   mcf_run             the step loop of MAC_CylinderFlow.solve  solver/main.py:271-434
 *   mcf_project         GhostCellIntegrator.project             solver/optimized_ghost_cell.py:218-229
 *   mcf_forces          GhostCellIntegrator.surface_forces      solver/optimized_ghost_cell.py:280-286
 *   mcf_rect_inverse    rectangle(...)[1] (the separable solve) solver/rectangular_inverse.py:48-53
 *   mcf_schur_solve     boundary_row_inverse(...)[0] / monitored_pressure_solver(...)[0]
 *                                                               solver/pressure_schur.py:16-72, 103-124
 *   mcf_pressure_solve  monitored_pressure_solver(...)[0] with either backend
 *                                                               solver/pressure_schur.py:16-72
 *   mcf_set_pressure_backend  pressure_solver='direct' | 'iterative' for the ghost path
 *                                                               solver/main.py:40-56, optimized_ghost_cell.py:109-111
 *   mcf_divergence_stats calculate_mac_divergence_cpp           solver/src/bindings.cpp:343-390
 *   mcf_set_* / mcf_get_state   construction of GhostCellIntegrator and its history attributes
 *                                                               solver/optimized_ghost_cell.py:62-150,
 *                                                               solver/main.py:547-566 (checkpoint restore)
 */
#ifndef GHOSTFLOW_H
#define GHOSTFLOW_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mcf_handle mcf_handle;

/* Rectangle / operator kinds. */
enum { MCF_PRESSURE = 0, MCF_U = 1, MCF_V = 2 };
/* Fields of the ghost tables. */
enum { MCF_FIELD_U = 0, MCF_FIELD_V = 1, MCF_FIELD_P = 2 };
/* Operators carrying override rows (rows that differ from the structured stencil). */
enum { MCF_OP_GRAD_U = 0, MCF_OP_GRAD_V = 1, MCF_OP_DIV = 2, MCF_OP_HELM_U = 3, MCF_OP_HELM_V = 4,
       MCF_OP_PRESSURE = 5, MCF_N_OPS = 6 };

/* Per-step diagnostics written by mcf_step / mcf_run (MCF_DIAG_LEN doubles per step). */
enum {
  MCF_D_TIME = 0, MCF_D_DT, MCF_D_CD, MCF_D_CL, MCF_D_CD_P, MCF_D_CD_V, MCF_D_CL_P, MCF_D_CL_V,
  MCF_D_FX, MCF_D_FY, MCF_D_MAXDIV, MCF_D_CFL, MCF_D_MOMENTUM_RES, MCF_D_COUPLED_RES,
  MCF_D_P_RES0, MCF_D_P_RES1, MCF_D_P_REL0, MCF_D_P_REL1, MCF_D_P_REFINE, MCF_D_P_TARGET0, MCF_D_P_TARGET1,
  MCF_D_P_ROUND0, MCF_D_P_ROUND1, MCF_D_P_BACKWARD0, MCF_D_P_BACKWARD1, MCF_D_H_REFINE,
  MCF_D_STATUS, MCF_D_RHS_MAX, MCF_D_P_SOLVES, MCF_D_PROBE0, MCF_D_PROBE1,
  MCF_D_SOLVE0 = 32,   /* MCF_SOLVE_LEN doubles per pressure solve of the step, in call order (up to 3) */
  MCF_DIAG_LEN = 64
};
/* Record of one monitored pressure solve (pressure_schur.py:64-69, `info['last_solve']`). */
enum {
  MCF_S_ITERATIONS = 0, MCF_S_RES_LINF, MCF_S_REL_L2, MCF_S_REFINE, MCF_S_TARGET, MCF_S_ROUND,
  MCF_S_BACKWARD, MCF_S_STATUS, MCF_S_SECONDS, MCF_S_SPARE,
  MCF_SOLVE_LEN = 10
};

/* Creation: grid, physics and the inlet profile (u at x=0, length ny). */
int mcf_create(int nx, int ny, double dx, double dy, double nu, double rho, mcf_handle** out);
void mcf_destroy(mcf_handle* h);
const char* mcf_last_error(const mcf_handle* h);
int mcf_device_name(char* buf, int len);

int mcf_set_inlet(mcf_handle* h, const double* u_inlet);
/* Transient inflow perturbation (solver/main.py:280-286): every following step first sets its inlet
   from its own end time t, u_inlet = base * (1 + shape * (1 - t / duration)) while t <= duration and
   base afterwards, with shape = perturb_amp * sin(2 pi y / Ly) (both ny values).  duration <= 0
   removes it (the inlet then stays as the last step or mcf_set_inlet left it). */
int mcf_set_inlet_perturbation(mcf_handle* h, const double* base, const double* shape, double duration);
/* Rows of the solid interval per array row: lo[j] <= i <= hi[j] are solid (hi < lo: none).
   u: ny rows over nx+1 columns; v: ny+1 rows over nx columns; p: ny rows over nx columns. */
int mcf_set_solid_intervals(mcf_handle* h, int field, const int* lo, const int* hi);
/* Ghost extension rows (optimized_ghost_cell.py:14-58): node[k] = flat node index of ghost k,
   ptr[n+1], col[] = flat index of an active node of the same field, w[] = weight. */
int mcf_set_ghost(mcf_handle* h, int field, int n, const int* node, const int* ptr, const int* col,
                  const double* w);
/* Override rows of an operator: row[k] = flat output index, CSR over the input array
   (vel layout for DIV/HELM, p for GRAD/PRESSURE).  Box = bounding box of the rows. */
int mcf_set_override(mcf_handle* h, int op, int n, const int* row, const int* ptr, const int* col,
                     const double* val);
/* Separable rectangle factor: operator = a I + b (lx + ly); V (yy*yy) orthonormal y-eigenbasis,
   lam (yy) its eigenvalues, lx given by lower/diag/upper (xx-1, xx, xx-1). */
int mcf_set_rectangle(mcf_handle* h, int kind, int yy, int xx, double a, double b, const double* V,
                      const double* lam, const double* lower, const double* diag, const double* upper);
/* Boundary-row capacitance (pressure_schur.py:85-101): rows (r, rectangle flat indices of the
   rows where the embedded operator differs from the rectangle), points (npts, rectangle flat
   indices of the columns those rows touch) and SC = S^{-1} C restricted to the points, dense
   r*npts row-major, with S = I + C R^{-1} E_rows the capacitance matrix. */
int mcf_set_schur(mcf_handle* h, int kind, int r, const int* rows, int npts, const int* points,
                  const double* SC);
/* Surface quadrature: n points; pressure stencil CSR over p, du/dn over u, dv/dn over v;
   normals (n*2: nx, ny) and weights (n*2). */
int mcf_set_surface(mcf_handle* h, int n, const int* pp, const int* pc, const double* pv,
                    const int* up, const int* uc, const double* uv, const int* vp, const int* vc,
                    const double* vv, const double* normals, const double* normal_weights,
                    double diameter, double u_ref);

/* Pressure probes (postprocessing/cylinder_observables.py:37-38 reconstructs the front and back
   stagnation pressures from p every step): n <= 2 CSR rows over p, evaluated on each step's output
   pressure and recorded in the diagnostics columns MCF_D_PROBE0 / MCF_D_PROBE1 (0 without probes).
   Call after mcf_set_surface; n = 0 removes them. */
int mcf_set_probes(mcf_handle* h, int n, const int* ptr, const int* col, const double* val);

/* State: u/v in vel layout (extended arrays, ghost values included), p (ny*nx, full array),
   history: prev_adv (vel layout q-array, zeros off the active faces; NULL = none),
   half_p (ny*nx q-array; NULL = none), prev_dt. */
int mcf_set_state(mcf_handle* h, const double* vel, const double* p, const double* prev_adv,
                  const double* half_p, double prev_dt, double time);
int mcf_get_state(mcf_handle* h, double* vel, double* p, double* prev_adv, double* half_p,
                  double* prev_dt, double* time, int* has_history);

/* One step; diag receives MCF_DIAG_LEN doubles (may be NULL). */
int mcf_step(mcf_handle* h, double dt, double* diag);
/* n steps at a fixed dt; diag receives n*MCF_DIAG_LEN doubles (may be NULL). */
int mcf_run(mcf_handle* h, int n, double dt, double* diag);

/* Projection of a vel-layout q-array (zeros off the active faces): out = q - dt G phi. */
int mcf_project(mcf_handle* h, const double* q, double dt, double* out, double* phi, double* info4);
/* Ghost-fill a vel-layout q-array in place (unpack: E q + offset). */
int mcf_extend(mcf_handle* h, const double* q, double* out);
/* Surface forces (Fx, Fy, Fx_p, Fy_p, Fx_v, Fy_v) of the given full arrays. */
int mcf_forces(mcf_handle* h, const double* vel, const double* p, double mu, double* out6);
/* Separable rectangle inverse of kind on a (yy, xx) right-hand side. */
int mcf_rect_inverse(mcf_handle* h, int kind, const double* rhs, double* out);
/* Allocation-free mcf_rect_inverse.  No counterpart in the reference: it is for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  Same arguments and the same results, bit for bit; it calls no hipMalloc or hipFree.
   The staged rectangle, its buffer table and its descriptor live in one fixed file-scope device
   array of 6 MiB per GPU (its address is looked up once per device by mcf_create), so a rectangle
   may hold at most 786,432 values (yy * xx); larger ones return status 1.  Every call on one GPU
   shares that array and it keeps the last call's contents: calls that overlap on one GPU (from
   several host threads, or on handles whose streams run concurrently) race, and the caller must
   serialise them. */
int mcf_rect_inverse_noalloc(mcf_handle* h, int kind, const double* rhs, double* out);
/* Boundary-row corrected solve of the embedded operator (values on the active nodes given as
   a q-array in the kind's full array layout); info4 = residual_linf, refinements, roundoff, target. */
int mcf_schur_solve(mcf_handle* h, int kind, const double* rhs, double* out, double* info4);
/* Pressure backends (pressure_schur.py:16-72):
   MCF_PB_DIRECT  boundary-row (Woodbury) corrected rectangle inverse, <= 2 refinements ('fft_boundary_schur');
   MCF_PB_GMRES   restarted GMRES on the pressure operator, left-preconditioned by the rectangle inverse,
                  with <= 2 refinement solves and the same residual gate ('gmres_rectangle').
   GMRES follows SciPy's gmres (the reference's call: restart 50, at most 100 restart cycles,
   rtol = atol = 1e-12, the adaptive inner tolerance on the preconditioned residual, stop on the true
   residual norm); its orthogonalisation is classical Gram-Schmidt applied twice.  restart <= 64,
   max_cycles >= 1, rtol >= 0, atol >= 0.  Steps of a handle on the GMRES backend run without step
   graphs or speculation and cannot join a batch. */
enum { MCF_PB_DIRECT = 0, MCF_PB_GMRES = 1 };
int mcf_set_pressure_backend(mcf_handle* h, int backend, int restart, int max_cycles, double rtol, double atol);
int mcf_get_pressure_backend(mcf_handle* h, int* backend);
/* One monitored pressure solve on the selected backend (the reference's `pressure_solve(values)`):
   rhs and out are p-arrays (ny, nx), values on the active cells (others ignored / returned 0).
   record receives MCF_SOLVE_LEN doubles (may be NULL).  Status 1 for a non-finite rhs, 2 when the
   solve fails its gate or GMRES does not converge (the record is still written). */
int mcf_pressure_solve(mcf_handle* h, const double* rhs, double* out, double* record);
/* Operators on packed-equivalent arrays (optimized_ghost_cell.py:186-216, 231-234):
   ADVECTION -N(q), DIFFUSION L q + diff_offset (vel -> vel), GRADIENT G p (p -> vel),
   DIVERGENCE D q + div_offset (vel -> p), HMATVEC q - dt/2 L q (vel -> vel). */
enum { MCF_APPLY_ADVECTION = 0, MCF_APPLY_DIFFUSION = 1, MCF_APPLY_GRADIENT = 2, MCF_APPLY_DIVERGENCE = 3,
       MCF_APPLY_HMATVEC = 4 };
int mcf_apply_operator(mcf_handle* h, int op, const double* in, double* out, double dt);
/* max|div|, sum div^2, count, core sum div^2, core count over cells with mask==0. */
int mcf_divergence_stats(mcf_handle* h, const double* vel, const unsigned char* mask, double* out5);

/* Time the dominant kernels with HIP events on the library's stream (state advances): out =
   [ms per launch of the pressure problem's forward transform over 50 back-to-back launches,
    ms per launch of its tridiagonal sweep (50 back to back), ms per step over n lean steps at dt,
    ms per launch of the forward transform over 50 launches each bracketed by its own events and
    completed before the next, count of the n steps that would have refined]. */
int mcf_profile(mcf_handle* h, int n, double dt, double* out5);

/* The capacitance (Schur) system S y = c of the boundary-row correction (pressure_schur.py:92-111:
   S = I + C R^{-1} E_rows, c = C R^{-1} b) in boundary-angle order, for the pair dump and the
   iterative solve.  S: r x r in the row order of mcf_set_schur; C: CSR over the capacitance points
   (columns index the points array of mcf_set_schur); angle_order[t] = row at angle position t. */
int mcf_set_schur_system(mcf_handle* h, int kind, const double* S, const int* Cptr, const int* Ccol,
                         const double* Cval, const int* angle_order);
/* Record the (c, y) pair of the first pressure solve of every projection of the steps that follow
   (a ring of `capacity` pairs; 0 disables), in angle order, tagged 4*step + projection where step
   counts from the start of each mcf_run / mcf_step call.  mcf_get_dump copies up to max_pairs
   recorded pairs and empties the ring.  Steps run without speculation while dumping.
   Build-only: the reference solves S directly. */
int mcf_set_dump(mcf_handle* h, int capacity);
int mcf_get_dump(mcf_handle* h, int max_pairs, double* c, double* y, int* tags, int* count);
/* GMRES(restart) on S (angle order) for n right-hand sides (n x r), from x0 (null: zero) until
   the true residual (compensated) |c - S y| <= rtol |c|, max_iter Arnoldi steps, or a restart that
   no longer lowers it by 10%; returns y, the Arnoldi steps taken and the final true relative
   residual per system.  The iterative counterpart of the direct S^{-1} apply
   (pressure_schur.py:20-57 runs GMRES on the full pressure system instead). */
int mcf_schur_gmres(mcf_handle* h, int kind, int n, const double* rhs, const double* x0, double rtol,
                    int max_iter, int restart, double* y, int* iters, double* relres);

/* Warm-start network for S y = c (no reference counterpart; the reference solves S directly,
   pressure_schur.py:92-111).  A periodic 1D convolution network over the boundary angle: `layers`
   convolutions of 5 taps with the given dilations, channels cin -> channels -> ... -> 1, ReLU
   between layers; input per angle position: c / |c| and cin-1 geometry features; output y / |c|.
   Forward and backward run on the matrix cores in BF16 with FP32 accumulation; weights, gradients
   and Adam moments are FP32.  Parameters are laid out per layer as W[out][in][tap] then bias[out].
   Training groups hold n samples of one boundary (length L): c, y (n x L, already scaled) and the
   features (L x (cin-1)).  x for predict / loss_grad: n x L x cin.  Loss: mean squared error over
   all positions of the output against y, or, when the group (or loss_grad) is given S (L x L,
   angle order, FP32), of the residual S yhat - c against the input c (y is then not used). */
typedef struct mcf_net mcf_net;
int mcf_net_create(int layers, int cin, int channels, const int* dilations, unsigned long long seed, mcf_net** out);
void mcf_net_destroy(mcf_net* n);
const char* mcf_net_last_error(const mcf_net* n);
int mcf_net_num_params(const mcf_net* n);
int mcf_net_get_params(mcf_net* n, float* out);
int mcf_net_set_params(mcf_net* n, const float* in);
int mcf_net_clear_groups(mcf_net* n);
int mcf_net_add_group(mcf_net* n, int count, int L, const float* c, const float* y, const float* features,
                      const float* S);
int mcf_net_train(mcf_net* n, int steps, int positions, float lr, unsigned long long seed, float* losses);
int mcf_net_predict(mcf_net* n, int nsamp, int L, const float* x, float* y);
int mcf_net_loss_grad(mcf_net* n, int nsamp, int L, const float* x, const float* y, const float* S, float* loss,
                      float* grad);
int mcf_net_profile(mcf_net* n, int nsamp, int L, int reps, float* ms_out);

/* Launch shapes of the rectangle inverse (no reference counterpart): the FP64 transform tile
   (0: 32x64, 1: 32x32, 2: 64x64 staged in LDS; 3: 32x32, 4: 16x32 split over K in 8-wave blocks;
   5: 32x32, 6: 16x32 split over K in 4-wave blocks; 7: the x-FFT form, a DCT-IV / DST-IV /
   mirrored-outlet sine transform along x through a half-length FFT with tridiagonal sweeps along y,
   which brings its own sweep, reported as 3; 8, 9: the parity-split transform in 8- and 4-wave
   blocks, two half-size products using V[yy-1-j][m] = (-1)^m V[j][m], which mcf_set_rectangle checks:
   a class whose bases lack it runs 4 instead) and the tridiagonal
   sweep width (0: 128, 1: 256, 2: 512 threads per mode).  By default the first mcf_run / mcf_step of a handle times every shape (and the
   x-FFT form's whole solve chain against the best dense chain, where every problem of a class has it) on the
   handle's own problems and keeps the fastest per problem class (pressure, velocity); the staged
   transform shapes give identical results, and so do the two split-K shapes of one block size; the
   families (staged, 8-wave and 4-wave split-K, parity-split, x-FFT) and the sweep shapes differ in
   round-off.  mcf_set_launch(h, g, t) pins
   both classes to (g, t) (g = 7: the x-FFT pair for each class that has it, the 16-row split-K transform with width t for
   the other); (-1, -1) returns to autotuning.  mcf_get_launch: [transform pressure,
   transform velocity, sweep pressure, sweep velocity, chosen (1) or default (0)]. */
int mcf_set_launch(mcf_handle* h, int gemm, int tri);
int mcf_get_launch(mcf_handle* h, int* out5);
/* The shapes of each class as mcf_get_launch reports them: [transform pressure, transform velocity,
   sweep pressure, sweep velocity] (the x-FFT pair is 7 with 3).  Restores a run's autotuned choice,
   e.g. after a checkpoint, so that a resumed run takes the same rounding path. */
int mcf_set_launch_classes(mcf_handle* h, const int* in4);

/* Speculative steps (no reference counterpart).  mode 1 (default): steady steps of mcf_run leave
   the Woodbury refinement launches out (pressure_schur.py:116-124 refines at most twice, and only
   when the first residual misses max(1e-10, 32 eps max|b|)); the device flags any step whose first
   residual would have refined, and the whole chunk is then replayed from a checkpoint with the
   refinements in place, so results equal mode 0 bit for bit.  Chunks that follow a refinement
   launch the pressure solves' first refinement pass and leave the velocity solve's and the
   second pressure pass out, flagged and replayed the same way.  mode 0: always launch them
   (device-predicated).  mode 2: as 1 with every lean step flagged (tests the replay).
   mcf_get_speculation: [mode, chunks replayed]. */
int mcf_set_speculation(mcf_handle* h, int mode);
int mcf_get_speculation(mcf_handle* h, int* out2);

/* Batched ensemble (no reference counterpart; the reference runs one case per process):
   cases are fully configured handles (tables, state, Helmholtz factors for dts[c]); all share the
   step parity and history state.  Every stage of the step is one launch over all cases and
   steady steps replay one HIP graph.  A handle in a live batch must not be stepped on its own
   concurrently.  diag receives n_cases * nsteps * MCF_DIAG_LEN doubles (case-major). */
typedef struct mcf_batch mcf_batch;
int mcf_batch_create(mcf_handle* const* cases, const double* dts, int n, mcf_batch** out);
void mcf_batch_destroy(mcf_batch* b);
const char* mcf_batch_last_error(const mcf_batch* b);
int mcf_batch_run(mcf_batch* b, int nsteps, double* diag);
/* Launch shapes of the batch (as mcf_set_launch / mcf_get_launch; autotuned at the first run). */
int mcf_batch_set_launch(mcf_batch* b, int gemm, int tri);
int mcf_batch_get_launch(mcf_batch* b, int* out5);
/* Speculative steps of the batch (as mcf_set_speculation; a replay re-runs every case). */
int mcf_batch_set_speculation(mcf_batch* b, int mode);
int mcf_batch_get_speculation(mcf_batch* b, int* out2);

#ifdef __cplusplus
}
#endif
#endif
