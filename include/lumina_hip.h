/*
 * lumina_hip.h -- C ABI of the MI355X (gfx950) CDA solve hot path.
 *
 * The entry points replace the reference's device interface for the CDA
 * solve (Rust crate lumina-compute plus the on-the-fly branch of
 * lumina-core). Each declaration names the reference item it stands in for
 * (paths relative to the reference repository root).
 *
 * Conventions (mirroring lumina-compute/src/backend.rs):
 *  - complex numbers are lh_c64 = {re, im}, the layout of
 *    num_complex::Complex<f64> (#[repr(C)]);
 *  - vectors are interleaved per dipole, element 3*i + c (cda/mod.rs:397-400);
 *  - polarisability blocks are row-major [9] (types.rs:12-18);
 *  - inputs are borrowed; outputs are caller-allocated; the library never
 *    keeps a caller pointer after a call returns;
 *  - every call returns an lh_status; lh_last_error() gives the message of
 *    the last failure on the calling thread;
 *  - sessions own their device memory and HIP stream; different sessions
 *    may be used from different threads concurrently (no global lock); a
 *    single session must not be used by two threads at once.
 */
#ifndef LUMINA_HIP_H
#define LUMINA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* num_complex::Complex<f64> */
typedef struct {
  double re, im;
} lh_c64;

/* ComputeError (backend.rs:13-22) plus SolverError::ConvergenceFailure
 * (solver/mod.rs:18-31). */
typedef enum {
  LH_OK = 0,
  LH_UNAVAILABLE = 1,          /* ComputeError::Unavailable */
  LH_DEVICE_ERROR = 2,         /* ComputeError::DeviceError (also dimension mismatch, gpu.rs:411-416) */
  LH_OUT_OF_MEMORY = 3,        /* ComputeError::OutOfMemory */
  LH_INVALID_ARGUMENT = 4,
  LH_CONVERGENCE_FAILURE = 5   /* SolverError::ConvergenceFailure { max_iter, residual } */
} lh_status;

/* Arithmetic of the far-field pair sum. FP64: everything complex FP64.
 * MIXED: Arnoldi matvecs evaluate and accumulate the far-field Green's
 * tensor in FP32 (compensated accumulation); restart residuals, Gram-Schmidt,
 * Givens rotations, the near-field (FCD) tensors and alpha^-1 stay FP64. */
typedef enum { LH_PREC_FP64 = 0, LH_PREC_MIXED = 1 } lh_precision;

typedef struct lh_backend lh_backend;
typedef struct lh_session lh_session;
typedef struct lh_pairsum lh_pairsum;

/* Message of the last failed call on this thread ("" if none). */
const char* lh_last_error(void);

/* DeviceInfo (backend.rs:26-39). backend_type: 0 Cpu, 1 Gpu, 2 Distributed. */
typedef struct {
  char name[256];
  char arch[64];
  int backend_type;
  uint64_t memory_bytes;
  uint32_t compute_units;
} lh_device_info;

/* GpuBackend::new_blocking (gpu.rs:301) -- one backend per device. */
int lh_backend_create(int device_ordinal, lh_backend** out);
void lh_backend_destroy(lh_backend* backend);
/* ComputeBackend::device_info (backend.rs:64). */
int lh_backend_device_info(const lh_backend* backend, lh_device_info* out);

/* ---- stored-matrix session: MatvecSession (backend.rs:47-55) ----------- */

/* ComputeBackend::create_session(n) (backend.rs:70-72; gpu.rs:310-399);
 * n is the number of unknowns (3N). */
int lh_session_create(lh_backend* backend, size_t n, lh_session** out);
/* MatvecSession::upload_matrix (backend.rs:51; gpu.rs:401-426); row-major n x n. */
int lh_session_upload_matrix(lh_session* session, const lh_c64* matrix, size_t n);
/* MatvecSession::matvec (backend.rs:54; gpu.rs:428-488): y = A x, length n. */
int lh_session_matvec(lh_session* session, const lh_c64* x, lh_c64* y, size_t n);
/* ComputeBackend::assemble_and_create_session (backend.rs:109-117; the wgpu
 * form gpu.rs:508-607 with shaders/matrix_assemble.wgsl), called by
 * CdaSolver for FCD off, no lattice, no substrate, off-grid dipoles
 * (cda/mod.rs:298-311): a session of n = 3 n_dipoles unknowns holding
 * A = blockdiag(diagonal_blocks) - G, G the point tensor (greens.rs:34-71)
 * of r_i - r_j, assembled on the device in FP64 (the wgpu form assembles
 * in FP32). diagonal_blocks: [n_dipoles][9] row-major alpha^-1
 * (compute_diagonal_blocks, assembly.rs:355-363). A pair closer than
 * 1e-10 nm gets a zero block, as in the wgpu shader. The reference's
 * None for n_dipoles == 0 is LH_INVALID_ARGUMENT here. */
int lh_session_assemble(lh_backend* backend, const double* positions, size_t n_dipoles, double k,
                        const lh_c64* diagonal_blocks, lh_session** out);
void lh_session_destroy(lh_session* session);

/* ---- on-the-fly pair-sum session --------------------------------------- */

/* The arguments of matvec_on_the_fly (assembly.rs:209-218) with the
 * Dipole list (types.rs:12-18) split into positions and polarisabilities,
 * plus the LatticeSpec (lattice.rs:29-44) and SubstrateRuntime::Fresnel
 * (substrate.rs:103-107) it may carry. */
typedef struct {
  size_t n_dipoles;
  const double* positions;        /* [n][3], nm */
  const lh_c64* polarisability;   /* [n][9] row-major, or [9] if polarisability_uniform */
  int polarisability_uniform;
  double k;                       /* 2 pi n_env / lambda, nm^-1 */
  int use_fcd;                    /* CdaSolver::use_fcd (cda/mod.rs:53) */
  double cell_size;               /* CdaSolver::cell_size (cda/mod.rs:55) */
  int has_lattice;                /* 2D lattice only (a2 present) */
  double a1[3], a2[3];
  double ewald_eta;               /* <= 0: automatic sqrt(pi / A) (lattice.rs:96-102) */
  int truncation_real, truncation_recip;
  double k_bloch[3];              /* SimulationParams::k_bloch (types.rs:92) */
  int has_substrate;
  double z_interface;
  lh_c64 delta_eps;               /* reflection factor (substrate.rs:191-197) */
  int precision;                  /* lh_precision */
  /* Optional [n][3] positions on which the FCD near/far decision of each
   * pair is taken (NULL: the positions themselves). No reference
   * counterpart: it holds the classification fixed while a dipole is moved,
   * for finite differences and shape derivatives. */
  const double* classify_positions;
} lh_pairsum_spec;

/* Incident plane wave (IncidentField, types.rs:37-74). The polarisation is
 * complex so that circular polarisation (e.g. (x +- i y)/sqrt 2) can be
 * expressed; a real polarisation reproduces the reference. */
typedef struct {
  double direction[3];
  lh_c64 polarisation[3];
  double amplitude;
} lh_incident;

/* Solve statistics. iterations = total Arnoldi steps (iterative.rs:129);
 * cycles = restart cycles entered; matvecs = matvecs performed on the device
 * (the reference's redundant cycle-top matvec, iterative.rs:55, is not
 * repeated: its result equals the previous cycle's explicit residual);
 * matvecs_fp32 = how many of those used the FP32 far field;
 * residual = final ||rhs - A x|| / ||rhs|| (FP64);
 * pair_evals = Green's pair evaluations performed (N(N-1) per matvec, plus
 * N^2 image pairs with a substrate); solve_group = the solve group of a
 * batch whose stream issued this solve's kernels (0 outside batches). */
typedef struct {
  uint64_t iterations;
  uint64_t cycles;
  uint64_t matvecs;
  uint64_t matvecs_fp32;
  double residual;
  double pair_evals;
  int status;                     /* LH_OK or LH_CONVERGENCE_FAILURE */
  int32_t solve_group;
} lh_solve_info;

/* Builds the device state for the closure of cda/mod.rs:329-345 / 375-391
 * (on-the-fly matvec): positions, alpha^-1 blocks (assembly.rs:325-327),
 * the FCD near-pair list and tensors (greens.rs:94-167). */
int lh_pairsum_create(lh_backend* backend, const lh_pairsum_spec* spec, lh_pairsum** out);
/* Re-targets a session to a new wavelength: k, polarisabilities and the
 * substrate factor change, geometry is kept (runner.rs:150-186). */
int lh_pairsum_set_wavelength(lh_pairsum* s, double k, const lh_c64* polarisability,
                              int polarisability_uniform, lh_c64 delta_eps);
/* matvec_on_the_fly (assembly.rs:209-318): y = A x, host vectors of 3N.
 * precision selects the far-field arithmetic of this one product. */
int lh_pairsum_matvec(lh_pairsum* s, const lh_c64* x, lh_c64* y, int precision);
/* solve_gmres (iterative.rs:33-160) on this session's operator: restarted
 * GMRES(min(30, 3N)), zero initial guess, stop at ||r|| < tol ||rhs||.
 * Returns LH_CONVERGENCE_FAILURE (x holds the last iterate) like
 * SolverError::ConvergenceFailure. */
int lh_pairsum_solve(lh_pairsum* s, const lh_c64* rhs, double tol, size_t max_iter,
                     lh_c64* x, lh_solve_info* info);
/* build_incident_field_vector (assembly.rs:178-194) for this session's dipoles. */
int lh_pairsum_incident(lh_pairsum* s, const lh_incident* inc, lh_c64* e_inc);
/* C_ext, C_abs, C_sca of moments p (cda/mod.rs:187-235) -> out[3]. */
int lh_pairsum_cross_sections(lh_pairsum* s, const lh_c64* p, const lh_c64* e_inc,
                              double amplitude, double out[3]);
/* CdaSolver::compute_cross_sections (cda/mod.rs:175-236) on the on-the-fly
 * path: builds E_inc, solves, reduces. p_out (3N, may be NULL) receives the
 * moments. */
int lh_pairsum_compute_cross_sections(lh_pairsum* s, const lh_incident* inc, double tol,
                                      size_t max_iter, double out[3], lh_c64* p_out,
                                      lh_solve_info* info);
/* Adjoint gradients of C_ext (no reference counterpart; the backward of the
 * solve, as a torch.autograd.Function would expose it). Solves A p = E (the
 * forward solve of lh_pairsum_compute_cross_sections) and A lam = conj(E)
 * on the session's operator, which is complex symmetric (free space, with or
 * without the Fresnel image), then
 *   grad_eps[2i], grad_eps[2i+1] = dC_ext/dRe(eps_i), dC_ext/dIm(eps_i),
 *     from dainv_deps = d(alpha_i^-1)/d eps_i ([n][9] or [9] if
 *     dainv_uniform, row-major complex: the caller's material model);
 *   grad_pos[3i+c] = dC_ext/d r_i,c, with the FCD near/far classification
 *     of every pair held fixed; NULL skips it.
 * cross_sections receives C_ext, C_abs, C_sca of the forward solve.
 * LH_UNAVAILABLE: a lattice (the Ewald operator is not symmetric: pairs at
 * equal height take sgn 0 = +1 for r and -r in the spectral sum), or grad_pos
 * with a substrate. LH_CONVERGENCE_FAILURE if either solve hits max_iter
 * (outputs are still written). */
int lh_pairsum_adjoint_grad(lh_pairsum* s, const lh_incident* inc, double tol, size_t max_iter,
                            const lh_c64* dainv_deps, int dainv_uniform,
                            double cross_sections[3], double* grad_eps, double* grad_pos,
                            lh_solve_info* forward_info, lh_solve_info* adjoint_info);
/* fields.rs:18-48 field_at_point, for m observation points (the loop of
 * compute_near_field_map, fields.rs:74-91): E = E_inc(obs) + sum_i
 * G(obs, r_i) p_i with the free-space point tensor (greens.rs:34-71: no
 * FCD, image or lattice term, as the reference), a dipole with
 * |obs - r_i|^2 < 1e-20 left out. moments (3N) = the solved moments, or NULL
 * for the solution of this session's last solve (already on the device).
 * points: [m][3]; field ([m][3], may be NULL); intensity ([m], may be NULL)
 * = |E|^2 (fields.rs:87). FP64 throughout. */
int lh_pairsum_field_at_points(lh_pairsum* s, const lh_c64* moments, const lh_incident* inc,
                               size_t m, const double* points, lh_c64* field,
                               double* intensity);
/* No reference counterpart: lh_pairsum_field_at_points for callers that
 * cannot allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc). moments (3N complex, or NULL for this session's last
 * solution), points ([m][3]), field ([m][3] complex, may be NULL) and
 * intensity ([m], may be NULL) are DEVICE pointers; the call allocates
 * nothing: its working memory is a fixed 64 MiB array held by each GPU the
 * library runs on until the process ends, which limits the session to
 * N <= 262,144 dipoles (LH_INVALID_ARGUMENT beyond). Calls on one GPU share
 * that array and are serialised. Returns when the results are written. */
int lh_pairsum_field_at_points_noalloc(lh_pairsum* s, const lh_c64* moments,
                                       const lh_incident* inc, size_t m, const double* points,
                                       lh_c64* field, double* intensity);
/* fields.rs:156-221 compute_far_field: theta = pi it / max(n_theta - 1, 1),
 * phi = 2 pi ip / n_phi (theta-major order), intensity = |(I - rhat rhat) F|^2
 * with F = sum_i p_i e^{-ik rhat.r_i}. moments as for
 * lh_pairsum_field_at_points. theta, phi (may be NULL) and intensity:
 * [n_theta * n_phi]. */
int lh_pairsum_far_field(lh_pairsum* s, const lh_c64* moments, size_t n_theta, size_t n_phi,
                         double* theta, double* phi, double* intensity);
void lh_pairsum_destroy(lh_pairsum* s);

/* ---- batched solves ------------------------------------------------------ */

/* One solve of a batch: a pair-sum spec (its precision field is ignored:
 * the batch precision applies) and its incident field. Problems whose
 * positions pointer, n_dipoles, use_fcd and cell_size agree share one
 * device copy of the geometry and its near-pair list. */
typedef struct {
  lh_pairsum_spec spec;
  lh_incident incident;
} lh_batch_problem;

/* Many independent compute_cross_sections calls (the rayon wavelength
 * sweep of runner.rs:145-218 and the GUI sweep, app.rs:297-392) advanced in
 * lock-step by one stream of batched kernels, the restart cycle captured
 * once as a HIP graph and replayed. cross_sections: [count][3];
 * moments (may be NULL): per problem a caller buffer of 3N;
 * infos (may be NULL): [count]. Returns LH_CONVERGENCE_FAILURE if any
 * problem failed to converge (its info.status says which). */
int lh_batch_solve(lh_backend* backend, const lh_batch_problem* problems, size_t count,
                   double tol, size_t max_iter, int precision, double* cross_sections,
                   lh_c64* const* moments, lh_solve_info* infos);

/* ---- wavelength sweep with the per-wavelength set-up on the device ------ */

/* A tabulated permittivity eps(lambda), interpolated by natural cubic
 * splines of Re eps and Im eps through the points (spline.rs:32-90) and
 * defined on [wavelength_nm[0], wavelength_nm[n_points - 1]]
 * (johnson_christy.rs:266-279; a Johnson-Christy table enters as
 * eps = (n + i k)^2 at its wavelengths). */
typedef struct {
  size_t n_points;               /* >= 2 */
  const double* wavelength_nm;   /* strictly increasing */
  const lh_c64* epsilon;
} lh_material;

/* The per-dipole optical model of the CLI run (runner.rs:150-186). The
 * geometry part of `geometry` is used as for lh_pairsum_create (positions,
 * use_fcd, cell_size, lattice, k_bloch, has_substrate and z_interface,
 * precision, classify_positions); its polarisability, k and delta_eps are
 * ignored: they are set up on the device for every wavelength. Dipole i has
 * material material[i] (NULL: all 0), volume spacing[i]^3 (NULL:
 * uniform_spacing^3) and, when hint[i] >= 0, the ellipsoid tensor of
 * types.rs:286-317 with depolarisation factors hint_depol[3 hint[i] ..] and
 * rotation hint_rotation[9 hint[i] ..] (row-major); otherwise the
 * Clausius-Mossotti polarisability with the radiative correction
 * (types.rs:243-259). The substrate reflection factor (has_substrate) comes
 * from material substrate_material, retarded or quasi-static
 * (substrate.rs:148-197). */
typedef struct {
  lh_pairsum_spec geometry;
  const int32_t* material;
  const double* spacing;
  double uniform_spacing;
  const int32_t* hint;
  size_t n_hints;
  const double* hint_depol;
  const double* hint_rotation;
  double n_env;                  /* k = 2 pi n_env / lambda, eps_m = n_env^2 */
  int substrate_material;
  int substrate_retarded;
} lh_sweep_spec;

/* The CLI wavelength sweep (runner.rs:145-218) as one batched solve: the
 * geometry and its near list are built once, then for every wavelength the
 * permittivities (splines), polarisabilities and their inverses, the
 * substrate factor and the incident field are computed by device kernels
 * straight into the batch, and all wavelengths are solved together.
 * cross_sections: [n_wavelengths][3] (C_ext, C_abs, C_sca); a wavelength
 * that does not converge gets a NaN row (runner.rs:189-203) and the call
 * returns LH_CONVERGENCE_FAILURE after writing every row. infos may be NULL.
 * A wavelength outside a used material's table is LH_INVALID_ARGUMENT
 * (MaterialError::OutOfRange). */
int lh_sweep_solve(lh_backend* backend, const lh_sweep_spec* spec, const lh_material* materials,
                   size_t n_materials, const double* wavelengths_nm, size_t n_wavelengths,
                   const lh_incident* incident, double tol, size_t max_iter,
                   double* cross_sections, lh_solve_info* infos);

/* The same, split so that a batch stays resident on the device and can be
 * solved repeatedly: create uploads geometry, alpha^-1, near tensors and
 * incident fields; run solves every problem; results reads them back. */
typedef struct lh_batch lh_batch;
int lh_batch_create(lh_backend* backend, const lh_batch_problem* problems, size_t count,
                    lh_batch** out);
int lh_batch_run(lh_batch* batch, double tol, size_t max_iter, int precision);
/* No reference counterpart: how many restart cycles one graph launch of
 * lh_batch_run replays. cycles = 0 (the default): 1, 2, 4, then 8 per
 * launch, one host read of the running-solve counter after each; cycles > 0:
 * that many (at most 64); cycles = -1: every cycle up to max_iter in one
 * launch, so that the whole batch is solved by one hipGraphLaunch and one
 * host read (solves that finish early skip the rest of the graph, at the
 * cost of one status read per work item). Results are identical for every
 * setting. LH_INVALID_ARGUMENT for cycles < -1. */
int lh_batch_set_graph_cycles(lh_batch* batch, int cycles);
int lh_batch_results(lh_batch* batch, double* cross_sections, lh_c64* const* moments,
                     lh_solve_info* infos);
void lh_batch_destroy(lh_batch* batch);

/* Instrumentation (no reference counterpart): when enabled, the far-field
 * pair-sum launches of lh_batch_run are bracketed by HIP events on the
 * batch's stream (event nodes of the captured restart-cycle graphs, which
 * keep running as graphs) and their device time is accumulated. pairs_*
 * count the Green's pair evaluations those launches performed (N(N-1), plus
 * N^2 image pairs, per participating solve). host_round_trips = graph
 * launches (or host-issued cycles) of the last lh_batch_run, each followed
 * by one read of the running-solve counter; cycles_issued = restart cycles
 * they issued (a graph launch replays several). solve_groups = 1, or 2 when
 * the batch's solves were split into two groups of about equal pair work
 * whose restart cycles run on two streams, the far fields of the groups
 * taking turns (each group's reduction and Gram-Schmidt run beside the other
 * group's far field); the far-field launches are then counted per group. */
typedef struct {
  double far64_ms, far32_ms;
  uint64_t far64_launches, far32_launches;
  double far64_pairs, far32_pairs;
  double run_ms;          /* wall time of the last lh_batch_run */
  int32_t far64_block;    /* far-field targets per work item in use (FP64, FP32) */
  int32_t far32_block;
  double tune_ms;         /* time spent choosing them (0 when fixed) */
  int32_t host_round_trips;
  int32_t solve_groups;
  uint64_t cycles_issued;
} lh_kernel_times;
int lh_batch_set_timing(lh_batch* batch, int enable);
int lh_batch_kernel_times(lh_batch* batch, lh_kernel_times* out);

#ifdef __cplusplus
}
#endif

#endif /* LUMINA_HIP_H */
