/*
 * nemoh_core -- MI355X solver core for the NEMOH Solver stage (C ABI).
 *
 * Replaces influence-matrix assembly and the dense complex solve of the
 * reference Solver (Solver/Core/SOLVE_BEM_INFD_DIRECT.f90,
 * Solver/Core/SOLVE_BEM_FD_DIRECT.f90 and the Green-function / solver
 * modules they call).  The Fortran host binds these with ISO_C_BINDING:
 * every argument is a plain pointer or scalar, arrays are owned by the
 * caller (column-major where 2-D), complex values are interleaved (re, im)
 * doubles as in COMPLEX(8).  The core copies inputs to the device and
 * writes outputs into the caller's buffers before returning; device copies
 * of geometry, tables, influence matrices and factors live in the handle.
 *
 * Every call returns NEMOH_OK (0) or an error code; nemoh_core_last_error()
 * gives the message.  One handle must not be used from two threads at once.
 */
#ifndef NEMOH_CORE_H
#define NEMOH_CORE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nemoh_core nemoh_core;

enum {
  NEMOH_OK = 0,
  NEMOH_ERR_ARG = 1,         /* invalid argument                               */
  NEMOH_ERR_HIP = 2,         /* HIP runtime failure (no device, out of memory) */
  NEMOH_ERR_PIVOT = 3,       /* pivot below 1e-20 (M_SOLVER.f90:49-52,61-64)   */
  NEMOH_ERR_OMEGA = 4,       /* |w| < 1e-4 (COMPUTE_GREEN_INFD.f90:252-255)    */
  NEMOH_ERR_FIT = 5,         /* Prony fit failure (COMPUTE_GREEN_FD.f90:1073, 1311-1314) */
  NEMOH_ERR_PANEL = 6,       /* degenerate panel (PREPARE_MESH.f90:112-120)    */
  NEMOH_ERR_CONVERGENCE = 7, /* mixed precision refinement did not converge    */
  NEMOH_ERR_WAVENUMBER = 8   /* X0 root search failed (COMPUTE_GREEN_FD.f90:1656-1660) */
};

enum { NEMOH_FP64 = 0, NEMOH_MIXED = 1 };

/* Creates a handle holding the mesh, the panel geometry and the wave
 * tables on the device.  Replaces the COM_VAR module state set up by
 * INITIALIZE (Solver/Core/INITIALIZATION.f90:28-69, COM_VAR.f90:29-95):
 * x,y,z = X,Y,Z(NP); m1..m4 = M1..M4(NFA), 1-based as in Fortran;
 * xg..tdis = XG,YG,ZG,XN,YN,ZN,AIRE,TDIS(NFA) and zer = ZER from GOMG
 * (PREPARE_MESH.f90:30-180); xr(328), xz(46), apd*(328,46) from CREK
 * (Solver/NEMOH.f90:179-208); nsym = NSYMY. */
int nemoh_core_create(nemoh_core **ctx, int np, int nfa, int nsym, const double *x, const double *y,
                      const double *z, const int *m1, const int *m2, const int *m3, const int *m4,
                      const double *xg, const double *yg, const double *zg, const double *xn,
                      const double *yn, const double *zn, const double *aire, const double *tdis,
                      double zer, double g, double rho, double depth, const double *xr, const double *xz,
                      const double *apd1x, const double *apd1z, const double *apd2x,
                      const double *apd2z);

/* One boundary-value problem: the exact drop-in for the calls
 * SOLVE_POTENTIAL_INFD_DIRECT(NVEL) (Solver/Core/SOLVE_BEM.f90:71) and
 * SOLVE_POTENTIAL_FD_DIRECT(NVEL,AMH,NEXP) (SOLVE_BEM.f90:83).
 * period = T; the regime (infinite vs finite depth) is chosen as at
 * SOLVE_BEM.f90:66.  Finite depth reads amh = AMH (X0 of w^2 h/g), nexp,
 * ambda(nexp), ar(nexp) = NEXP, AMBDA, AR from LISC; infinite depth reads
 * nq, qq(nq), cq(nq) = the module NQ, QQ, CQ at call time.  nvel =
 * NVEL(NFA*2**NSYMY).  Outputs ZIGB, ZIGS, ZPB, ZPS (NFA each).  The
 * factorisation is kept on the device while the period is unchanged,
 * mirroring w_previous (SOLVE_BEM_INFD_DIRECT.f90:82). */
int nemoh_core_solve(nemoh_core *ctx, double period, double amh, int nexp, const double *ambda,
                     const double *ar, int nq, const double *qq, const double *cq, const double *nvel,
                     double *zigb, double *zigs, double *zpb, double *zps);

/* Every problem of one frequency at once: nvel(NFA*2**NSYMY, nrhs) ->
 * pressure(NFA*2**NSYMY, nrhs) = rho i w phi (SOLVE_BEM.f90:91-96), the
 * argument SOLVE_BVP returns (SOLVE_BEM.f90:28).  When nint > 0 and nds
 * (nint, NFA*2**NSYMY) is given, forces(nint, nrhs) = -NDS . pressure
 * (Solver/NEMOH.f90:137-141) is returned as well (either output may be
 * NULL).  Same frequency parameters as nemoh_core_solve. */
int nemoh_core_solve_multi(nemoh_core *ctx, double period, double amh, int nexp, const double *ambda,
                           const double *ar, int nq, const double *qq, const double *cq, int nrhs,
                           const double *nvel, double *pressure, int nint, const double *nds,
                           double *forces);

/* Standalone mode: the frequency set-up of SOLVE_BVP and the direct
 * drivers done by the core itself (X0, theta rule, LISC), with the
 * reference's module NQ carried across calls in the handle (infinite-depth
 * calls reuse the NQ left by the last finite-depth call,
 * SOLVE_BEM_INFD_DIRECT.f90:79).  Then as nemoh_core_solve_multi. */
int nemoh_core_solve_frequency(nemoh_core *ctx, double period, int nrhs, const double *nvel,
                               double *pressure, int nint, const double *nds, double *forces);

/* Standalone mode with no per-frequency host input: the preProcessor's body
 * conditions, force weights and Froude-Krylov pressures are generated on
 * the device from the handle's mesh (preProcessor/Mesh.f90:207-233 panel
 * geometry, BodyConditions.f90:27-137, Integration.f90:27-86, Main.f90:
 * 199-243, Compute_Wave and CIH/SIH in Common/Environment.f90:86-123, the
 * preProcessor's own wave number Environment.f90:45-84).
 *
 * set_dofs: body_of_panel(NFA) = the body of each panel (1-based; NULL = all
 * body 1, the mesh files' cPanel); rad(nrad) the radiation cases and
 * force(nforce) the force-integration cases of Nemoh.cal; xeff, yeff the
 * wave reference point (Nemoh.cal line 5).  Fails with NEMOH_ERR_PANEL on a
 * panel the preProcessor rejects (area < 1e-7, Mesh.f90:222-225).  Binding
 * the handle to another mesh clears the set-up.
 *
 * body_conditions: at omega (rad/s) for nbeta headings beta (radians),
 * nvel(NFA*2**NSYMY, nbeta + nrad) = the normal velocities of the
 * diffraction problems then the radiation problems (the preProcessor's
 * order), pinc(NFA*2**NSYMY, nbeta) the incident-wave pressures and
 * nds(nforce, NFA*2**NSYMY) the force weights; any output may be NULL.
 *
 * solve_waves: the same problems solved at T = 2 pi / omega (the Solver's
 * main program, NEMOH.f90:135) with the set-up of nemoh_core_solve_frequency:
 * forces(nforce, nbeta + nrad) = -NDS . pressure and
 * fk(nforce, nbeta) = -NDS . pinc, the Froude-Krylov forces (Main.f90:
 * 229-234); nbeta + nrad <= 16. */
typedef struct {
  int body;      /* 1-based body index */
  int kind;      /* 1 = translation along dir, 2 = rotation about the axis through `axis` along dir */
  double dir[3];
  double axis[3];
} nemoh_dof;
int nemoh_core_set_dofs(nemoh_core *ctx, const int *body_of_panel, int nrad, const nemoh_dof *rad,
                        int nforce, const nemoh_dof *force, double xeff, double yeff);
int nemoh_core_body_conditions(nemoh_core *ctx, double omega, int nbeta, const double *beta, double *nvel,
                               double *pinc, double *nds);
int nemoh_core_solve_waves(nemoh_core *ctx, double omega, int nbeta, const double *beta, double *forces,
                           double *fk);

/* Kochin function (COMPUTE_KOCHIN.f90:25-153, called at SOLVE_BEM.f90:
 * 98-103) of the sources of every problem of the last solve on this handle,
 * at ntheta angles theta (radians): hkochin(ntheta, nrhs) complex, nrhs the
 * problem count of that call; the wave number is the Solver's kwave of that
 * frequency (w^2/g, or AMH/Depth in finite depth, SOLVE_BEM.f90:66-84) and
 * (xeff, yeff) the wave reference point (Nemoh.cal line 5, COM_VAR XEFF,
 * YEFF). */
int nemoh_core_kochin(nemoh_core *ctx, double xeff, double yeff, int ntheta, const double *theta,
                      double *hkochin);

/* No counterpart in the reference: nemoh_core_kochin for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results;
 * it calls no hipMalloc/hipFree: the angles and the values go through a
 * fixed 2.1 MB device array of the library, one per GPU, so ntheta is at
 * most 4096 (NEMOH_ERR_ARG beyond).  Calls on one GPU share that array:
 * calls that overlap in time on one GPU (two threads, two handles) race and
 * must be serialised by the caller. */
int nemoh_core_kochin_noalloc(nemoh_core *ctx, double xeff, double yeff, int ntheta, const double *theta,
                              double *hkochin);

/* Potential of the sources of every problem of the last solve on this
 * handle at npoints field points xyz[3 p + {0,1,2}] (x, y, z):
 * phi(npoints, nrhs) complex.  Replaces COMPUTE_POTENTIAL_DOMAIN
 * (COMPUTE_POTENTIAL_DOMAIN.f90:24-82) called per point of the free-surface
 * mesh at SOLVE_BEM.f90:104-110 (z = 0 there; the elevation is
 * ETA = i w / g PHI); the regime and wave number are those of the solve.
 * NEMOH_ERR_ARG if a point has AK0 (z + ZG) > -1.5e-6 for some panel, where
 * the reference stops. */
int nemoh_core_potential(nemoh_core *ctx, int npoints, const double *xyz, double *phi);

/* Frequency parameters as the standalone mode computes them. */
typedef struct {
  int finite_depth;
  double amh;
  int nexp;
  double ambda[31], ar[31];
  int nq;
  double qq[101], cq[101];
} nemoh_freq;
int nemoh_core_freq_setup(nemoh_core *ctx, double period, nemoh_freq *out);

/* Precision of the solve: NEMOH_FP64 (complex128 LU) or NEMOH_MIXED
 * (complex64 LU on MFMA refined in FP64 until max|r| <= tol * max|b|, which
 * bounds the normwise backward error by tol; tol <= 0 keeps the current
 * value, default 1e-8; max_iter <= 0 keeps the current limit, default 12). */
int nemoh_core_set_precision(nemoh_core *ctx, int mode, double tol, int max_iter);

/* Execution options (no counterpart in the reference).
 * NEMOH_OPT_STEPWISE_SOLVE: 1 = triangular solves as one launch per 64-row
 * block step (the path taken when the row blocks outnumber the compute
 * units), 0 = automatic (a single chained launch per triangle, whose
 * workgroups poll the solved blocks they need, when the row blocks fit). */
/* NEMOH_OPT_FP32_ASSEMBLY (mixed precision only): 1 = influence matrices
 * with FP32 far-field Rankine terms, FP32 tabulated wave terms and
 * exponential images, and complex64 storage of D and S (FP64 kept for the
 * selectors, table indices, near-panel Rankine integrals, the residual's and
 * phi's sums and the forces); 0 (default) = FP64 assembly and FP64 D, S.
 * Its error against FP64 reaches 2e-6..9e-6 on the 4000-panel barge and
 * 7e-4 in shallow water (h = 20 m), beyond the 1e-6 the mixed path keeps by
 * default. */
/* NEMOH_OPT_SOLVE_SPIN_LIMIT: polls a chained solve's wave makes while
 * waiting for an earlier row block before it gives up (default 1048576,
 * one to two seconds); when any gives up, the solves of that call are redone
 * with one launch per block step (nemoh_stats.solve_fallbacks counts it).
 * 0 makes every wait give up at once, which forces that path (tests). */
/* NEMOH_OPT_GRAPHS: 1 (default) = each frequency's assembly,
 * factorisation, solves and phi replay from a HIP graph that the handle
 * captures on the third use of a launch shape (precision, regime, problem
 * count and solve form; the first two run one by one, so that shapes used
 * once or twice, e.g. a batch of distinct bodies, pay no capture); graphs
 * are dropped when the handle's buffers or mesh change; 0 = always launched
 * one by one.  Timing (nemoh_core_enable_timing) always launches one by one. */
/* NEMOH_OPT_LU_RANK: 128 (default) = block steps in pairs with rank-128
 * trailing updates; 64 = one rank-64 trailing update per block step.
 * NEMOH_OPT_LU_LOOKAHEAD (rank 128 only): 1 = each pair's trailing update
 * split into the next pair's rows and columns and the bulk, which runs on a
 * second stream of the handle beside the next pair's panels; 0 (default) =
 * one launch per pair.  Not validated: the one run with it on (graphs
 * captured over the two streams by 8 request-loop threads) ended in a host
 * segmentation fault whose cause is not established. */
/* NEMOH_OPT_LU_FUSED_PANEL (rank 128 only): 1 = the panel of each pair's
 * second block column runs inside the pair's cross launch, whose workgroups
 * wait (bounded) for the pivot block that launch inverts; 0 (default) = a
 * launch of its own.  Both take the same time (the launch boundary is not
 * what bounds a pair's steps); a failed wait is reported as NEMOH_ERR_HIP. */
/* NEMOH_OPT_PAIRED_ASSEMBLY (FP64 assembly): 1 (default) = the wave part of
 * the Green function, which depends on a panel pair through its horizontal
 * distance and image depths only, is evaluated once per unordered pair
 * {i, j} for both elements D_ij and D_ji (the Rankine part, whose near-field
 * switch is the source panel's, once per element); 0 = every element on its
 * own.  Both agree to rounding. */
/* NEMOH_OPT_LU_BAND: 16 (default), 32 or 64 = height, in 64-row tiles, of
 * the bands in which the trailing update deals its tiles to the XCDs (each
 * XCD walks one contiguous run of band-high strips column by column); the
 * factors do not depend on it.
 * NEMOH_OPT_LU_TUNED: 1 = the update rank and band come from the
 * autotuner's table (nemoh_core_autotune) for the handle's precision and
 * number of 64-row blocks when it has an entry; 0 (default) = the handle's
 * own NEMOH_OPT_LU_RANK and NEMOH_OPT_LU_BAND. */
enum { NEMOH_OPT_STEPWISE_SOLVE = 1, NEMOH_OPT_FP32_ASSEMBLY = 2, NEMOH_OPT_SOLVE_SPIN_LIMIT = 3,
       NEMOH_OPT_GRAPHS = 4, NEMOH_OPT_LU_RANK = 5, NEMOH_OPT_LU_LOOKAHEAD = 6, NEMOH_OPT_LU_FUSED_PANEL = 7,
       NEMOH_OPT_PAIRED_ASSEMBLY = 8, NEMOH_OPT_LU_BAND = 9, NEMOH_OPT_LU_TUNED = 10 };
int nemoh_core_set_option(nemoh_core *ctx, int option, int value);

/* Timings (ms) of the last solve's phases and refinement statistics. */
typedef struct {
  double ms_assembly, ms_factor, ms_solve, ms_total;
  int refine_iterations;
  double refine_residual;
  int cache_hit;
  /* trailing-update kernel of the last factorisation (timing on): summed
   * launch durations, launch count and algorithmic real flops
   * (8 * r * m^2 per launch, r = the update's rank, m = order of the
   * trailing matrix) */
  double ms_lu_update;
  int lu_update_launches;
  double lu_update_flops;
  int lu_precision; /* NEMOH_FP64 or NEMOH_MIXED (complex64 factors) */
  int solve_fallbacks; /* chained solves redone stepwise after a bounded wait gave up (this handle) */
  int graph_replays;   /* frequencies replayed from a HIP graph (this handle) */
  int lu_rank, lu_band; /* factorisation form of the last factorisation (0 before the first) */
} nemoh_stats;
int nemoh_core_get_stats(const nemoh_core *ctx, nemoh_stats *out);

/* Autotuner (no counterpart in the reference): times the factorisation of
 * an order-nfa test matrix (well conditioned, pivots near 4) in the handle's
 * precision under each of the NEMOH_TUNE_CANDIDATES forms -- update rank 64
 * or 128 x tile band 16, 32 or 64 -- taking the fastest of reps runs of each,
 * records the fastest form in a process-wide table for that precision and
 * number of 64-row blocks, and turns NEMOH_OPT_LU_TUNED on for this handle.
 * entries (optional, NEMOH_TUNE_CANDIDATES of them) get each form's time,
 * chosen (optional) the index of the one recorded.  The handle's matrices
 * are overwritten (the next solve factorises again).  The two ranks
 * eliminate in a different order, so factors (and results, within
 * rounding) depend on the choice; a process whose table is fixed gives the
 * same results for the same inputs. */
#define NEMOH_TUNE_CANDIDATES 6
typedef struct {
  int rank, band;
  double ms; /* fastest factorisation time of the form */
} nemoh_tune_entry;
int nemoh_core_autotune(nemoh_core *ctx, int reps, nemoh_tune_entry *entries, int *chosen);
int nemoh_core_enable_timing(nemoh_core *ctx, int on);

/* The handle's last error message; with ctx = NULL, the calling thread's
 * last nemoh_core_create failure. */
const char *nemoh_core_last_error(const nemoh_core *ctx);
void nemoh_core_destroy(nemoh_core *ctx);

/* Host-side set-up routines of the reference, for callers without the
 * Fortran host.  prepare_mesh = PRE_PROC_MESH/GOMG (PREPARE_MESH.f90:30-180,
 * 1-based m1..m4); build_tables = CREK (Solver/NEMOH.f90:179-274) on the
 * GPU; wavenumber = X0 (COMPUTE_GREEN_FD.f90:1635); theta_rule = CINT_INFD
 * (COMPUTE_GREEN_INFD.f90:505), returns NQ; prony = LISC
 * (COMPUTE_GREEN_FD.f90:1007). */
int nemoh_core_prepare_mesh(int np, int nfa, int nsym, const double *x, const double *y, const double *z,
                            const int *m1, const int *m2, const int *m3, const int *m4, double *xg,
                            double *yg, double *zg, double *xn, double *yn, double *zn, double *aire,
                            double *tdis, double *zer);
int nemoh_core_build_tables(double *xr, double *xz, double *apd1x, double *apd1z, double *apd2x,
                            double *apd2z);
double nemoh_core_wavenumber(double ak);
/* The preProcessor's wave number for frequency w (Common/Environment.f90:
 * 45-84); NaN if its bisection fails. */
double nemoh_core_incident_wavenumber(double w, double depth, double g);
int nemoh_core_theta_rule(double akk, double *qq, double *cq);
int nemoh_core_prony(double akh, double amh, int *nexp, double *ambda, double *ar);

/* Batches (no counterpart in the reference, whose Solver runs one body per
 * process: Solver/NEMOH.f90:129-143 loops over the problems of one mesh).
 * Many independent bodies, each an ordinary handle (its mesh, tables and
 * degrees of freedom set), solved together one frequency at a time: the
 * assembly of every body is one launch, every step of the block LU one
 * launch over all matrices that have it (max_i ceil(nfa_i / 128) block
 * pairs of launches instead of the sum), each triangular solve one launch,
 * with the light per-body kernels on the same stream.  Every body's results
 * equal those of nemoh_core_solve_waves on its handle alone.  The handles
 * stay the caller's, must outlive the batch and are not to be used by other
 * threads during a batch call.  All bodies share nsym, depth and g, solve
 * the same number of problems (headings + radiation dofs <= max_rhs) and
 * take the batch's precision (FP32 assembly is not batched). */
typedef struct nemoh_batch nemoh_batch;
int nemoh_batch_create(nemoh_batch **batch, int nbody, nemoh_core *const *bodies, int precision, int max_rhs);
/* One frequency for every body: body i's forces (nforce_i x (nbeta +
 * nrad_i), column-major complex) at forces + 2 * offset_i, offset_i = the
 * sum of nforce_j (nbeta + nrad_j) over j < i; fk likewise (nforce_i x
 * nbeta); either may be NULL.  A body whose refinement needs more sweeps, or
 * whose chained solve gave up waiting, is solved again on its own handle
 * from the factors already on the device. */
int nemoh_batch_solve_waves(nemoh_batch *batch, double omega, int nbeta, const double *beta, double *forces,
                            double *fk);
/* The same with a frequency per body (omega[i] for body i; the bodies may
 * then fall in both depth regimes, which take one assembly launch each):
 * e.g. the frequencies of one mesh, one handle per frequency.  Each handle
 * keeps its own module NQ, as for nemoh_core_solve_waves on it. */
int nemoh_batch_solve_waves_multi(nemoh_batch *batch, const double *omega, int nbeta, const double *beta,
                                  double *forces, double *fk);
/* Timing of a batch's grouped rank-128 LU trailing updates (a measurement
 * hook, no counterpart in the reference).  on != 0 records events around
 * every update launch of the calls that follow and resets the sums; then
 * nemoh_batch_update_time gives, summed over those calls, the launches'
 * milliseconds, their algorithmic flops (8 x 128 x m^2 per matrix and
 * launch, m its trailing order) and bytes (2 m^2 + 256 m elements of 8
 * bytes on the mixed path, of 16 on FP64: the trailing matrix read and
 * written, the two 128-wide panels read), and the launch count. */
int nemoh_batch_set_timing(nemoh_batch *batch, int on);
int nemoh_batch_update_time(const nemoh_batch *batch, double *ms, double *flops, double *bytes, int *launches);
const char *nemoh_batch_last_error(const nemoh_batch *batch);
void nemoh_batch_destroy(nemoh_batch *batch);

/* Diagnostics used by the tests: the 16 Green-function values of every
 * (ISP, IFP) pair (the VAV*+VNS* outputs SP1 SM1 SP2 SM2 VSXP1 VSYP1 VSZP1
 * VSXM1 VSYM1 VSZM1 VSXP2 VSYP2 VSZP2 VSXM2 VSYM2 VSZM2, COM_VAR.f90:85-90)
 * into out[(isp*nfa + ifp)*16 + k]; and a check of the MFMA lane layout
 * (out: 16x16 product of fixed 16x4 and 4x16 operands, row-major). */
int nemoh_core_pairs(nemoh_core *ctx, double period, double amh, int nexp, const double *ambda,
                     const double *ar, int nq, const double *qq, const double *cq, double *out);
int nemoh_core_mfma_probe(int fp32, double *out);
/* GM of the handle's mesh: the largest horizontal vertex distance, to the
 * y-images as well when NSYMY = 1 (SOLVE_BEM_INFD_DIRECT.f90:66-75,
 * SOLVE_BEM_FD_DIRECT.f90:86-95), computed on the device. */
int nemoh_core_get_extent(const nemoh_core *ctx, double *gm);

/* Request loop -- no counterpart in the reference (one case per process);
 * for callers that serve many independent solves, of different meshes and
 * sizes, on one GPU.  A request is a mesh (as nemoh_core_create, 1-based
 * m1..m4), depth, rho, g, precision, nfreq periods, and per frequency the
 * normal velocities nvel(NFA*2**nsym, nrhs) (nfreq such blocks, complex);
 * nds(nint, NFA*2**nsym) is shared by the frequencies.  The reply is
 * forces(nint, nrhs, nfreq) = -NDS . pressure at each frequency, as
 * nemoh_core_solve_frequency of a session that solves the request's
 * frequencies in order.  Work is admitted one (request, frequency) at a
 * time, first come first served, while the device memory of the workers'
 * handles stays within budget_bytes: a busy worker counts the larger of what
 * its handle holds and what its frequency needs (FP64 influence matrices,
 * the complex64 factor copy in mixed precision, factors, right-hand sides,
 * geometry: about 32 N^2 bytes per symmetry block in FP64 and 40 N^2 in
 * mixed, N = NFA), an idle worker what its handle still holds; idle
 * workers' buffers are freed when they alone stand in the way, and with
 * nothing running the next frequency is always admitted.  Each of the
 * `workers` threads owns a solver handle (own HIP stream, resident matrices
 * and factors), so admitted work overlaps on the device; a result does not
 * depend on what else was in flight.  xr..apd2z are the wave tables as for
 * nemoh_core_create.  submit copies its inputs and returns at once with an
 * id; wait blocks until that request is done, writes the forces and
 * timing_ms[3] = {submit to done, submit to first start, first start to
 * done}, and returns the request's status (the first
 * error of its frequencies; others are unaffected).  Calls are thread-safe;
 * destroy finishes the queued work first. */
typedef struct nemoh_server nemoh_server;
int nemoh_server_create(nemoh_server **srv, int workers, double budget_bytes, const double *xr,
                        const double *xz, const double *apd1x, const double *apd1z,
                        const double *apd2x, const double *apd2z);
int nemoh_server_submit(nemoh_server *srv, int np, int nfa, int nsym, const double *x, const double *y,
                        const double *z, const int *m1, const int *m2, const int *m3, const int *m4,
                        double depth, double rho, double g, int precision, int nfreq,
                        const double *periods, int nrhs, const double *nvel, int nint,
                        const double *nds, long long *id);
/* A request whose problems are generated on the device, as
 * nemoh_core_solve_waves does: nfreq frequencies omegas (rad/s), nbeta
 * headings betas (radians), radiation and force degrees of freedom,
 * body_of_panel(NFA) (NULL: all body 1), wave reference point xeff, yeff.
 * Its reply: forces(nforce, nbeta + nrad, nfreq) and fk(nforce, nbeta,
 * nfreq), the Froude-Krylov forces. */
int nemoh_server_submit_waves(nemoh_server *srv, int np, int nfa, int nsym, const double *x,
                              const double *y, const double *z, const int *m1, const int *m2,
                              const int *m3, const int *m4, double depth, double rho, double g,
                              int precision, int nfreq, const double *omegas, int nbeta,
                              const double *betas, const int *body_of_panel, int nrad,
                              const nemoh_dof *rad, int nforce, const nemoh_dof *force, double xeff,
                              double yeff, long long *id);
/* fk is written for wave requests only; forces, fk and timing_ms may be NULL. */
int nemoh_server_wait(nemoh_server *srv, long long id, double *forces, double *fk, double *timing_ms);
/* Device bytes the workers' handles hold now, and the largest total
 * charged against the budget while other frequencies were running (both
 * may be NULL). */
int nemoh_server_memory(nemoh_server *srv, double *held, double *peak);
/* Most frequencies of one wave request a worker solves at once (1 to 64;
 * 1, the default, takes one frequency at a time).  A worker then takes as
 * many of its request's pending frequencies as the budget allows at the
 * request's footprint each, binds that many handles of its own to the mesh
 * and solves them as one batch (nemoh_batch_solve_waves_multi); results are
 * those of one frequency at a time, bit for bit.  It pays for requests with
 * many frequencies of one small mesh; when each request brings a new mesh,
 * re-binding and re-tabling the handles costs more than the batch gains. */
int nemoh_server_set_batch(nemoh_server *srv, int frequencies);
const char *nemoh_server_last_error(nemoh_server *srv);
void nemoh_server_destroy(nemoh_server *srv);

#ifdef __cplusplus
}
#endif
#endif
