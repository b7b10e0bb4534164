// Launch interface of the solver-core kernels (host side).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include <algorithm>
#include <cstddef>

#include "green.h"
#include "linalg.h"

namespace nemoh {

constexpr int kMaxRhs = 16;  // right-hand sides per triangular-solve pass

// Influence-matrix outputs of one assembly launch (row-major, leading
// dimension ld).  dm/sm are the antisymmetric block (nsym = 1 only);
// dp32/dm32 are optional single-precision copies of D for the mixed path.
struct MatOut {
  int ld;
  double2* dp;
  double2* sp;
  double2* dm;
  double2* sm;
  float2* dp32;
  float2* dm32;
  int ld32;  // leading dimension of dp32 / dm32
};

// Iterative refinement of one symmetry block in device memory: bmax = max|b|,
// rmax = max|r| of the current residual, iters = sweeps done, done = 1 once
// rmax <= tol bmax (or on a non-finite residual, bad = 1).  Kernels given a
// state return at once when it is done.
struct RefineState {
  double bmax, rmax;
  int done, iters, bad, pad;
};

// Influence matrices of the mixed path (complex64, row-major, leading
// dimension ld): D for the FP64 residual, a copy of D that the LU factors in
// place, and S; m: the antisymmetric block (nsym = 1 only).
struct MatOut32 {
  int ld;
  float2* dp;
  float2* dp_fac;
  float2* sp;
  float2* dm;
  float2* dm_fac;
  float2* sm;
};

void launch_crek(const double* xr, const double* xz, double* tab, hipStream_t st);
// Mixed-precision assembly (green_mixed.h): FP64 selectors, indices and
// near-panel Rankine integrals, FP32 far-field, wave and exponential-image
// terms on the FP32 tables tab32; pairs with a point at the free surface in
// FP64, stored in single precision.
void launch_assemble_mixed(const Geometry& G, const Tables& T, const float4* tab32, const FreqParams* P, int fd,
                           int nsym, const MatOut32& out, const int* surf, int nsurf, hipStream_t st);
// fd: the regime of *P (finite depth 1, infinite 0), known to the host.
// surf[0..nsurf) (device): the panels with zg >= ZER, whose rows and columns
// take the free-surface branch of the wave part.  paired: the wave part of
// each unordered pair once for both elements (assemble_pair_kernel), else
// once per element (assemble_kernel).
void launch_assemble(const Geometry& G, const Tables& T, const FreqParams* P, int fd, int nsym, const MatOut& out,
                     const int* surf, int nsurf, bool paired, hipStream_t st);
void launch_pairs(const Geometry& G, const Tables& T, const FreqParams* P, int fd, double* out, hipStream_t st);
// One body of a grouped paired assembly: its geometry, frequency
// parameters and matrices, its nt x nt grid of 8-panel blocks (nwaves =
// nt (nt + 1) / 2 pair blocks, four per workgroup) on workgroups [wg0, wg1)
// of the launch.
struct AsmJob {
  Geometry G;
  Tables T;
  const FreqParams* P;
  MatOut out;
  int nt, nwaves, wg0, wg1;
};
// The paired assembly of every body of jobs[0, count) in one launch of grid
// workgroups (the bulk pairs; free-surface panels are the caller's, through
// launch_assemble's surface pass).  All bodies share the regime (fd) and
// symmetry (nsym); f32: every job writes the complex64 copy of D.
void launch_assemble_group(const AsmJob* jobs, int count, int grid, int fd, int nsym, bool f32, hipStream_t st);
// The free-surface pass of launch_assemble alone (rows and columns of the
// panels surf[0, nsurf) with zg >= ZER), after a grouped assembly.
void launch_assemble_surface(const Geometry& G, const Tables& T, const FreqParams* P, int fd, int nsym,
                             const MatOut& out, const int* surf, int nsurf, hipStream_t st);
// Block LU without pivoting in place in A (n x n, leading dimension ld):
// afterwards the blocks below the diagonal hold L (unit block lower), the
// blocks above hold U (unit block upper) and P (ceil(n/64) blocks of 64x64,
// row-major) the inverses of the pivot blocks.  *info receives the first row
// whose pivot has modulus <= 1e-20 (1-based), if any.  rank 64: one rank-64
// trailing update per block step; rank 128: block steps in pairs, a small
// cross update of the pair's second block row and column, then one rank-128
// update of the rest of the trailing matrix; with st2 (look-ahead) that update
// is split between the next pair's block rows and columns on st and the
// bulk on st2, which overlaps the next pair's panels (dep: 2 events per
// pair, ceil(n/128) pairs; st waits for st2 before returning).  upd_ev
// (optional, 2 per block step): each trailing-update launch (the rank-128
// ones in the paired form) is made with hipExtLaunchKernel, which stamps
// [2 t] at its start and [2 t + 1] at its end, and upd_flops[t] gets its
// algorithmic flops; returns the number t of such launches.
// luctl (optional, 2 lu_pairs(n) ints of device memory): with it, the
// paired form folds each pair's second panel into its cross launch, whose
// workgroups wait (at most spin_limit polls; info[1] |= 2 when one gives up)
// for the pivot block inverted inside the same launch.
// band: height, in 64-row tiles, of the bands in which the trailing update
// deals its tiles to the XCDs (update_tile; the factors do not depend on it).
template <typename T>
int launch_lu(C2<T>* A, int ld, int n, C2<T>* P, int* info, hipStream_t st, int rank = 128, hipStream_t st2 = nullptr,
              hipEvent_t* dep = nullptr, hipEvent_t* upd_ev = nullptr, double* upd_flops = nullptr,
              int* luctl = nullptr, int spin_limit = 1 << 20, int band = 16);
// Writes a well-conditioned n x n test matrix (pivots near 4) into A, for
// timing factorisations.
template <typename T>
void launch_fill_test_matrix(C2<T>* A, int ld, int n, hipStream_t st);
// block pairs of an order-n factorisation, with one to spare
inline int lu_pairs(int n) { return ((n + kNB - 1) / kNB + 1) / 2 + 1; }
// Panel launch of the paired form: L panel of block column lc (nL blocks
// below the diagonal) and the U rows ur0 (nU0 blocks) and ur1 (nU1).
struct PanelJob {
  int lc, nL, ur0, nU0, ur1, nU1;
};
// One matrix's share of a grouped LU launch (launch_lu_group): its
// workgroups are [wg0, wg1) of the launch, wg0 a multiple of 8.
template <typename T>
struct LuJob {
  C2<T>* A;
  C2<T>* P;
  int* info;
  int ld, n, k, wg0, wg1;
  PanelJob J;
  int band;  // the update's tile bands (0: the default, 16)
};
// Control block of the chained triangular solves of one handle: the two
// triangles' ticket counters (reset on the device at the start of every
// chained solve).
struct ChainCtl {
  unsigned ticket[2];
  int pad[2];
};
// One system's share of a grouped chained solve (launch_trsv_group):
// factors A (leading dimension ld) and P, right-hand sides / solution B,
// scratch W, its control block and status word, and its refinement state;
// workgroups [wg0, wg1) of the launch.
template <typename T>
struct TrsvJob {
  const C2<T>* A;
  const C2<T>* P;
  C2<T>* B;
  C2<T>* W;
  ChainCtl* ctl;
  int* status;
  const RefineState* skip;
  int ld, n, nrhs, wg0, wg1;
};
// One body (and symmetry block) of a batch's grouped per-body kernels: its
// buffers, sizes and the workgroup range it takes in each kind of grouped
// launch (kind 0: elementwise over n x nrhs, 256 per workgroup; 1: gemv, 16
// rows per workgroup; 2: forces, nforce x nrhs; 3: Froude-Krylov forces,
// nforce x nbeta).  Block 0's job also carries block 1's Z and F (Z2, F2)
// for the combination.
struct BodyJob {
  const double2* V;
  double2 *B, *R, *Z, *F, *Z2, *F2;
  float2* R32;
  RefineState* rs;
  double* red;  // 64 partial maxima
  const double2 *D, *S;
  int ld, n, nsym, nrhs, nbeta, nforce;
  double tol, rho;
  double2 *sig, *pot, *pres;
  const double2* pinc;
  const double* nds;
  double2* force;  // nforce x nrhs, then nforce x nbeta
  int* info;       // the body's pivot report and wait status
  int wg[4], nwg[4];
};
// A body's pivot report, wait status and refinement states after a batch's
// solves (launch_group_status).
struct BatchStatus {
  int info[2];
  int pad[2];
  RefineState rs[2];
};
void launch_group_info_reset(const BodyJob* jobs, int count, hipStream_t st);
void launch_group_status(const BodyJob* jobs, int count, int nb, BatchStatus* out, hipStream_t st);
// operations of launch_group_elementwise
enum { kGroupRhs = 0, kGroupCopyBZ, kGroupCopyBR, kGroupCopyBRSkip, kGroupToF32, kGroupToF64Set, kGroupToF64Add };
void launch_group_elementwise(const BodyJob* jobs, int count, int grid, int op, double bx, hipStream_t st);
// the right-hand side's (first) or the residual's max and the convergence test
void launch_group_residual_max(const BodyJob* jobs, int count, int first, hipStream_t st);
void launch_group_refine_reset(const BodyJob* jobs, int count, hipStream_t st);
// phi = 0: R -= D Z (skipped once converged); 1: F = S Z
void launch_group_gemv(const BodyJob* jobs, int count, int grid, int phi, hipStream_t st);
// pressure = rho i w phi with each body's rho and frequency (omega: device
// array, one per body)
void launch_group_combine(const BodyJob* jobs, int count, int grid, const double* omega, hipStream_t st);
// fk = 0: forces into force[0, nforce nrhs); 1: Froude-Krylov forces after them
void launch_group_force(const BodyJob* jobs, int count, int grid, int fk, hipStream_t st);
// Grouped factorisation of many independent matrices: the launch sequence
// of launch_lu's paired form (rank 128, one stream), each launch doing that
// step for every matrix that has it (one workgroup range per matrix), so a
// batch of bodies of uneven sizes takes max_i ceil(n_i / 128) block pairs of
// launches instead of the sum.  Each matrix ends with the same factors, bit
// for bit, as launch_lu's.
template <typename T>
struct LuGroup {
  struct Launch {
    int kind;   // 0: first pivot inversion, 1: panel, 2: cross update, 3: rank-128 update
    int first;  // its jobs: d_jobs[first, first + count)
    int count;
    int grid;
    double flops;  // rank-128 updates: 8 x 128 x m^2 summed over the jobs (m: trailing order)
    double bytes;  // and (2 m^2 + 256 m) elements (C read and written, the two 128-wide panels read)
  };
  std::vector<Launch> launches;
  LuJob<T>* d_jobs = nullptr;  // device memory (hipFree by the owner)
};
// mats: A, P, info, ld, n of each matrix (the other fields are set here).
template <typename T>
hipError_t build_lu_group(const std::vector<LuJob<T>>& mats, LuGroup<T>* g);
// ev (optional): two events per rank-128 update launch, recorded around it
// in order; returns the number of update launches.
template <typename T>
int launch_lu_group(const LuGroup<T>& g, hipStream_t st, const hipEvent_t* ev = nullptr);
// Grouped chained solves (one launch per triangle for every system): the
// systems' TrsvJob fields except wg0 / wg1, which are set here.
template <typename T>
struct TrsvGroup {
  TrsvJob<T>* d_begin = nullptr;  // chain_begin's workgroup ranges
  TrsvJob<T>* d_tri = nullptr;    // the systems, largest first
  int* d_cum = nullptr;           // [nlevel + 1]: tickets dealt before each chain position
  unsigned* d_ticket = nullptr;   // the two triangles' launch-wide ticket counters
  int count = 0, grid_begin = 0, grid_tri = 0, nlevel = 0;
};
template <typename T>
hipError_t build_trsv_group(const std::vector<TrsvJob<T>>& sys, TrsvGroup<T>* g);
template <typename T>
void free_trsv_group(TrsvGroup<T>* g);
template <typename T>
void launch_trsv_group(const TrsvGroup<T>& g, int spin_limit, hipStream_t st);
// Solves A x = b for nrhs <= kMaxRhs right-hand sides (B: n x nrhs
// row-major) with the factors of launch_lu; on return B holds the solution,
// W is scratch of the same size.  chained: one launch per triangle whose
// workgroups take 64-row blocks in ticket order and poll the solved blocks
// they need in W (forward) and B (backward), which hold an unsolved marker
// until written (a kernel at the start of each solve resets ctl's tickets and
// marks W; status set nonzero if a bounded wait timed out), with no per-call
// argument, so the launches replay from a HIP graph; else one launch per
// block step.  Skipped when skip (optional) is done.
template <typename T>
void launch_lu_solve(const C2<T>* A, int ld, int n, const C2<T>* P, C2<T>* B, C2<T>* W, int nrhs, bool chained,
                     ChainCtl* ctl, int* status, int spin_limit, hipStream_t st, const RefineState* skip = nullptr);
// info[0] = info[1] = 0 (a kernel rather than a memset, so it replays from a
// HIP graph)
void launch_info_reset(int* info, hipStream_t st);
// Y (n x nrhs) = beta Y + alpha M X, FP64, nrhs <= kMaxRhs, all row-major;
// skipped when skip (optional) is done.
void launch_matvec(const double2* M, int ld, int n, const double2* X, int nrhs, double2* Y, double alpha,
                   double beta, hipStream_t st, const RefineState* skip = nullptr);
// The same with M complex64 (ld a multiple of 4), X, Y and the sums in FP64.
void launch_matvec32(const float2* M, int ld, int n, const double2* X, int nrhs, double2* Y, double alpha,
                     double beta, hipStream_t st, const RefineState* skip = nullptr);
void launch_rhs(const double2* V, int n, int nsym, int nrhs, double bx, double2* B, hipStream_t st);
void launch_combine(const double2* Z1, const double2* Z2, const double2* F1, const double2* F2, int n, int nsym,
                    int nrhs, double rho_w, double2* sig, double2* pot, double2* pres, hipStream_t st);
void launch_force(const double2* pres, const double* nds, int nint, int ntot, int nrhs, double2* F, hipStream_t st);
// b = a / rs->rmax (single precision); b (+)= rs->rmax a; b = a.  All skipped
// once rs is done (copy: when skip is given and done).
void launch_f64_to_f32(const double2* a, float2* b, size_t count, const RefineState* rs, hipStream_t st);
void launch_f32_to_f64(const float2* a, double2* b, size_t count, const RefineState* rs, int accumulate,
                       hipStream_t st);
void launch_copy(const double2* a, double2* b, size_t count, const RefineState* skip, hipStream_t st);
// max|r| (NaN kept) into rs: first = the right-hand side before any sweep
// (sets bmax and rmax), else the residual after a sweep (iters + 1, done
// when rmax <= tol bmax); part: 64 doubles of scratch.
void launch_refine_residual_max(const double2* r, size_t count, double* part, RefineState* rs, double tol, int first,
                                hipStream_t st);
void launch_refine_reset(RefineState* rs, hipStream_t st);
// Degree of freedom of the preProcessor's radiation / force cases: body
// (1-based), kind 1 = translation along dir, 2 = rotation about the axis
// through `axis` along dir (same layout as nemoh_dof of the C ABI).
struct Dof {
  int body, kind;
  double dir[3], axis[3];
};
// preProcessor panel geometry pg = [xm(3n) | normal(3n) | area(n)]
void launch_panel_props(const Geometry& G, double* pg, hipStream_t st);
// radiation right-hand sides into out[c * ld + col0 + q], q < nrad
void launch_radiation(const double* pg, const int* body, const Dof* dofs, int nrad, int n, int ntot, double2* out,
                      int ld, int col0, hipStream_t st);
// force weights, row-major (nforce x ntot)
void launch_nds(const double* pg, const int* body, const Dof* dofs, int nforce, int n, int ntot, double* nds,
                hipStream_t st);
// incident-wave right-hand sides into out[c * ld + b] and pressures into
// pinc[c * nbeta + b], b < nbeta (beta in radians)
void launch_diffraction(const double* pg, double k, double w, double g, double rho, double depth, double xeff,
                        double yeff, const double* beta, int nbeta, int n, int ntot, double2* out, int ld,
                        double2* pinc, hipStream_t st);
// Kochin function of the sources sig (2n x nrhs: ZIGB rows then ZIGS rows)
// at ntheta angles into H (ntheta x nrhs)
void launch_kochin(const Geometry& G, const double2* sig, int nsym, int nrhs, double k, double depth, double xeff,
                   double yeff, const double* theta, int ntheta, double2* H, hipStream_t st);
// Potential of the sources sig (ZIGB, then ZIGS when nsym = 1; nrhs
// problems) at npts field points pts[3 p + {0,1,2}] into phi[p * nrhs + q];
// bad[p] = 1 where the point has no value.  fd: the regime of *P; dhsl = 2
// Depth; a: the finite-depth amplitude factor of the frequency.
void launch_potential(const Geometry& G, const Tables& T, const FreqParams* P, int fd, double dhsl, double a,
                      const double* pts, int npts, const double2* sig, int nrhs, double2* phi, int* bad,
                      hipStream_t st);
// out (cols x rows, row-major) = transpose of in (rows x cols, row-major)
void launch_transpose_c128(const double2* in, int rows, int cols, double2* out, hipStream_t st);
// out[ntile*(ntile+1)/2] (ntile = ceil(np/256)): per-workgroup maxima of the
// squared horizontal vertex distances (GM^2)
void launch_extent(const double* x, const double* y, int np, int nsym, double* out, hipStream_t st);
void launch_mfma_probe(const double* a, const double* b, double* d, hipStream_t st);
void launch_mfma_probe_f32(const float* a, const float* b, float* d, hipStream_t st);

}  // namespace nemoh
