// C-ABI of the solver core: handle lifetime, per-frequency pipeline and the
// mapping of the reference's Solver module state onto device-resident data.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "core_internal.h"
#include "nemoh_core.h"
#include "host_math.h"
#include "kernels.h"

using namespace nemoh;

// Device memory of nemoh_core_kochin_noalloc, which allocates nothing: the
// angles (kNoallocMaxAngles doubles), the Kochin values in the kernel's
// (angle x problem) order and the same transposed to the caller's order
// (kNoallocMaxAngles x kMaxRhs complex each): 2.1 MB, held by every GPU the
// library runs on until the process ends, shared by every handle on that
// GPU.  Host code reaches it only through the address hipGetSymbolAddress
// gives for the current device (noalloc_scratch below), never by its name.
constexpr int kNoallocMaxAngles = 4096;
constexpr size_t kNoallocScratchBytes =
    kNoallocMaxAngles * sizeof(double) + 2 * (size_t)kNoallocMaxAngles * kMaxRhs * sizeof(double2);
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

struct nemoh_core {
  int np = 0, nfa = 0, nsym = 0, nb = 1;
  double zer = 0, g = 9.81, rho = 1025, depth = 0;
  double gm = 0;  // max horizontal vertex distance (theta-rule argument)
  hipStream_t st = nullptr;
  hipStream_t st2 = nullptr;  // the factorisation's look-ahead stream (bulk trailing updates)
  std::vector<hipEvent_t> lu_dep;  // st <-> st2 dependencies of one factorisation (2 per block pair)
  // NEMOH_OPT_LU_LOOKAHEAD, off: the one run with it on (its two-stream
  // launches captured into graphs by 8 request-loop threads, 256 bodies)
  // ended in a host segmentation fault whose cause is not established
  bool lookahead = false;
  // element capacity of every device buffer, keyed by the address of the
  // handle field that holds it: buffers only grow, so re-binding the handle
  // to a mesh no larger than before allocates (and frees) nothing
  std::unordered_map<const void*, size_t> caps;
  std::unordered_map<const void*, size_t> cap_bytes;  // the same in bytes
  // preProcessor side (nemoh_core_set_dofs): bodies, degrees of freedom,
  // its own panel geometry, force weights, incident-wave buffers
  bool dofs_set = false;
  int nrad = 0, nforce = 0;
  double xeff = 0, yeff = 0;
  double* d_pg = nullptr;     // [xm(3n) | normal(3n) | area(n)]
  int* d_body = nullptr;      // body of each panel (1-based)
  nemoh_dof* d_dofs = nullptr;  // nrad radiation dofs, then nforce force dofs
  double* d_fnds = nullptr;   // force weights (nforce x ntot)
  double* d_beta = nullptr;   // headings of the current call (<= 16)
  double2* d_pinc = nullptr;  // incident pressure (ntot x nbeta)
  int last_nrhs = 0;          // problems of the last solve whose sources are in sig (0: none)
  double* d_theta = nullptr;  // Kochin angles of the current call
  double2* d_kochin = nullptr;
  double amh = 0;               // AMH of the last frequency set-up (finite depth)
  double* d_field = nullptr;    // field points of the current potential call (x y z each)
  double2* d_phi = nullptr;     // their potentials (npoints x nrhs)
  int* d_fbad = nullptr;        // per point: 1 if it has no value
  double* d_pts = nullptr;  // vertex x, y (np each) for the extent reduction
  double* d_ext = nullptr;  // per-workgroup maxima of the extent reduction
  // resident geometry and tables
  double* d_geo = nullptr;  // xg yg zg xn yn zn aire tdis, n each
  double* d_vtx = nullptr;  // vx vy vz, 4n each
  int* d_surf = nullptr;    // panels with zg >= ZER (free-surface branch of the wave part)
  int nsurf = 0;
  double* d_xr = nullptr;
  double* d_xz = nullptr;
  double* d_tab = nullptr;
  float4* d_tab32 = nullptr;  // the tables in single precision (mixed assembly)
  double* d_rden = nullptr;  // reciprocal Lagrange denominators (XR, XZ stencils)
  Geometry G{};
  Tables T{};
  FreqParams* d_par = nullptr;
  FreqParams h_par{};
  // influence matrices of the current frequency (per symmetry block)
  double2* D[2] = {nullptr, nullptr};
  double2* S[2] = {nullptr, nullptr};
  // mixed path: complex64 D factored in place (rows of ld32 = nfa rounded up
  // to a multiple of 4 elements); with NEMOH_OPT_FP32_ASSEMBLY the residual's
  // D and S are complex64 too (D32r, S32), else FP64 (D, S)
  float2* D32[2] = {nullptr, nullptr};
  float2* D32r[2] = {nullptr, nullptr};
  float2* S32[2] = {nullptr, nullptr};
  int ld32 = 0;
  double2* Pinv[2] = {nullptr, nullptr};  // inverses of the pivot blocks (64x64 each)
  float2* Pinv32[2] = {nullptr, nullptr};
  double w_prev = -1.0;
  int factored_prec = -1;
  int nq_state = 0;  // module NQ of the standalone session
  // precision
  int prec = NEMOH_FP64;
  double tol = 1e-8;  // mixed refinement: max|r| <= tol max|b|
  int max_iter = 12;
  // right-hand-side workspace (capacity in right-hand sides)
  int cap = 0;
  double2 *V = nullptr, *B = nullptr, *W = nullptr, *Z[2] = {nullptr, nullptr}, *F[2] = {nullptr, nullptr};
  double2 *R = nullptr, *sig = nullptr, *pot = nullptr, *pres = nullptr;
  float2 *R32 = nullptr, *W32 = nullptr;
  double* d_nds = nullptr;
  double2* d_force = nullptr;
  double* d_red = nullptr;  // reduction scratch
  int* d_info = nullptr;     // [0] first bad pivot row, [1] solve wait timeout
  RefineState* d_rs = nullptr;  // mixed refinement state of each symmetry block
  // the chained triangular solves' control block (ChainCtl: the tickets,
  // kept on the device), then the fused cross launches' control words
  // (2 lu_pairs(n))
  int* d_chain = nullptr;
  int cu_count = 0;
  bool stepwise_solve = false;
  bool fp32_assembly = false;  // NEMOH_OPT_FP32_ASSEMBLY (mixed path only)
  int spin_limit = 1 << 20;    // polls before a chained solve's wait gives up (one to two seconds)
  // HIP graphs of the per-frequency pipeline (NEMOH_OPT_GRAPHS), keyed by
  // what shapes the launch sequence; layout_gen moves on whenever a device
  // buffer is replaced or the handle is bound to another mesh, which
  // invalidates every graph (their kernel arguments hold the old pointers)
  bool use_graphs = true;
  struct GraphEntry {
    int refactor, mode, fd, nrhs, stepwise, form;
    int uses;               // launches of this shape since the graphs were last dropped
    hipGraphExec_t exec;    // captured on the kGraphAfter-th use
  };
  std::vector<GraphEntry> graphs;
  unsigned layout_gen = 0, graphs_gen = 0;
  // statistics
  bool timing = false;
  nemoh_stats stats{};
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<hipEvent_t> upd_ev;  // around each trailing-update launch (timing on)
  std::vector<double> upd_flops;   // their algorithmic flops
  int upd_launches = 0;            // how many upd_ev pairs the last factorisation stamped
  int lu_rank = 128;               // NEMOH_OPT_LU_RANK: rank of the trailing updates (64 or 128)
  int lu_band = 16;                // NEMOH_OPT_LU_BAND: tile bands of the trailing update (16, 32, 64)
  // NEMOH_OPT_LU_TUNED: the factorisation form comes from the autotuner's
  // table for this precision and size class when it has an entry
  bool lu_tuned = false;
  int lu_rank_used = 0, lu_band_used = 0;  // form of the last factorisation
  // NEMOH_OPT_LU_FUSED_PANEL: second panel inside the cross launch (off:
  // measured equal to the separate launch, 3.36 ms per 4000-panel
  // factorisation either way, without the in-launch wait)
  bool lu_fused = false;
  // NEMOH_OPT_PAIRED_ASSEMBLY: the wave part of each unordered panel pair
  // once for both matrix elements
  bool paired_assembly = true;
  unsigned char* d_scratch = nullptr;  // this device's `scratch` (nemoh_core_kochin_noalloc)
  std::string err;
};

namespace {

// Graph capture and device-memory management never overlap in this process:
// a capture holds this lock from its beginning to its end, and so does every
// allocation, free and synchronous set-up call of the library, which HIP may
// refuse (or which may invalidate the capture) while another thread's stream
// is capturing.  Captures happen once per handle and launch shape, and take
// about a millisecond.
std::recursive_mutex& device_lock() {
  static std::recursive_mutex m;
  return m;
}

// The autotuner's choices (nemoh_core_autotune): factorisation form (update
// rank, tile-band height) by precision and number of 64-row blocks, for the
// handles of this process that opt in (NEMOH_OPT_LU_TUNED).
struct TunedForm {
  int rank, band;
};
std::mutex& tune_mutex() {
  static std::mutex m;
  return m;
}
std::map<std::pair<int, int>, TunedForm>& tune_table() {
  static std::map<std::pair<int, int>, TunedForm> t;
  return t;
}

// The last creation error of this thread (nemoh_core_last_error(NULL)).
thread_local std::string create_error;

// The device address of `scratch` on device dev, looked up once per device
// with that device current (each GPU holds its own copy of the array).
hipError_t noalloc_scratch(int dev, unsigned char** out) {
  static std::mutex mu;
  static std::unordered_map<int, unsigned char*> addr;
  std::lock_guard<std::mutex> lk(mu);
  auto it = addr.find(dev);
  if (it != addr.end()) {
    *out = it->second;
    return hipSuccess;
  }
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return e != hipSuccess ? e : hipErrorInvalidSymbol;
  addr[dev] = static_cast<unsigned char*>(p);
  *out = static_cast<unsigned char*>(p);
  return hipSuccess;
}

int fail(nemoh_core* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(c, call)                                                                    \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail((c), NEMOH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Makes *p hold at least count elements (a buffer is only ever replaced by
// a larger one).  The whole call holds the device lock, the capacity check
// too: with the check outside it (a handle's records are used by one thread
// at a time) the 4000-panel throughput rose 2 % but the 256-body batch fell
// from about 910 to 720-760 freq-solves/s and the request loop's latency
// grew, as workers kept launching through each other's captures and binds.
template <typename T>
int dalloc(nemoh_core* c, T** p, size_t count) {
  std::lock_guard<std::recursive_mutex> lk(device_lock());
  size_t& have = c->caps[(const void*)p];
  size_t& bytes = c->cap_bytes[(const void*)p];
  if (*p && have >= count) return NEMOH_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  have = bytes = 0;
  HIPCHK(c, hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T)));
  have = count;
  bytes = std::max<size_t>(count, 1) * sizeof(T);
  ++c->layout_gen;
  return NEMOH_OK;
}

template <typename T>
void dfree(nemoh_core* c, T** p) {
  std::lock_guard<std::recursive_mutex> lk(device_lock());
  if (*p) (void)hipFree(*p);
  if (*p) ++c->layout_gen;
  *p = nullptr;
  c->caps[(const void*)p] = 0;
  c->cap_bytes[(const void*)p] = 0;
}

// The mixed path with NEMOH_OPT_FP32_ASSEMBLY: complex64 D and S.
bool c64_matrices(const nemoh_core* c) { return c->prec == NEMOH_MIXED && c->fp32_assembly; }

int ensure_matrices(nemoh_core* c) {
  const size_t ninv = (size_t)((c->nfa + kNB - 1) / kNB) * kNB * kNB;
  const size_t ndep = 2 * (size_t)((c->nfa + 2 * kNB - 1) / (2 * kNB));  // look-ahead events: 2 per block pair
  while (c->lu_dep.size() < ndep) {
    hipEvent_t e;
    HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    c->lu_dep.push_back(e);
  }
  for (int b = 0; b < c->nb; ++b) {
    if (c64_matrices(c)) {
      const size_t n2 = (size_t)c->nfa * c->ld32;
      if (int r = dalloc(c, &c->D32r[b], n2)) return r;
      if (int r = dalloc(c, &c->S32[b], n2)) return r;
    } else {
      const size_t n2 = (size_t)c->nfa * c->nfa;
      if (int r = dalloc(c, &c->D[b], n2)) return r;
      if (int r = dalloc(c, &c->S[b], n2)) return r;
    }
    if (c->prec == NEMOH_MIXED) {
      if (int r = dalloc(c, &c->D32[b], (size_t)c->nfa * c->ld32)) return r;
      if (int r = dalloc(c, &c->Pinv32[b], ninv)) return r;
    } else {
      if (int r = dalloc(c, &c->Pinv[b], ninv)) return r;
    }
  }
  return NEMOH_OK;
}

// r = b - D z (FP64 sums) and f = S z with the matrices of the current mode
void residual(nemoh_core* c, int b, int nrhs, const RefineState* skip) {
  if (c64_matrices(c))
    launch_matvec32(c->D32r[b], c->ld32, c->nfa, c->Z[b], nrhs, c->R, -1.0, 1.0, c->st, skip);
  else
    launch_matvec(c->D[b], c->nfa, c->nfa, c->Z[b], nrhs, c->R, -1.0, 1.0, c->st, skip);
}
void potential_of_block(nemoh_core* c, int b, int nrhs) {
  if (c64_matrices(c))
    launch_matvec32(c->S32[b], c->ld32, c->nfa, c->Z[b], nrhs, c->F[b], 1.0, 0.0, c->st);
  else
    launch_matvec(c->S[b], c->nfa, c->nfa, c->Z[b], nrhs, c->F[b], 1.0, 0.0, c->st);
}

// cap = right-hand sides the buffers hold for the current mesh (reset to 0
// when the handle is bound to another mesh)
int ensure_rhs(nemoh_core* c, int nrhs) {
  if (nrhs <= c->cap) return NEMOH_OK;
  c->cap = 0;
  const size_t n = c->nfa, nt = (size_t)c->nfa * c->nb;
  const size_t k = nrhs;
  int r = 0;
  if ((r = dalloc(c, &c->V, nt * k)) || (r = dalloc(c, &c->B, n * k)) || (r = dalloc(c, &c->W, n * k)) ||
      (r = dalloc(c, &c->R, n * k)) || (r = dalloc(c, &c->sig, 2 * n * k)) || (r = dalloc(c, &c->pot, 2 * n * k)) ||
      (r = dalloc(c, &c->pres, 2 * n * k)) || (r = dalloc(c, &c->R32, n * k)) || (r = dalloc(c, &c->W32, n * k)))
    return r;
  for (int b = 0; b < c->nb; ++b)
    if ((r = dalloc(c, &c->Z[b], n * k)) || (r = dalloc(c, &c->F[b], n * k))) return r;
  c->cap = nrhs;
  return NEMOH_OK;
}

// FreqParams for one call: the per-frequency scalars of VNSINFD / VNSFD.
int build_params(nemoh_core* c, double period, double amh, int nexp, const double* ambda, const double* ar, int nq,
                 const double* qq, const double* cq, bool* fd_out, double* w_out) {
  const double pi = 4.0 * std::atan(1.0);
  const double w = 2.0 * pi / period;
  if (!(std::fabs(w) >= 1.E-4)) return fail(c, NEMOH_ERR_OMEGA, "ABS(WR) < 1.E-4");
  const double kwave = w * w / c->g;
  const bool fd = !(c->depth == 0.0 || kwave * c->depth >= 20);  // SOLVE_BEM.f90:66
  c->last_nrhs = 0;  // the sources in sig belong to the previous parameters
  c->amh = amh;
  FreqParams& P = c->h_par;
  std::memset(&P, 0, sizeof P);
  P.fd = fd ? 1 : 0;
  P.nsym = c->nsym;
  P.zer = c->zer;
  P.depth = c->depth;
  if (!fd) {
    P.ak0 = w * w / c->g;
    P.pik = M_PI / P.ak0;
    P.pik3 = M_PI / (P.ak0 * P.ak0 * P.ak0);
    if (nq < 0 || nq > 101 || (nq > 0 && (!qq || !cq))) return fail(c, NEMOH_ERR_ARG, "nq out of range");
    P.nq = nq;
    for (int k = 0; k < nq; ++k) {
      P.qq[k] = qq[k];
      P.cq[k] = cq[k];
    }
  } else {
    if (nexp < 0 || nexp > 30 || (nexp > 0 && (!ambda || !ar))) return fail(c, NEMOH_ERR_ARG, "nexp out of range");
    const double h = c->depth;
    const double akh = amh * std::tanh(amh);
    const double a = (amh + akh) * (amh + akh) / (h * (amh * amh - akh * akh + akh));
    const double am0 = amh / h;
    const double adpi2 = -a / (8. * pi * pi), adpi = -a / (8 * pi);
    P.ak0 = am0;
    P.pik = M_PI / am0;
    P.pik3 = P.pik / (am0 * am0);
    P.coe1 = adpi2 / am0;
    P.coe2 = adpi / am0;
    P.coe3 = adpi2;
    P.coe4 = adpi;
    for (int k = 0; k < nexp; ++k) {
      P.ambda[k] = ambda[k];
      P.ar[k] = ar[k];
    }
    P.ambda[nexp] = 0.0;  // VNSFD.f90:262-263
    P.ar[nexp] = 2.0;
    P.nexp1 = nexp + 1;
  }
  *fd_out = fd;
  *w_out = w;
  HIPCHK(c, hipMemcpyAsync(c->d_par, &P, sizeof P, hipMemcpyHostToDevice, c->st));
  return NEMOH_OK;
}

int check_info(nemoh_core* c) {
  int h[2] = {0, 0};
  HIPCHK(c, hipMemcpyAsync(h, c->d_info, 2 * sizeof(int), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (h[1] & 2) {
    c->w_prev = -1.0;
    return fail(c, NEMOH_ERR_HIP, "factorisation: a bounded wait for an inverted pivot block timed out");
  }
  if (h[1]) {
    c->w_prev = -1.0;
    return fail(c, NEMOH_ERR_HIP, "triangular solve: a bounded wait for a solved block timed out");
  }
  const int info = h[0];
  if (info) {
    c->w_prev = -1.0;
    char buf[96];
    std::snprintf(buf, sizeof buf, "PIVOT INFERIEUR A 1e-20 (row %d)", info);
    return fail(c, NEMOH_ERR_PIVOT, buf);
  }
  return NEMOH_OK;
}

// What the resident matrices hold: precision, and complex64 D and S.
int matrix_mode(const nemoh_core* c) { return c->prec + (c64_matrices(c) ? 2 : 0); }

// The factorisation form the handle uses now: the tuner's entry for its
// precision and size class when it opted in and one exists, else its own.
TunedForm lu_form(const nemoh_core* c) {
  TunedForm f{c->lu_rank, c->lu_band};
  if (c->lu_tuned) {
    std::lock_guard<std::mutex> lk(tune_mutex());
    auto it = tune_table().find({c->prec, (c->nfa + kNB - 1) / kNB});
    if (it != tune_table().end()) f = it->second;
  }
  return f;
}

// Queues assembly + factorisation of the current frequency (matrices
// allocated; no host wait and, without timing, no event: capturable).
int queue_factor(nemoh_core* c) {
  if (c->timing) HIPCHK(c, hipEventRecord(c->ev[0], c->st));
  if (c64_matrices(c)) {
    const MatOut32 out{c->ld32, c->D32r[0], c->D32[0], c->S32[0], c->D32r[1], c->D32[1], c->S32[1]};
    launch_assemble_mixed(c->G, c->T, c->d_tab32, c->d_par, c->h_par.fd, c->nsym, out, c->d_surf, c->nsurf, c->st);
  } else {
    MatOut out{c->nfa, c->D[0], c->S[0], c->D[1], c->S[1], nullptr, nullptr};
    if (c->prec == NEMOH_MIXED) {  // the copy of D that the complex64 LU factors in place
      out.dp32 = c->D32[0];
      out.dm32 = c->D32[1];
      out.ld32 = c->ld32;
    }
    launch_assemble(c->G, c->T, c->d_par, c->h_par.fd, c->nsym, out, c->d_surf, c->nsurf, c->paired_assembly,
                    c->st);
  }
  HIPCHK(c, hipGetLastError());
  if (c->timing) HIPCHK(c, hipEventRecord(c->ev[1], c->st));
  launch_info_reset(c->d_info, c->st);
  const int nblk = (c->nfa + kNB - 1) / kNB;
  c->upd_launches = 0;
  if (c->timing) {
    const size_t need = 2 * (size_t)nblk * c->nb;
    while (c->upd_ev.size() < need) {
      hipEvent_t e;
      HIPCHK(c, hipEventCreate(&e));
      c->upd_ev.push_back(e);
    }
    c->upd_flops.resize(need);
  }
  // dependency events of the look-ahead form (2 per block pair, created by
  // ensure_matrices, outside any graph capture)
  hipStream_t st2 = c->lookahead && c->lu_dep.size() >= 2 * (size_t)((nblk + 1) / 2) ? c->st2 : nullptr;
  const TunedForm form = lu_form(c);
  c->lu_rank_used = form.rank;
  c->lu_band_used = form.band;
  for (int b = 0; b < c->nb; ++b) {
    hipEvent_t* ue = c->timing ? c->upd_ev.data() + 2 * (size_t)c->upd_launches : nullptr;
    double* uf = c->timing ? c->upd_flops.data() + c->upd_launches : nullptr;
    // polls of a fused cross launch's wait for its pivot block (one to two
    // seconds; not the solves' spin limit, which tests set to 0)
    constexpr int kLuSpinLimit = 1 << 20;
    int* luctl = c->lu_fused ? c->d_chain + 4 : nullptr;
    if (c->prec == NEMOH_MIXED)
      c->upd_launches += launch_lu<float>(reinterpret_cast<C2<float>*>(c->D32[b]), c->ld32, c->nfa,
                                          reinterpret_cast<C2<float>*>(c->Pinv32[b]), c->d_info, c->st, form.rank,
                                          st2, c->lu_dep.data(), ue, uf, luctl, kLuSpinLimit, form.band);
    else
      c->upd_launches += launch_lu<double>(reinterpret_cast<C2<double>*>(c->D[b]), c->nfa, c->nfa,
                                           reinterpret_cast<C2<double>*>(c->Pinv[b]), c->d_info, c->st, form.rank,
                                           st2, c->lu_dep.data(), ue, uf, luctl, kLuSpinLimit, form.band);
  }
  HIPCHK(c, hipGetLastError());
  if (c->timing) HIPCHK(c, hipEventRecord(c->ev[2], c->st));
  // the pivot flag is read back after the solves, so the host never waits
  // between factorisation and solves
  return NEMOH_OK;
}

// A x = b in place in B (W scratch): one flag-chained launch per triangle,
// or one launch per block step (forced by NEMOH_OPT_STEPWISE_SOLVE, or after
// a chained solve's bounded wait timed out).  skip: a refinement state whose
// `done` turns the launches into no-ops.
template <typename T>
void lu_solve(nemoh_core* c, const C2<T>* A, int ld, const C2<T>* P, C2<T>* B, C2<T>* W, int nrhs, bool stepwise,
              const RefineState* skip = nullptr) {
  const int n = c->nfa;
  static_assert(sizeof(ChainCtl) == 4 * sizeof(int), "ChainCtl takes the four ints of d_chain");
  launch_lu_solve<T>(A, ld, n, P, B, W, nrhs, !stepwise, reinterpret_cast<ChainCtl*>(c->d_chain),
                     c->d_info + 1, c->spin_limit, c->st, skip);
}

// Mixed path, one refinement sweep of symmetry block b, every kernel a no-op
// once the block's state is done: x (+)= rmax A32^-1 (r / rmax) in complex64,
// then the FP64 residual r = b - D x and its convergence test.
void mixed_sweep(nemoh_core* c, int b, int nrhs, bool first, bool stepwise) {
  const int n = c->nfa;
  const size_t cnt = (size_t)n * nrhs;
  RefineState* rs = c->d_rs + b;
  launch_f64_to_f32(c->R, c->R32, cnt, rs, c->st);
  lu_solve(c, reinterpret_cast<const C2<float>*>(c->D32[b]), c->ld32,
           reinterpret_cast<const C2<float>*>(c->Pinv32[b]), reinterpret_cast<C2<float>*>(c->R32),
           reinterpret_cast<C2<float>*>(c->W32), nrhs, stepwise, rs);
  launch_f32_to_f64(c->R32, c->Z[b], cnt, rs, first ? 0 : 1, c->st);
  launch_copy(c->B, c->R, cnt, rs, c->st);
  residual(c, b, nrhs, rs);
  launch_refine_residual_max(c->R, cnt, c->d_red, rs, c->tol, 0, c->st);
}

// Sweeps queued per block before the host looks at convergence: two is what
// the 1e-8 backward error takes on every workload measured (the host adds
// sweeps one at a time when a block needs more).
constexpr int kQueuedSweeps = 2;

// Solves symmetry block b for the right-hand sides in c->B (n x nrhs,
// row-major) into c->Z[b] and forms phi = S sigma into c->F[b].
void solve_block(nemoh_core* c, int b, int nrhs, bool stepwise) {
  const int n = c->nfa;
  const size_t cnt = (size_t)n * nrhs;
  if (c->prec == NEMOH_FP64) {
    launch_copy(c->B, c->Z[b], cnt, nullptr, c->st);
    lu_solve(c, reinterpret_cast<const C2<double>*>(c->D[b]), n, reinterpret_cast<const C2<double>*>(c->Pinv[b]),
             reinterpret_cast<C2<double>*>(c->Z[b]), reinterpret_cast<C2<double>*>(c->W), nrhs, stepwise);
  } else {
    RefineState* rs = c->d_rs + b;
    launch_refine_reset(rs, c->st);
    launch_copy(c->B, c->R, cnt, nullptr, c->st);
    launch_refine_residual_max(c->B, cnt, c->d_red, rs, c->tol, 1, c->st);
    for (int k = 0; k < kQueuedSweeps && k < c->max_iter; ++k) mixed_sweep(c, b, nrhs, k == 0, stepwise);
  }
  potential_of_block(c, b, nrhs);
}

// The pivot report first, then a non-finite or unconverged refinement.
int refine_outcome(nemoh_core* c, const RefineState* h, int pivot_row) {
  if (pivot_row) {
    c->w_prev = -1.0;
    char buf[96];
    std::snprintf(buf, sizeof buf, "PIVOT INFERIEUR A 1e-20 (row %d)", pivot_row);
    return fail(c, NEMOH_ERR_PIVOT, buf);
  }
  if (c->prec != NEMOH_MIXED) return NEMOH_OK;
  for (int b = 0; b < c->nb; ++b) {
    c->stats.refine_iterations = std::max(c->stats.refine_iterations, h[b].iters);
    if (h[b].bmax > 0) c->stats.refine_residual = std::fmax(c->stats.refine_residual, h[b].rmax / h[b].bmax);
  }
  for (int b = 0; b < c->nb; ++b) {
    char buf[160];
    if (h[b].bad) {
      std::snprintf(buf, sizeof buf, "refinement met a non-finite %s (max = %g)",
                    h[b].iters == 0 && h[b].bmax == 0 ? "right-hand side" : "residual", h[b].rmax);
      return fail(c, NEMOH_ERR_CONVERGENCE, buf);
    }
    if (!h[b].done) {
      std::snprintf(buf, sizeof buf, "refinement stopped at max|r|/max|b| = %.3e after %d iterations",
                    h[b].bmax > 0 ? h[b].rmax / h[b].bmax : 0.0, h[b].iters);
      return fail(c, NEMOH_ERR_CONVERGENCE, buf);
    }
  }
  return NEMOH_OK;
}

// Queues the right-hand sides and solves of every symmetry block, with phi
// (no host wait: capturable).
int queue_solves(nemoh_core* c, int nrhs, bool stepwise) {
  if (c->timing) HIPCHK(c, hipEventRecord(c->ev[2], c->st));
  for (int b = 0; b < c->nb; ++b) {
    launch_rhs(c->V, c->nfa, c->nsym, nrhs, b == 0 ? 1.0 : -1.0, c->B, c->st);
    solve_block(c, b, nrhs, stepwise);
  }
  HIPCHK(c, hipGetLastError());
  return NEMOH_OK;
}

// Uses of a launch shape before the handle captures it in a HIP graph: a
// capture and instantiation cost about as much as several small-body
// frequencies, so shapes met once or twice (a batch of distinct bodies on a
// request-loop worker) run eagerly.
constexpr int kGraphAfter = 3;

void drop_graphs(nemoh_core* c) {
  for (auto& g : c->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  c->graphs.clear();
  c->graphs_gen = c->layout_gen;
}

// Assembly + factorisation (when refactor) and the solves of one frequency:
// eagerly with timing on or graphs off, else by replaying the handle's HIP
// graph of that launch sequence, captured on its first use.  The frequency's
// parameters (d_par) and right-hand sides (V) are in device memory, which
// the graph reads, and every kernel argument is fixed for a given binding,
// precision, regime, problem count and solve form.
int launch_frequency(nemoh_core* c, bool refactor, int nrhs, bool stepwise) {
  auto eager = [&]() -> int {
    if (refactor)
      if (int r = queue_factor(c)) return r;
    return queue_solves(c, nrhs, stepwise);
  };
  if (!c->use_graphs || c->timing) return eager();
  if (c->graphs_gen != c->layout_gen) drop_graphs(c);
  const TunedForm form = lu_form(c);
  const int key[6] = {refactor ? 1 : 0, matrix_mode(c), c->h_par.fd, nrhs, stepwise ? 1 : 0,
                      form.rank * 1000 + form.band};
  nemoh_core::GraphEntry* g = nullptr;
  for (auto& e : c->graphs)
    if (e.refactor == key[0] && e.mode == key[1] && e.fd == key[2] && e.nrhs == key[3] && e.stepwise == key[4] &&
        e.form == key[5])
      g = &e;
  if (!g) {
    c->graphs.push_back({key[0], key[1], key[2], key[3], key[4], key[5], 0, nullptr});
    g = &c->graphs.back();
  }
  if (!g->exec && ++g->uses < kGraphAfter) return eager();
  if (!g->exec) {
    std::lock_guard<std::recursive_mutex> lk(device_lock());
    HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    int r = refactor ? queue_factor(c) : NEMOH_OK;
    if (!r) r = queue_solves(c, nrhs, stepwise);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(c->st, &graph);
    if (r) {
      if (graph) (void)hipGraphDestroy(graph);
      return r;
    }
    if (e != hipSuccess) return fail(c, NEMOH_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(c, NEMOH_ERR_HIP, std::string("graph instantiation: ") + hipGetErrorString(ei));
    g->exec = exec;
  }
  HIPCHK(c, hipGraphLaunch(g->exec, c->st));
  ++c->stats.graph_replays;
  return NEMOH_OK;
}

// Everything after the frequency set-up: c->V holds (ntot x nrhs) row-major
// normal velocities on the device.  The solves of every symmetry block, phi
// and the combination are queued without a host wait; one wait at the end
// reads the pivot report, the chained solves' timeout flag and the mixed
// refinement states.  A block that needs more sweeps gets them one at a time;
// a timed-out chained solve (which the ticket order makes unexpected) is
// redone with one launch per block step.
int run_pipeline(nemoh_core* c, double w, int nrhs) {
  const int mode = matrix_mode(c);
  const bool refactor = !(w == c->w_prev && c->factored_prec == mode);  // the factors cached per frequency
  c->stats.cache_hit = refactor ? 0 : 1;
  if (refactor)
    if (int r = ensure_matrices(c)) return r;
  const int n = c->nfa;
  c->stats.refine_iterations = 0;
  c->stats.refine_residual = 0;
  bool stepwise = c->stepwise_solve;
  const size_t cnt = (size_t)n * nrhs;
  for (int attempt = 0;; ++attempt) {
    if (int r = launch_frequency(c, attempt == 0 && refactor, nrhs, stepwise)) return r;
    if (attempt == 0 && refactor) {
      c->w_prev = w;
      c->factored_prec = mode;
    }
    RefineState h[2] = {};
    int info[2] = {0, 0};
    for (;;) {
      HIPCHK(c, hipMemcpyAsync(info, c->d_info, sizeof info, hipMemcpyDeviceToHost, c->st));
      if (c->prec == NEMOH_MIXED)
        HIPCHK(c, hipMemcpyAsync(h, c->d_rs, c->nb * sizeof(RefineState), hipMemcpyDeviceToHost, c->st));
      HIPCHK(c, hipStreamSynchronize(c->st));
      if (info[0] || info[1] || c->prec != NEMOH_MIXED) break;
      int more = -1;
      for (int b = 0; b < c->nb; ++b)
        if (!h[b].done && !h[b].bad && h[b].iters < c->max_iter) more = b;
      if (more < 0) break;
      // one more sweep of that block: its right-hand side and residual again
      // (the buffers were reused by the blocks after it), then phi
      launch_rhs(c->V, n, c->nsym, nrhs, more == 0 ? 1.0 : -1.0, c->B, c->st);
      launch_copy(c->B, c->R, cnt, nullptr, c->st);
      residual(c, more, nrhs, nullptr);
      mixed_sweep(c, more, nrhs, false, stepwise);
      potential_of_block(c, more, nrhs);
      HIPCHK(c, hipGetLastError());
    }
    if (info[1] & 2) {  // a fused cross launch gave up waiting for its pivot block: the factors are unfinished
      c->w_prev = -1.0;
      return fail(c, NEMOH_ERR_HIP, "factorisation: a bounded wait for an inverted pivot block timed out");
    }
    if (info[1] && !info[0] && attempt == 0 && !stepwise) {
      const int zero = 0;
      HIPCHK(c, hipMemcpyAsync(c->d_info + 1, &zero, sizeof zero, hipMemcpyHostToDevice, c->st));
      stepwise = true;
      ++c->stats.solve_fallbacks;
      continue;
    }
    if (info[1]) {
      c->w_prev = -1.0;
      return fail(c, NEMOH_ERR_HIP, "triangular solve: a bounded wait for a solved block timed out");
    }
    if (int r = refine_outcome(c, h, info[0])) return r;
    break;
  }
  launch_combine(c->Z[0], c->Z[1], c->F[0], c->F[1], n, c->nsym, nrhs, c->rho * w, c->sig, c->pot, c->pres, c->st);
  HIPCHK(c, hipGetLastError());
  c->last_nrhs = nrhs;
  if (c->timing) HIPCHK(c, hipEventRecord(c->ev[3], c->st));
  return NEMOH_OK;
}

int upload_nvel_colmajor(nemoh_core* c, const double* nvel, int nrhs) {
  // caller: (ntot, nrhs) column-major complex; device: (ntot x nrhs) row-major
  const size_t nt = (size_t)c->nfa * c->nb;
  std::vector<double> tmp(2 * nt * nrhs);
  for (int q = 0; q < nrhs; ++q)
    for (size_t i = 0; i < nt; ++i) {
      tmp[2 * (i * nrhs + q)] = nvel[2 * (q * nt + i)];
      tmp[2 * (i * nrhs + q) + 1] = nvel[2 * (q * nt + i) + 1];
    }
  HIPCHK(c, hipMemcpyAsync(c->V, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return NEMOH_OK;
}

}  // namespace

// nds staging is kept outside the anonymous namespace helpers above
static int stage_nds(nemoh_core* c, int nint, const double* nds) {
  const size_t nt = (size_t)c->nfa * c->nb;
  const size_t need = (size_t)nint * nt;
  if (int r = dalloc(c, &c->d_nds, need)) return r;
  // caller: NDS(nint, ntot) column-major -> device (nint x ntot) row-major
  std::vector<double> tmp(need);
  for (int k = 0; k < nint; ++k)
    for (size_t i = 0; i < nt; ++i) tmp[k * nt + i] = nds[k + (size_t)nint * i];
  HIPCHK(c, hipMemcpyAsync(c->d_nds, tmp.data(), need * sizeof(double), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return NEMOH_OK;
}

// F(k, q) = -sum_c P(c, q) nds(k, c) on the device (P: ntot x nrhs, nds:
// nint x ntot, both row-major), written to the caller column-major.
static int download_forces(nemoh_core* c, const double2* P, const double* d_nds, int nint, int nrhs, double* out) {
  const size_t nt = (size_t)c->nfa * c->nb;
  if (int r = dalloc(c, &c->d_force, (size_t)nint * nrhs)) return r;
  launch_force(P, d_nds, nint, (int)nt, nrhs, c->d_force, c->st);
  HIPCHK(c, hipGetLastError());
  std::vector<double> f(2 * (size_t)nint * nrhs);
  HIPCHK(c, hipMemcpyAsync(f.data(), c->d_force, f.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  for (int k = 0; k < nint; ++k)  // device (nint x nrhs) row-major -> caller column-major
    for (int q = 0; q < nrhs; ++q) {
      out[2 * (k + (size_t)nint * q)] = f[2 * (k * nrhs + q)];
      out[2 * (k + (size_t)nint * q) + 1] = f[2 * (k * nrhs + q) + 1];
    }
  return NEMOH_OK;
}

// End of a solve call: wait, report a pivot or wait failure, timing stats.
static int finish_call(nemoh_core* c) {
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (int r = check_info(c)) return r;
  if (c->timing) {
    if (!c->stats.cache_hit) {
      float a = 0, f = 0;
      HIPCHK(c, hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
      HIPCHK(c, hipEventElapsedTime(&f, c->ev[1], c->ev[2]));
      c->stats.ms_assembly = a;
      c->stats.ms_factor = f;
      // trailing-update kernels: summed durations and algorithmic flops
      double ms = 0, flops = 0;
      for (int t = 0; t < c->upd_launches; ++t) {
        float e = 0;
        HIPCHK(c, hipEventElapsedTime(&e, c->upd_ev[2 * (size_t)t], c->upd_ev[2 * (size_t)t + 1]));
        ms += e;
        flops += c->upd_flops[t];
      }
      c->stats.ms_lu_update = ms;
      c->stats.lu_update_launches = c->upd_launches;
      c->stats.lu_update_flops = flops;
      c->stats.lu_precision = c->factored_prec & 1;
    }
    float s = 0, t = 0;
    HIPCHK(c, hipEventElapsedTime(&s, c->ev[2], c->ev[3]));
    c->stats.ms_solve = s;
    HIPCHK(c, hipEventElapsedTime(&t, c->ev[c->stats.cache_hit ? 2 : 0], c->ev[3]));
    c->stats.ms_total = t;
  }
  return NEMOH_OK;
}

static int finish_outputs(nemoh_core* c, int nrhs, double* pressure, int nint, const double* nds, double* forces) {
  const size_t nt = (size_t)c->nfa * c->nb;
  if (forces && nint > 0) {
    if (!nds) return fail(c, NEMOH_ERR_ARG, "nds is NULL");
    if (int r = stage_nds(c, nint, nds)) return r;
    if (int r = download_forces(c, c->pres, c->d_nds, nint, nrhs, forces)) return r;
  }
  if (pressure) {
    std::vector<double> p(2 * nt * nrhs);
    HIPCHK(c, hipMemcpyAsync(p.data(), c->pres, p.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (size_t i = 0; i < nt; ++i)
      for (int q = 0; q < nrhs; ++q) {
        pressure[2 * (i + nt * q)] = p[2 * (i * nrhs + q)];
        pressure[2 * (i + nt * q) + 1] = p[2 * (i * nrhs + q) + 1];
      }
  }
  return finish_call(c);
}

// Body conditions of one frequency into c->V (ntot x (nbeta + nrad),
// headings first) and the incident pressures into c->d_pinc.
static int generate_waves(nemoh_core* c, double omega, int nbeta, const double* beta, bool wait = true) {
  if (!c->dofs_set) return fail(c, NEMOH_ERR_ARG, "nemoh_core_set_dofs has not been called for this mesh");
  const int nrhs = nbeta + c->nrad;
  if (nbeta < 0 || nrhs < 1 || nrhs > kMaxRhs || (nbeta > 0 && !beta))
    return fail(c, NEMOH_ERR_ARG, "need 1 <= headings + radiation dofs <= 16");
  if (!(std::fabs(omega) >= 1.E-4)) return fail(c, NEMOH_ERR_OMEGA, "ABS(WR) < 1.E-4");
  const double k = host::incident_wavenumber(omega, c->depth, c->g);
  if (std::isnan(k)) return fail(c, NEMOH_ERR_WAVENUMBER, "Error: unable to find the wavenumber");
  if (int r = ensure_rhs(c, nrhs)) return r;
  const int n = c->nfa, nt = n * c->nb;
  if (nbeta > 0) {
    HIPCHK(c, hipMemcpyAsync(c->d_beta, beta, nbeta * sizeof(double), hipMemcpyHostToDevice, c->st));
    launch_diffraction(c->d_pg, k, omega, c->g, c->rho, c->depth, c->xeff, c->yeff, c->d_beta, nbeta, n, nt, c->V,
                       nrhs, c->d_pinc, c->st);
  }
  launch_radiation(c->d_pg, c->d_body, reinterpret_cast<const Dof*>(c->d_dofs), c->nrad, n, nt, c->V, nrhs, nbeta,
                   c->st);
  HIPCHK(c, hipGetLastError());
  // the headings are read from host memory by the copy above before return
  // (a batch waits once for all its bodies instead)
  if (wait) HIPCHK(c, hipStreamSynchronize(c->st));
  return NEMOH_OK;
}

static bool mesh_args_ok(int np, int nfa, int nsym, const double* x, const double* y, const double* z,
                         const int* m1, const int* m2, const int* m3, const int* m4, const double* const geo[8],
                         double depth) {
  if (np <= 0 || nfa <= 0 || (nsym != 0 && nsym != 1) || !x || !y || !z || !m1 || !m2 || !m3 || !m4 || depth < 0)
    return false;
  for (int k = 0; k < 8; ++k)
    if (!geo[k]) return false;
  return true;
}

namespace nemoh {

int core_bind_mesh(nemoh_core* c, int np, int nfa, int nsym, const double* x, const double* y, const double* z,
                   const int* m1, const int* m2, const int* m3, const int* m4, const double* const geo[8], double zer,
                   double g, double rho, double depth) {
  if (!c) return NEMOH_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(device_lock());
  if (!mesh_args_ok(np, nfa, nsym, x, y, z, m1, m2, m3, m4, geo, depth))
    return fail(c, NEMOH_ERR_ARG, "bad mesh arguments");
  std::vector<double> vtx(12 * (size_t)nfa);
  for (int j = 0; j < nfa; ++j) {
    const int m[4] = {m1[j] - 1, m2[j] - 1, m3[j] - 1, m4[j] - 1};
    for (int l = 0; l < 4; ++l) {
      if (m[l] < 0 || m[l] >= np) return fail(c, NEMOH_ERR_ARG, "panel vertex index out of range");
      vtx[4 * (size_t)j + l] = x[m[l]];
      vtx[4 * (size_t)nfa + 4 * (size_t)j + l] = y[m[l]];
      vtx[8 * (size_t)nfa + 4 * (size_t)j + l] = z[m[l]];
    }
  }
  // nothing queued on the handle may still read the buffers being replaced
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->w_prev = -1.0;
  c->factored_prec = -1;
  const int ntile = (np + 255) / 256;
  const size_t nred = (size_t)ntile * (ntile + 1) / 2;
  // panels whose pairs take the free-surface branch of the wave part
  // (the reference's test ZG < ZER fails for them)
  std::vector<int> surf;
  for (int j = 0; j < nfa; ++j)
    if (!(geo[2][j] < zer)) surf.push_back(j);
  int r = 0;
  if ((r = dalloc(c, &c->d_geo, 8 * (size_t)nfa)) || (r = dalloc(c, &c->d_vtx, 12 * (size_t)nfa)) ||
      (r = dalloc(c, &c->d_surf, surf.size())) ||
      (r = dalloc(c, &c->d_chain, 4 + 2 * (size_t)lu_pairs(nfa))) || (r = dalloc(c, &c->d_pts, 2 * (size_t)np)) ||
      (r = dalloc(c, &c->d_ext, nred)))
    return r;
  for (int k = 0; k < 8; ++k)
    HIPCHK(c, hipMemcpyAsync(c->d_geo + (size_t)k * nfa, geo[k], nfa * sizeof(double), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->d_vtx, vtx.data(), vtx.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
  if (!surf.empty())
    HIPCHK(c, hipMemcpyAsync(c->d_surf, surf.data(), surf.size() * sizeof(int), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->d_pts, x, np * sizeof(double), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->d_pts + np, y, np * sizeof(double), hipMemcpyHostToDevice, c->st));
  // the chained triangular solves' control block (every solve resets its
  // tickets; cleared here so that no launch ever reads an unset counter)
  HIPCHK(c, hipMemsetAsync(c->d_chain, 0, 4 * sizeof(int), c->st));
  // GM = max over vertex pairs of the horizontal distance (and to the y
  // images when nsym = 1), the argument of the theta rule
  launch_extent(c->d_pts, c->d_pts + np, np, nsym, c->d_ext, c->st);
  HIPCHK(c, hipGetLastError());
  std::vector<double> ext(nred);
  HIPCHK(c, hipMemcpyAsync(ext.data(), c->d_ext, nred * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  double gm2 = 0.0;
  for (double e : ext) gm2 = std::max(e, gm2);
  // the reference keeps the max of SQRT over the pairs; a correctly rounded
  // square root is monotonic, so one root of the max gives the same value
  c->gm = std::sqrt(gm2);
  c->np = np;
  c->nfa = nfa;
  c->ld32 = (nfa + 3) / 4 * 4;
  ++c->layout_gen;  // kernel arguments of the graphs (geometry, sizes) change
  c->nsym = nsym;
  c->nb = 1 << nsym;
  c->zer = zer;
  c->nsurf = (int)surf.size();
  c->g = g;
  c->rho = rho;
  c->depth = depth;
  c->nq_state = 0;
  c->cap = 0;
  c->dofs_set = false;  // body conditions belong to the previous mesh
  c->last_nrhs = 0;
  c->stats = nemoh_stats{};
  c->upd_launches = 0;
  const double* d = c->d_geo;
  c->G = Geometry{nfa, d, d + nfa, d + 2 * (size_t)nfa, d + 3 * (size_t)nfa, d + 4 * (size_t)nfa,
                  d + 5 * (size_t)nfa, d + 6 * (size_t)nfa, d + 7 * (size_t)nfa,
                  c->d_vtx, c->d_vtx + 4 * (size_t)nfa, c->d_vtx + 8 * (size_t)nfa};
  return NEMOH_OK;
}

int core_nq_sequence(nemoh_core* c, int nfreq, const double* periods, int* nq_before) {
  if (!c || nfreq < 0 || (nfreq > 0 && (!periods || !nq_before))) return NEMOH_ERR_ARG;
  const double pi = 4.0 * std::atan(1.0);
  int state = c->nq_state;
  for (int f = 0; f < nfreq; ++f) {
    nq_before[f] = state;
    const double w = 2.0 * pi / periods[f];
    const std::string at = "frequency " + std::to_string(f) + ": ";
    if (!(std::fabs(w) >= 1.E-4)) return fail(c, NEMOH_ERR_OMEGA, at + "ABS(WR) < 1.E-4");
    const double kwave = w * w / c->g;
    if (c->depth == 0.0 || kwave * c->depth >= 20) continue;
    const double amh = host::x0(w * w * c->depth / c->g);
    if (std::isnan(amh)) return fail(c, NEMOH_ERR_WAVENUMBER, at + "ERREUR DANS LA RECHERCHE DE LA RACINE");
    double qq[101], cq[101];
    state = host::theta_rule(amh / c->depth * c->gm, qq, cq);
  }
  return NEMOH_OK;
}

void core_set_nq_state(nemoh_core* c, int nq) { c->nq_state = nq; }

size_t core_held_bytes(const nemoh_core* c) {
  size_t b = 0;
  for (const auto& kv : c->cap_bytes) b += kv.second;
  return b;
}

size_t core_footprint_bytes(int np, int nfa, int nsym, int precision, int nrhs) {
  const size_t n = nfa, nb = (size_t)1 << nsym, nt = n * nb, k = nrhs;
  const size_t nblk = (n + kNB - 1) / kNB, ninv = nblk * kNB * kNB, ld32 = (n + 3) / 4 * 4;
  size_t b = nb * (2 * n * n * sizeof(double2));  // D, S
  if (precision == NEMOH_MIXED) b += nb * (n * ld32 * sizeof(float2) + ninv * sizeof(float2));
  else b += nb * ninv * sizeof(double2);
  b += nt * k * sizeof(double2) + 3 * n * k * sizeof(double2) + 3 * 2 * n * k * sizeof(double2) +
       2 * n * k * sizeof(float2) + nb * 2 * n * k * sizeof(double2);  // right-hand-side workspace
  b += 20 * n * sizeof(double) + 4 * sizeof(int) + 2 * (size_t)np * sizeof(double);  // geometry
  b += 7 * n * sizeof(double) + nt * 16 * sizeof(double) + nt * kMaxRhs * sizeof(double2);  // body conditions
  return b;
}

int core_release_matrices(nemoh_core* c) {
  if (!c) return NEMOH_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->st));
  for (int b = 0; b < 2; ++b) {
    dfree(c, &c->D[b]);
    dfree(c, &c->S[b]);
    dfree(c, &c->D32[b]);
    dfree(c, &c->D32r[b]);
    dfree(c, &c->S32[b]);
    dfree(c, &c->Pinv[b]);
    dfree(c, &c->Pinv32[b]);
    dfree(c, &c->Z[b]);
    dfree(c, &c->F[b]);
  }
  for (double2** q : {&c->V, &c->B, &c->W, &c->R, &c->sig, &c->pot, &c->pres}) dfree(c, q);
  dfree(c, &c->R32);
  dfree(c, &c->W32);
  c->cap = 0;
  c->w_prev = -1.0;  // the factors are gone
  c->factored_prec = -1;
  c->last_nrhs = 0;  // and the sources
  return NEMOH_OK;
}

int core_device_count() {
  int n = 0;
  return hipGetDeviceCount(&n) == hipSuccess ? n : 0;
}

}  // namespace nemoh

extern "C" {

int nemoh_core_create(nemoh_core** out, int np, int nfa, int nsym, const double* x, const double* y, const double* z,
                      const int* m1, const int* m2, const int* m3, const int* m4, const double* xg, const double* yg,
                      const double* zg, const double* xn, const double* yn, const double* zn, const double* aire,
                      const double* tdis, double zer, double g, double rho, double depth, const double* xr,
                      const double* xz, const double* apd1x, const double* apd1z, const double* apd2x,
                      const double* apd2z) {
  if (!out) return NEMOH_ERR_ARG;
  *out = nullptr;
  if (!xr || !xz || !apd1x || !apd1z || !apd2x || !apd2z) return NEMOH_ERR_ARG;
  const double* geo[8] = {xg, yg, zg, xn, yn, zn, aire, tdis};
  if (!mesh_args_ok(np, nfa, nsym, x, y, z, m1, m2, m3, m4, geo, depth)) return NEMOH_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(device_lock());
  create_error.clear();
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    create_error = "no HIP device";
    return NEMOH_ERR_HIP;
  }
  nemoh_core* c = new nemoh_core();
  auto bail = [&](int code) {
    if (create_error.empty()) {
      const hipError_t e = hipGetLastError();
      create_error = c->err.empty() ? std::string("handle creation: ") + hipGetErrorString(e) : c->err;
    }
    nemoh_core_destroy(c);
    return code;
  };
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking) != hipSuccess)
    return bail(NEMOH_ERR_HIP);
  for (auto& e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) return bail(NEMOH_ERR_HIP);
  std::vector<double> tab(4 * (size_t)kNR * kNZ);
  for (int j = 0; j < kNZ; ++j)
    for (int i = 0; i < kNR; ++i) {
      const size_t f = i + (size_t)kNR * j;  // Fortran (i, j)
      tab[4 * f + 0] = apd1x[f];
      tab[4 * f + 1] = apd1z[f];
      tab[4 * f + 2] = apd2x[f];
      tab[4 * f + 3] = apd2z[f];
    }
  // reciprocal Lagrange denominators of every 3-node stencil (PL2's
  // (U3-U1)*(U3-U2) for the three weight orderings), XR then XZ
  std::vector<double> rden(3 * (size_t)(kNR - 2 + kNZ - 2));
  auto stencils = [&](const double* u, int n, double* out) {
    for (int k = 0; k + 2 < n; ++k) {
      const double a = u[k], b = u[k + 1], cc = u[k + 2];
      out[3 * k + 0] = 1.0 / ((a - b) * (a - cc));
      out[3 * k + 1] = 1.0 / ((b - cc) * (b - a));
      out[3 * k + 2] = 1.0 / ((cc - a) * (cc - b));
    }
  };
  stencils(xr, kNR, rden.data());
  stencils(xz, kNZ, rden.data() + 3 * (kNR - 2));
  std::vector<float> tab32(tab.begin(), tab.end());
  if (hipMalloc((void**)&c->d_tab, tab.size() * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&c->d_tab32, tab32.size() * sizeof(float)) != hipSuccess ||
      hipMalloc((void**)&c->d_rden, rden.size() * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&c->d_xr, kNR * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&c->d_xz, kNZ * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&c->d_par, sizeof(FreqParams)) != hipSuccess ||
      hipMalloc((void**)&c->d_red, 64 * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&c->d_info, 2 * sizeof(int)) != hipSuccess ||
      hipMalloc((void**)&c->d_rs, 2 * sizeof(RefineState)) != hipSuccess)
    return bail(NEMOH_ERR_HIP);
  {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return bail(NEMOH_ERR_HIP);
    c->cu_count = prop.multiProcessorCount;
    if (noalloc_scratch(dev, &c->d_scratch) != hipSuccess) return bail(NEMOH_ERR_HIP);
  }
  // copies on the handle's own stream (no legacy-stream synchronisation)
  if (hipMemcpyAsync(c->d_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, c->st) != hipSuccess ||
      hipMemcpyAsync(c->d_tab32, tab32.data(), tab32.size() * sizeof(float), hipMemcpyHostToDevice, c->st) !=
          hipSuccess ||
      hipMemcpyAsync(c->d_rden, rden.data(), rden.size() * sizeof(double), hipMemcpyHostToDevice, c->st) !=
          hipSuccess ||
      hipMemcpyAsync(c->d_xr, xr, kNR * sizeof(double), hipMemcpyHostToDevice, c->st) != hipSuccess ||
      hipMemcpyAsync(c->d_xz, xz, kNZ * sizeof(double), hipMemcpyHostToDevice, c->st) != hipSuccess ||
      hipMemsetAsync(c->d_info, 0, 2 * sizeof(int), c->st) != hipSuccess || hipStreamSynchronize(c->st) != hipSuccess)
    return bail(NEMOH_ERR_HIP);
  c->T = Tables{c->d_xr, c->d_xz, c->d_tab, c->d_rden, c->d_rden + 3 * (kNR - 2)};
  // NEMOH_GRAPHS=0 in the environment: launch one by one (as
  // NEMOH_OPT_GRAPHS = 0), e.g. under a profiler that traces kernels
  if (const char* e = std::getenv("NEMOH_GRAPHS")) c->use_graphs = std::atoi(e) != 0;
  // NEMOH_LU_RANK=64 or 128: the factorisation form of every handle (as
  // NEMOH_OPT_LU_RANK), for handles a caller does not create itself (the
  // request loop's workers)
  if (const char* e = std::getenv("NEMOH_LU_RANK")) {
    const int r = std::atoi(e);
    if (r == 64 || r == 128) c->lu_rank = r;
  }
  // NEMOH_LU_TUNED=1: the handle takes its factorisation form from the
  // autotuner's table (as NEMOH_OPT_LU_TUNED), likewise
  if (const char* e = std::getenv("NEMOH_LU_TUNED")) c->lu_tuned = std::atoi(e) != 0;
  if (int r = core_bind_mesh(c, np, nfa, nsym, x, y, z, m1, m2, m3, m4, geo, zer, g, rho, depth)) return bail(r);
  *out = c;
  return NEMOH_OK;
}

void nemoh_core_destroy(nemoh_core* c) {
  if (!c) return;
  std::lock_guard<std::recursive_mutex> lk(device_lock());
  if (c->st) (void)hipStreamSynchronize(c->st);
  auto release = [](auto*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
  };
  release(c->d_geo);
  release(c->d_vtx);
  release(c->d_surf);
  release(c->d_xr);
  release(c->d_xz);
  release(c->d_tab);
  release(c->d_tab32);
  release(c->d_rden);
  release(c->d_par);
  release(c->d_red);
  release(c->d_info);
  release(c->d_rs);
  release(c->d_chain);
  release(c->d_nds);
  release(c->d_force);
  release(c->d_pts);
  release(c->d_ext);
  release(c->d_pg);
  release(c->d_theta);
  release(c->d_kochin);
  release(c->d_field);
  release(c->d_phi);
  release(c->d_fbad);
  release(c->d_body);
  release(c->d_dofs);
  release(c->d_fnds);
  release(c->d_beta);
  release(c->d_pinc);
  for (int b = 0; b < 2; ++b) {
    release(c->D[b]);
    release(c->S[b]);
    release(c->D32[b]);
    release(c->D32r[b]);
    release(c->S32[b]);
    release(c->Pinv[b]);
    release(c->Pinv32[b]);
    release(c->Z[b]);
    release(c->F[b]);
  }
  release(c->V);
  release(c->B);
  release(c->W);
  release(c->R);
  release(c->sig);
  release(c->pot);
  release(c->pres);
  release(c->R32);
  release(c->W32);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->upd_ev) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->lu_dep) (void)hipEventDestroy(e);
  drop_graphs(c);
  if (c->st2) (void)hipStreamSynchronize(c->st2);
  if (c->st2) (void)hipStreamDestroy(c->st2);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
}

const char* nemoh_core_last_error(const nemoh_core* c) {
  if (c) return c->err.c_str();
  return create_error.empty() ? "null handle" : create_error.c_str();
}

int nemoh_core_set_precision(nemoh_core* c, int mode, double tol, int max_iter) {
  if (!c || (mode != NEMOH_FP64 && mode != NEMOH_MIXED)) return NEMOH_ERR_ARG;
  c->prec = mode;
  if (tol > 0) c->tol = tol;
  if (max_iter > 0) c->max_iter = max_iter;
  return NEMOH_OK;
}

int nemoh_core_set_option(nemoh_core* c, int option, int value) {
  if (!c) return NEMOH_ERR_ARG;
  if (option == NEMOH_OPT_STEPWISE_SOLVE) {
    c->stepwise_solve = value != 0;
    return NEMOH_OK;
  }
  if (option == NEMOH_OPT_FP32_ASSEMBLY) {
    c->fp32_assembly = value != 0;
    return NEMOH_OK;
  }
  if (option == NEMOH_OPT_LU_RANK) {
    if (value != 64 && value != 128) return fail(c, NEMOH_ERR_ARG, "LU update rank must be 64 or 128");
    if (value != c->lu_rank) {
      c->lu_rank = value;
      c->w_prev = -1.0;   // the factors on the device came from the other form
      ++c->layout_gen;    // and so did the graphs' launch sequences
    }
    return NEMOH_OK;
  }
  if (option == NEMOH_OPT_LU_FUSED_PANEL) {
    if (c->lu_fused != (value != 0)) {
      c->lu_fused = value != 0;
      c->w_prev = -1.0;
      ++c->layout_gen;  // the graphs' launch sequences change
    }
    return NEMOH_OK;
  }
  if (option == NEMOH_OPT_LU_LOOKAHEAD) {
    if (c->lookahead != (value != 0)) {
      c->lookahead = value != 0;
      ++c->layout_gen;  // the graphs' launch sequences change
    }
    return NEMOH_OK;
  }
  if (option == NEMOH_OPT_LU_BAND) {
    if (value != 16 && value != 32 && value != 64) return fail(c, NEMOH_ERR_ARG, "LU tile band must be 16, 32 or 64");
    c->lu_band = value;  // the factors do not depend on it
    return NEMOH_OK;
  }
  if (option == NEMOH_OPT_LU_TUNED) {
    if (c->lu_tuned != (value != 0)) {
      c->lu_tuned = value != 0;
      c->w_prev = -1.0;  // the factors on the device may be of the other form
    }
    return NEMOH_OK;
  }
  if (option == NEMOH_OPT_PAIRED_ASSEMBLY) {
    if (c->paired_assembly != (value != 0)) {
      c->paired_assembly = value != 0;
      c->w_prev = -1.0;  // the matrices of the other form differ in rounding
      ++c->layout_gen;   // and so do the graphs' launch sequences
    }
    return NEMOH_OK;
  }
  if (option == NEMOH_OPT_GRAPHS) {
    c->use_graphs = value != 0;
    return NEMOH_OK;
  }
  if (option == NEMOH_OPT_SOLVE_SPIN_LIMIT) {
    if (value < 0) return fail(c, NEMOH_ERR_ARG, "spin limit must be >= 0");
    c->spin_limit = value;
    return NEMOH_OK;
  }
  return fail(c, NEMOH_ERR_ARG, "unknown option");
}

int nemoh_core_autotune(nemoh_core* c, int reps, nemoh_tune_entry* entries, int* chosen) {
  if (!c || reps < 1 || reps > 100) return NEMOH_ERR_ARG;
  if (int r = ensure_matrices(c)) return r;
  const bool mixed = c->prec == NEMOH_MIXED;
  const int n = c->nfa;
  static const TunedForm cand[NEMOH_TUNE_CANDIDATES] = {{64, 16}, {64, 32}, {64, 64}, {128, 16}, {128, 32}, {128, 64}};
  double best = 0;
  int ib = -1;
  for (int k = 0; k < NEMOH_TUNE_CANDIDATES; ++k) {
    float mn = 0;
    for (int r = 0; r < reps; ++r) {
      if (mixed) {
        auto* A = reinterpret_cast<C2<float>*>(c->D32[0]);
        launch_fill_test_matrix(A, c->ld32, n, c->st);
        HIPCHK(c, hipEventRecord(c->ev[0], c->st));
        launch_lu<float>(A, c->ld32, n, reinterpret_cast<C2<float>*>(c->Pinv32[0]), c->d_info, c->st, cand[k].rank,
                         nullptr, nullptr, nullptr, nullptr, nullptr, 1 << 20, cand[k].band);
      } else {
        auto* A = reinterpret_cast<C2<double>*>(c->D[0]);
        launch_fill_test_matrix(A, n, n, c->st);
        HIPCHK(c, hipEventRecord(c->ev[0], c->st));
        launch_lu<double>(A, n, n, reinterpret_cast<C2<double>*>(c->Pinv[0]), c->d_info, c->st, cand[k].rank,
                          nullptr, nullptr, nullptr, nullptr, nullptr, 1 << 20, cand[k].band);
      }
      HIPCHK(c, hipGetLastError());
      HIPCHK(c, hipEventRecord(c->ev[1], c->st));
      HIPCHK(c, hipEventSynchronize(c->ev[1]));
      float ms = 0;
      HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
      if (r == 0 || ms < mn) mn = ms;
    }
    if (entries) entries[k] = nemoh_tune_entry{cand[k].rank, cand[k].band, (double)mn};
    if (ib < 0 || mn < best) {
      best = mn;
      ib = k;
    }
  }
  launch_info_reset(c->d_info, c->st);
  HIPCHK(c, hipStreamSynchronize(c->st));
  {
    std::lock_guard<std::mutex> lk(tune_mutex());
    tune_table()[{c->prec, (n + kNB - 1) / kNB}] = cand[ib];
  }
  c->lu_tuned = true;
  c->w_prev = -1.0;  // the matrices now hold test data
  if (chosen) *chosen = ib;
  return NEMOH_OK;
}

int nemoh_core_enable_timing(nemoh_core* c, int on) {
  if (!c) return NEMOH_ERR_ARG;
  c->timing = on != 0;
  return NEMOH_OK;
}

int nemoh_core_get_stats(const nemoh_core* c, nemoh_stats* out) {
  if (!c || !out) return NEMOH_ERR_ARG;
  *out = c->stats;
  out->lu_rank = c->lu_rank_used;
  out->lu_band = c->lu_band_used;
  return NEMOH_OK;
}

int nemoh_core_set_dofs(nemoh_core* c, const int* body_of_panel, int nrad, const nemoh_dof* rad, int nforce,
                        const nemoh_dof* force, double xeff, double yeff) {
  static_assert(sizeof(Dof) == sizeof(nemoh_dof) && offsetof(Dof, dir) == offsetof(nemoh_dof, dir) &&
                    offsetof(Dof, axis) == offsetof(nemoh_dof, axis),
                "Dof mirrors nemoh_dof");
  if (!c) return NEMOH_ERR_ARG;
  c->dofs_set = false;
  if (nrad < 0 || nrad > kMaxRhs || nforce < 0 || (nrad > 0 && !rad) || (nforce > 0 && !force))
    return fail(c, NEMOH_ERR_ARG, "bad numbers of degrees of freedom");
  std::vector<nemoh_dof> dofs;
  if (nrad > 0) dofs.assign(rad, rad + nrad);
  if (nforce > 0) dofs.insert(dofs.end(), force, force + nforce);
  for (const nemoh_dof& d : dofs)
    if ((d.kind != 1 && d.kind != 2) || d.body < 1)
      return fail(c, NEMOH_ERR_ARG, "a degree of freedom needs kind 1 (translation) or 2 (rotation) and body >= 1");
  const int n = c->nfa, nt = n * c->nb;
  std::vector<int> body(n, 1);
  if (body_of_panel) body.assign(body_of_panel, body_of_panel + n);
  int r = 0;
  if ((r = dalloc(c, &c->d_pg, 7 * (size_t)n)) || (r = dalloc(c, &c->d_body, (size_t)n)) ||
      (r = dalloc(c, &c->d_dofs, std::max<size_t>(1, dofs.size()))) ||
      (r = dalloc(c, &c->d_fnds, std::max<size_t>(1, (size_t)nforce * nt))) || (r = dalloc(c, &c->d_beta, kMaxRhs)) ||
      (r = dalloc(c, &c->d_pinc, (size_t)nt * kMaxRhs)))
    return r;
  HIPCHK(c, hipMemcpyAsync(c->d_body, body.data(), n * sizeof(int), hipMemcpyHostToDevice, c->st));
  if (!dofs.empty())
    HIPCHK(c, hipMemcpyAsync(c->d_dofs, dofs.data(), dofs.size() * sizeof(nemoh_dof), hipMemcpyHostToDevice, c->st));
  launch_panel_props(c->G, c->d_pg, c->st);
  launch_nds(c->d_pg, c->d_body, reinterpret_cast<const Dof*>(c->d_dofs) + nrad, nforce, n, nt, c->d_fnds, c->st);
  HIPCHK(c, hipGetLastError());
  std::vector<double> area(n);
  HIPCHK(c, hipMemcpyAsync(area.data(), c->d_pg + 6 * (size_t)n, n * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  for (int i = 0; i < n; ++i)
    if (!(area[i] >= 1.0E-07)) {  // preProcessor/Mesh.f90:222-225
      char buf[96];
      std::snprintf(buf, sizeof buf, "Error: surface of panel %d is too small (%.7e)", i + 1, area[i]);
      return fail(c, NEMOH_ERR_PANEL, buf);
    }
  c->nrad = nrad;
  c->nforce = nforce;
  c->xeff = xeff;
  c->yeff = yeff;
  c->dofs_set = true;
  return NEMOH_OK;
}

int nemoh_core_body_conditions(nemoh_core* c, double omega, int nbeta, const double* beta, double* nvel,
                               double* pinc, double* nds) {
  if (!c) return NEMOH_ERR_ARG;
  if (int r = generate_waves(c, omega, nbeta, beta)) return r;
  const int nrhs = nbeta + c->nrad;
  const size_t nt = (size_t)c->nfa * c->nb;
  auto colmajor = [&](const double2* d, size_t rows, int cols, double* out) -> int {
    std::vector<double> h(2 * rows * cols);
    HIPCHK(c, hipMemcpyAsync(h.data(), d, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (size_t i = 0; i < rows; ++i)
      for (int q = 0; q < cols; ++q) {
        out[2 * (i + rows * q)] = h[2 * (i * cols + q)];
        out[2 * (i + rows * q) + 1] = h[2 * (i * cols + q) + 1];
      }
    return NEMOH_OK;
  };
  if (nvel)
    if (int r = colmajor(c->V, nt, nrhs, nvel)) return r;
  if (pinc && nbeta > 0)
    if (int r = colmajor(c->d_pinc, nt, nbeta, pinc)) return r;
  if (nds && c->nforce > 0) {
    std::vector<double> h((size_t)c->nforce * nt);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->d_fnds, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (int k = 0; k < c->nforce; ++k)
      for (size_t i = 0; i < nt; ++i) nds[k + (size_t)c->nforce * i] = h[k * nt + i];
  }
  return NEMOH_OK;
}

int nemoh_core_solve_waves(nemoh_core* c, double omega, int nbeta, const double* beta, double* forces, double* fk) {
  if (!c) return NEMOH_ERR_ARG;
  if (!(std::fabs(omega) >= 1.E-4)) return fail(c, NEMOH_ERR_OMEGA, "ABS(WR) < 1.E-4");
  const double pi = 4.0 * std::atan(1.0);
  const double period = 2.0 * pi / omega;  // T as the Solver's main program forms it
  nemoh_freq f;
  if (int r = nemoh_core_freq_setup(c, period, &f)) return r;
  bool fd;
  double w;
  if (int r = build_params(c, period, f.amh, f.nexp, f.ambda, f.ar, f.nq, f.qq, f.cq, &fd, &w)) return r;
  if (int r = generate_waves(c, omega, nbeta, beta)) return r;
  const int nrhs = nbeta + c->nrad;
  if (int r = run_pipeline(c, w, nrhs)) return r;
  if (forces && c->nforce > 0)
    if (int r = download_forces(c, c->pres, c->d_fnds, c->nforce, nrhs, forces)) return r;
  if (fk && c->nforce > 0 && nbeta > 0)
    if (int r = download_forces(c, c->d_pinc, c->d_fnds, c->nforce, nbeta, fk)) return r;
  return finish_call(c);
}

int nemoh_core_kochin(nemoh_core* c, double xeff, double yeff, int ntheta, const double* theta, double* hkochin) {
  if (!c) return NEMOH_ERR_ARG;
  if (ntheta <= 0 || !theta || !hkochin) return fail(c, NEMOH_ERR_ARG, "bad Kochin arguments");
  if (c->last_nrhs <= 0) return fail(c, NEMOH_ERR_ARG, "no solve on this mesh yet");
  const int nrhs = c->last_nrhs;
  int r = 0;
  if ((r = dalloc(c, &c->d_theta, (size_t)ntheta)) || (r = dalloc(c, &c->d_kochin, (size_t)ntheta * nrhs))) return r;
  HIPCHK(c, hipMemcpyAsync(c->d_theta, theta, ntheta * sizeof(double), hipMemcpyHostToDevice, c->st));
  // the Solver's kwave of that frequency (SOLVE_BEM.f90:66-84): w^2/g, or
  // AMH/Depth in finite depth -- the ak0 of the frequency parameters
  launch_kochin(c->G, c->sig, c->nsym, nrhs, c->h_par.ak0, c->depth, xeff, yeff, c->d_theta, ntheta, c->d_kochin,
                c->st);
  HIPCHK(c, hipGetLastError());
  std::vector<double> h(2 * (size_t)ntheta * nrhs);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->d_kochin, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  for (int t = 0; t < ntheta; ++t)  // device (ntheta x nrhs) -> caller column-major
    for (int q = 0; q < nrhs; ++q) {
      hkochin[2 * (t + (size_t)ntheta * q)] = h[2 * ((size_t)t * nrhs + q)];
      hkochin[2 * (t + (size_t)ntheta * q) + 1] = h[2 * ((size_t)t * nrhs + q) + 1];
    }
  return NEMOH_OK;
}

int nemoh_core_kochin_noalloc(nemoh_core* c, double xeff, double yeff, int ntheta, const double* theta,
                              double* hkochin) {
  if (!c) return NEMOH_ERR_ARG;
  if (ntheta <= 0 || !theta || !hkochin) return fail(c, NEMOH_ERR_ARG, "bad Kochin arguments");
  if (ntheta > kNoallocMaxAngles)
    return fail(c, NEMOH_ERR_ARG,
                "nemoh_core_kochin_noalloc takes at most " + std::to_string(kNoallocMaxAngles) + " angles");
  if (c->last_nrhs <= 0) return fail(c, NEMOH_ERR_ARG, "no solve on this mesh yet");
  if (!c->d_scratch) return fail(c, NEMOH_ERR_HIP, "no device scratch");
  const int nrhs = c->last_nrhs;
  double* d_theta = reinterpret_cast<double*>(c->d_scratch);
  double2* d_h = reinterpret_cast<double2*>(d_theta + kNoallocMaxAngles);
  double2* d_ht = d_h + (size_t)kNoallocMaxAngles * kMaxRhs;
  HIPCHK(c, hipMemcpyAsync(d_theta, theta, ntheta * sizeof(double), hipMemcpyHostToDevice, c->st));
  launch_kochin(c->G, c->sig, c->nsym, nrhs, c->h_par.ak0, c->depth, xeff, yeff, d_theta, ntheta, d_h, c->st);
  // (angle x problem) -> the caller's column-major (angle, problem)
  launch_transpose_c128(d_h, ntheta, nrhs, d_ht, c->st);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(hkochin, d_ht, 2 * (size_t)ntheta * nrhs * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return NEMOH_OK;
}

int nemoh_core_potential(nemoh_core* c, int npoints, const double* xyz, double* phi) {
  if (!c) return NEMOH_ERR_ARG;
  if (npoints <= 0 || !xyz || !phi) return fail(c, NEMOH_ERR_ARG, "bad potential arguments");
  if (c->last_nrhs <= 0) return fail(c, NEMOH_ERR_ARG, "no solve on this mesh yet");
  const int nrhs = c->last_nrhs;
  int r = 0;
  if ((r = dalloc(c, &c->d_field, 3 * (size_t)npoints)) || (r = dalloc(c, &c->d_phi, (size_t)npoints * nrhs)) ||
      (r = dalloc(c, &c->d_fbad, (size_t)npoints)))
    return r;
  HIPCHK(c, hipMemcpyAsync(c->d_field, xyz, 3 * (size_t)npoints * sizeof(double), hipMemcpyHostToDevice, c->st));
  // VNVF's A (COMPUTE_GREEN_FREESURFACE.f90:420-442) from the AMH of the solve
  double a = 0.0;
  if (c->h_par.fd) {
    const double akh = c->amh * std::tanh(c->amh);
    a = (c->amh + akh) * (c->amh + akh) / (c->depth * (c->amh * c->amh - akh * akh + akh));
  }
  launch_potential(c->G, c->T, c->d_par, c->h_par.fd, 2.0 * c->depth, a, c->d_field, npoints, c->sig, nrhs, c->d_phi,
                   c->d_fbad, c->st);
  HIPCHK(c, hipGetLastError());
  std::vector<double> h(2 * (size_t)npoints * nrhs);
  std::vector<int> bad(npoints);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->d_phi, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(bad.data(), c->d_fbad, bad.size() * sizeof(int), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  for (int p = 0; p < npoints; ++p)
    if (bad[p]) {
      char buf[128];
      std::snprintf(buf, sizeof buf, "AKZ < -1.5 E-6: field point %d (z = %g) too close to the free surface", p,
                    xyz[3 * (size_t)p + 2]);
      return fail(c, NEMOH_ERR_ARG, buf);
    }
  for (int p = 0; p < npoints; ++p)  // device (npoints x nrhs) -> caller column-major
    for (int q = 0; q < nrhs; ++q) {
      phi[2 * (p + (size_t)npoints * q)] = h[2 * ((size_t)p * nrhs + q)];
      phi[2 * (p + (size_t)npoints * q) + 1] = h[2 * ((size_t)p * nrhs + q) + 1];
    }
  return NEMOH_OK;
}

double nemoh_core_incident_wavenumber(double w, double depth, double g) {
  return host::incident_wavenumber(w, depth, g);
}

int nemoh_core_get_extent(const nemoh_core* c, double* gm) {
  if (!c || !gm) return NEMOH_ERR_ARG;
  *gm = c->gm;
  return NEMOH_OK;
}

int nemoh_core_solve_multi(nemoh_core* c, double period, double amh, int nexp, const double* ambda, const double* ar,
                           int nq, const double* qq, const double* cq, int nrhs, const double* nvel, double* pressure,
                           int nint, const double* nds, double* forces) {
  if (!c || nrhs <= 0 || nrhs > kMaxRhs || !nvel) return fail(c, NEMOH_ERR_ARG, "bad arguments");
  bool fd;
  double w;
  if (int r = build_params(c, period, amh, nexp, ambda, ar, nq, qq, cq, &fd, &w)) return r;
  if (int r = ensure_rhs(c, nrhs)) return r;
  if (int r = upload_nvel_colmajor(c, nvel, nrhs)) return r;
  if (int r = run_pipeline(c, w, nrhs)) return r;
  return finish_outputs(c, nrhs, pressure, nint, nds, forces);
}

int nemoh_core_solve(nemoh_core* c, double period, double amh, int nexp, const double* ambda, const double* ar, int nq,
                     const double* qq, const double* cq, const double* nvel, double* zigb, double* zigs, double* zpb,
                     double* zps) {
  if (!c || !nvel) return fail(c, NEMOH_ERR_ARG, "bad arguments");
  bool fd;
  double w;
  if (int r = build_params(c, period, amh, nexp, ambda, ar, nq, qq, cq, &fd, &w)) return r;
  if (int r = ensure_rhs(c, 1)) return r;
  if (int r = upload_nvel_colmajor(c, nvel, 1)) return r;
  if (int r = run_pipeline(c, w, 1)) return r;
  const size_t n = c->nfa;
  std::vector<double> s(4 * n), p(4 * n);
  HIPCHK(c, hipMemcpyAsync(s.data(), c->sig, s.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(p.data(), c->pot, p.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (int r = check_info(c)) return r;
  if (zigb) std::memcpy(zigb, s.data(), 2 * n * sizeof(double));
  if (zigs) std::memcpy(zigs, s.data() + 2 * n, 2 * n * sizeof(double));
  if (zpb) std::memcpy(zpb, p.data(), 2 * n * sizeof(double));
  if (zps) std::memcpy(zps, p.data() + 2 * n, 2 * n * sizeof(double));
  return NEMOH_OK;
}

int nemoh_core_freq_setup(nemoh_core* c, double period, nemoh_freq* f) {
  if (!c || !f) return NEMOH_ERR_ARG;
  std::memset(f, 0, sizeof *f);
  const double pi = 4.0 * std::atan(1.0);
  const double w = 2.0 * pi / period;
  if (!(std::fabs(w) >= 1.E-4)) return fail(c, NEMOH_ERR_OMEGA, "ABS(WR) < 1.E-4");
  const double kwave = w * w / c->g;
  if (c->depth == 0.0 || kwave * c->depth >= 20) {
    // CINT_INFD fills QQ/CQ; the module NQ stays as it was
    f->finite_depth = 0;
    const double ak0 = w * w / c->g;
    double qq[101], cq[101];
    const int n = host::theta_rule(ak0 * c->gm, qq, cq);
    f->nq = c->nq_state;
    for (int k = 0; k < 101; ++k) {
      f->qq[k] = (k < n) ? qq[k] : 0.0;
      f->cq[k] = (k < n) ? cq[k] : 0.0;
    }
  } else {
    f->finite_depth = 1;
    const double akh0 = w * w * c->depth / c->g;
    f->amh = host::x0(akh0);
    if (std::isnan(f->amh)) return fail(c, NEMOH_ERR_WAVENUMBER, "ERREUR DANS LA RECHERCHE DE LA RACINE");
    const double akh = f->amh * std::tanh(f->amh);
    double qq[101], cq[101];
    c->nq_state = host::theta_rule(f->amh / c->depth * c->gm, qq, cq);  // NQ=N; rule stays local
    f->nq = c->nq_state;
    const int e = host::prony_fit(akh, f->amh, &f->nexp, f->ambda, f->ar);
    if (e) return fail(c, NEMOH_ERR_FIT, "exponential fit of the finite-depth kernel failed");
  }
  return NEMOH_OK;
}

int nemoh_core_solve_frequency(nemoh_core* c, double period, int nrhs, const double* nvel, double* pressure, int nint,
                               const double* nds, double* forces) {
  nemoh_freq f;
  if (int r = nemoh_core_freq_setup(c, period, &f)) return r;
  return nemoh_core_solve_multi(c, period, f.amh, f.nexp, f.ambda, f.ar, f.nq, f.qq, f.cq, nrhs, nvel, pressure, nint,
                                nds, forces);
}

int nemoh_core_pairs(nemoh_core* c, double period, double amh, int nexp, const double* ambda, const double* ar, int nq,
                     const double* qq, const double* cq, double* out) {
  if (!c || !out) return NEMOH_ERR_ARG;
  bool fd;
  double w;
  if (int r = build_params(c, period, amh, nexp, ambda, ar, nq, qq, cq, &fd, &w)) return r;
  const size_t cnt = (size_t)c->nfa * c->nfa * 16;
  double* d = nullptr;
  std::lock_guard<std::recursive_mutex> lk(device_lock());
  HIPCHK(c, hipMalloc((void**)&d, cnt * sizeof(double)));
  launch_pairs(c->G, c->T, c->d_par, c->h_par.fd, d, c->st);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, cnt * sizeof(double), hipMemcpyDeviceToHost, c->st);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(c, NEMOH_ERR_HIP, hipGetErrorString(e));
  return NEMOH_OK;
}

int nemoh_core_prepare_mesh(int np, int nfa, int nsym, const double* x, const double* y, const double* z, const int* m1,
                            const int* m2, const int* m3, const int* m4, double* xg, double* yg, double* zg, double* xn,
                            double* yn, double* zn, double* aire, double* tdis, double* zer) {
  if (np <= 0 || nfa <= 0 || !x || !y || !z || !m1 || !m2 || !m3 || !m4 || !zer) return NEMOH_ERR_ARG;
  std::vector<int> a(nfa), b(nfa), cc(nfa), d(nfa);
  for (int i = 0; i < nfa; ++i) {
    a[i] = m1[i] - 1;
    b[i] = m2[i] - 1;
    cc[i] = m3[i] - 1;
    d[i] = m4[i] - 1;
    if (a[i] < 0 || b[i] < 0 || cc[i] < 0 || d[i] < 0 || a[i] >= np || b[i] >= np || cc[i] >= np || d[i] >= np)
      return NEMOH_ERR_ARG;
  }
  host::PanelGeometry g{xg, yg, zg, xn, yn, zn, aire, tdis, 0.0};
  const int r = host::prepare_mesh(np, nfa, nsym, x, y, z, a.data(), b.data(), cc.data(), d.data(), &g);
  *zer = g.zer;
  return r ? NEMOH_ERR_PANEL : NEMOH_OK;
}

int nemoh_core_build_tables(double* xr, double* xz, double* apd1x, double* apd1z, double* apd2x, double* apd2z) {
  if (!xr || !xz || !apd1x || !apd1z || !apd2x || !apd2z) return NEMOH_ERR_ARG;
  // grids (NEMOH.f90:184-198)
  for (int j = 1; j <= kNZ; ++j)
    xz[j - 1] = -std::fmin(std::fmin(std::pow(10.0, j / 5.0 - 6.0), std::pow(10.0, j / 8.0 - 4.5)), 16.0);
  xr[0] = 0.0;
  for (int i = 2; i <= kNR; ++i)
    xr[i - 1] = (i < 40) ? std::fmin(std::pow(10.0, (i - 1.0) / 5 - 6), 4.0 / 3.0 + std::fabs((i - 32.0) / 3.0))
                         : 4.0 / 3.0 + std::fabs((i - 32.0) / 3.0);
  std::lock_guard<std::recursive_mutex> lk(device_lock());
  double *dxr = nullptr, *dxz = nullptr, *dtab = nullptr;
  std::vector<double> tab(4 * (size_t)kNR * kNZ);
  hipError_t e = hipMalloc((void**)&dxr, kNR * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&dxz, kNZ * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&dtab, tab.size() * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(dxr, xr, kNR * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dxz, xz, kNZ * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    launch_crek(dxr, dxz, dtab, nullptr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(tab.data(), dtab, tab.size() * sizeof(double), hipMemcpyDeviceToHost);
  if (dxr) (void)hipFree(dxr);
  if (dxz) (void)hipFree(dxz);
  if (dtab) (void)hipFree(dtab);
  if (e != hipSuccess) return NEMOH_ERR_HIP;
  for (size_t f = 0; f < (size_t)kNR * kNZ; ++f) {  // node f = i + 328 j (Fortran order)
    apd1x[f] = tab[4 * f + 0];
    apd1z[f] = tab[4 * f + 1];
    apd2x[f] = tab[4 * f + 2];
    apd2z[f] = tab[4 * f + 3];
  }
  return NEMOH_OK;
}

double nemoh_core_wavenumber(double ak) { return host::x0(ak); }

int nemoh_core_theta_rule(double akk, double* qq, double* cq) {
  if (!qq || !cq) return -1;
  return host::theta_rule(akk, qq, cq);
}

int nemoh_core_prony(double akh, double amh, int* nexp, double* ambda, double* ar) {
  if (!nexp || !ambda || !ar) return NEMOH_ERR_ARG;
  return host::prony_fit(akh, amh, nexp, ambda, ar) ? NEMOH_ERR_FIT : NEMOH_OK;
}

int nemoh_core_mfma_probe(int fp32, double* out) {
  if (!out) return NEMOH_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(device_lock());
  double ha[64], hb[64];
  for (int i = 0; i < 16; ++i)
    for (int k = 0; k < 4; ++k) ha[i * 4 + k] = (i + 1) * 0.5 + k * 3.0;
  for (int k = 0; k < 4; ++k)
    for (int j = 0; j < 16; ++j) hb[k * 16 + j] = (k + 1) * 1.25 - j * 0.75;
  hipError_t e;
  if (fp32) {
    float fa[64], fb[64], fd[256];
    for (int k = 0; k < 64; ++k) {
      fa[k] = (float)ha[k];
      fb[k] = (float)hb[k];
    }
    float *da = nullptr, *db = nullptr, *dd = nullptr;
    e = hipMalloc((void**)&da, sizeof fa);
    if (e == hipSuccess) e = hipMalloc((void**)&db, sizeof fb);
    if (e == hipSuccess) e = hipMalloc((void**)&dd, sizeof fd);
    if (e == hipSuccess) e = hipMemcpy(da, fa, sizeof fa, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, fb, sizeof fb, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      launch_mfma_probe_f32(da, db, dd, nullptr);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(fd, dd, sizeof fd, hipMemcpyDeviceToHost);
    if (da) (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dd) (void)hipFree(dd);
    for (int k = 0; k < 256; ++k) out[k] = fd[k];
  } else {
    double *da = nullptr, *db = nullptr, *dd = nullptr;
    e = hipMalloc((void**)&da, sizeof ha);
    if (e == hipSuccess) e = hipMalloc((void**)&db, sizeof hb);
    if (e == hipSuccess) e = hipMalloc((void**)&dd, 256 * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(da, ha, sizeof ha, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, hb, sizeof hb, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      launch_mfma_probe(da, db, dd, nullptr);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, dd, 256 * sizeof(double), hipMemcpyDeviceToHost);
    if (da) (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dd) (void)hipFree(dd);
  }
  return e == hipSuccess ? NEMOH_OK : NEMOH_ERR_HIP;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches: many independent bodies (one handle each) solved together, each
// step one launch over all of them -- the paired assembly, every step of the
// block LU (max_i ceil(n_i / 128) block pairs of launches instead of the
// sum), the chained triangular solves, and the right-hand sides, refinement
// steps, phi, combination and forces -- on one stream; only the frequency
// set-up and the body conditions stay per body.  No counterpart in the
// reference, which solves one body per run; every body gets the arithmetic
// of its own handle's path, so its results equal a solve on its handle
// alone, bit for bit.
struct nemoh_batch {
  std::vector<nemoh_core*> bodies;
  int prec = NEMOH_FP64, nsym = 0, nb = 1, max_rhs = 0;
  hipStream_t st = nullptr;
  // job tables, built for the bodies' current buffers (layout_gen of each
  // body when built) and for the right-hand-side count they were built for
  std::vector<unsigned> gens;
  std::vector<int> shapes;  // per body when built: nfa, nb, nforce, nrad
  int built_rhs = -1;
  AsmJob* d_asm = nullptr;
  int asm_count = 0, asm_grid = 0;
  LuGroup<float> lu32;
  LuGroup<double> lu64;
  TrsvGroup<float> tr32[2];
  TrsvGroup<double> tr64[2];
  // the per-body steps: one job table per symmetry block, the workgroups of
  // each kind of grouped launch, the forces of every body in one buffer
  // (body i at force_off[i]: nforce x nrhs, then nforce x nbeta) and the
  // status words gathered after the solves
  BodyJob* d_body[2] = {nullptr, nullptr};
  int body_grid[4] = {0, 0, 0, 0};
  double2* d_force = nullptr;
  std::vector<size_t> force_off;
  BatchStatus* d_status = nullptr;
  // per call: each body's frequency (the combination's rho w) and, when the
  // bodies' frequencies fall in both depth regimes, the assembly jobs of
  // each regime (host copies of the jobs kept for that)
  double* d_omega = nullptr;
  std::vector<double> omega_host;  // the source of d_omega's copy (kept until the next call)
  std::vector<AsmJob> asm_host, asm_part;
  AsmJob* d_asm_part = nullptr;
  // timing of the grouped rank-128 updates (nemoh_batch_set_timing): two
  // events per update launch of a call, the durations summed over calls
  bool timing = false;
  std::vector<hipEvent_t> upd_ev;
  double upd_ms = 0.0, upd_flops = 0.0, upd_bytes = 0.0;
  int upd_launches = 0;
  std::string err;
};

namespace {

// The rank-128 update launches of a grouped factorisation: their count, and
// their algorithmic flops and bytes added to flops and bytes.
template <typename T>
size_t update_work(const LuGroup<T>& g, double& flops, double& bytes) {
  size_t n = 0;
  for (const auto& L : g.launches)
    if (L.kind == 3) {
      ++n;
      flops += L.flops;
      bytes += L.bytes;
    }
  return n;
}

// Workgroups of a body's paired assembly from which it takes a launch of its
// own in a batch (about five rounds of the GPU at three workgroups per CU).
constexpr int kOwnAssemblyWg = 4096;
// Workgroups of a body's paired assembly: 8 x 8 panel tiles of the lower
// triangle with the diagonal, four per workgroup.
int pair_assembly_wg(const nemoh_core* c) {
  const int nt = (c->nfa + 7) / 8, nw = nt * (nt + 1) / 2;
  return (nw + 3) / 4;
}

// Waits for the batch's stream whatever the launch status (copies into host
// memory may be pending), then reports whether both were clean.
bool batch_sync(nemoh_batch* b) {
  const hipError_t launched = hipGetLastError();
  const hipError_t synced = hipStreamSynchronize(b->st);
  return launched == hipSuccess && synced == hipSuccess;
}

int batch_fail(nemoh_batch* b, int code, const std::string& msg) {
  // work already queued may read the caller's headings: finished first
  (void)hipStreamSynchronize(b->st);
  b->err = msg;
  return code;
}

void batch_free_tables(nemoh_batch* b) {
  if (b->d_asm) (void)hipFree(b->d_asm);
  b->d_asm = nullptr;
  for (auto* p : {&b->d_body[0], &b->d_body[1]}) if (*p) (void)hipFree(*p), *p = nullptr;
  if (b->d_force) (void)hipFree(b->d_force);
  b->d_force = nullptr;
  if (b->d_status) (void)hipFree(b->d_status);
  b->d_status = nullptr;
  if (b->d_omega) (void)hipFree(b->d_omega);
  b->d_omega = nullptr;
  if (b->d_asm_part) (void)hipFree(b->d_asm_part);
  b->d_asm_part = nullptr;
  for (auto* p : {&b->lu32.d_jobs}) if (*p) (void)hipFree(*p), *p = nullptr;
  for (auto* p : {&b->lu64.d_jobs}) if (*p) (void)hipFree(*p), *p = nullptr;
  for (int k = 0; k < 2; ++k) {
    free_trsv_group(&b->tr32[k]);
    free_trsv_group(&b->tr64[k]);
  }
  b->gens.clear();
  b->shapes.clear();
  b->built_rhs = -1;
}

// The influence-matrix outputs of a handle's assembly (as queue_factor).
MatOut mat_out(const nemoh_core* c) {
  MatOut out{c->nfa, c->D[0], c->S[0], c->D[1], c->S[1], nullptr, nullptr, 0};
  if (c->prec == NEMOH_MIXED) {
    out.dp32 = c->D32[0];
    out.dm32 = c->D32[1];
    out.ld32 = c->ld32;
  }
  return out;
}

// Job tables of the batch for its bodies' current buffers and nrhs problems
// each (device lock held: allocations).
int batch_build(nemoh_batch* b, const std::vector<int>& nrhs) {
  std::lock_guard<std::recursive_mutex> lk(device_lock());
  batch_free_tables(b);
  std::vector<AsmJob> asmj;
  std::vector<LuJob<float>> m32;
  std::vector<LuJob<double>> m64;
  std::vector<TrsvJob<float>> t32[2];
  std::vector<TrsvJob<double>> t64[2];
  std::vector<BodyJob> bj[2];
  int wg = 0;
  int bwg[4] = {0, 0, 0, 0};
  size_t foff = 0;
  b->force_off.assign(1, 0);
  for (size_t i = 0; i < b->bodies.size(); ++i) {
    nemoh_core* c = b->bodies[i];
    const int nbeta = nrhs[i] - c->nrad;
    const int nwg[4] = {(c->nfa * nrhs[i] + 255) / 256, (c->nfa + 15) / 16, c->nforce * nrhs[i], c->nforce * nbeta};
    for (int k = 0; k < c->nb; ++k) {
      BodyJob j{};
      j.V = c->V;
      j.B = c->B;
      j.R = c->R;
      j.Z = c->Z[k];
      j.F = c->F[k];
      j.Z2 = c->Z[1];
      j.F2 = c->F[1];
      j.R32 = c->R32;
      j.rs = c->d_rs + k;
      j.red = c->d_red;
      j.D = c->D[k];
      j.S = c->S[k];
      j.ld = c->nfa;
      j.n = c->nfa;
      j.nsym = c->nsym;
      j.nrhs = nrhs[i];
      j.nbeta = nbeta;
      j.nforce = c->nforce;
      j.tol = c->tol;
      j.rho = c->rho;
      j.sig = c->sig;
      j.pot = c->pot;
      j.pres = c->pres;
      j.pinc = c->d_pinc;
      j.nds = c->d_fnds;
      j.info = c->d_info;
      for (int q = 0; q < 4; ++q) {
        j.wg[q] = bwg[q];
        j.nwg[q] = nwg[q];
      }
      bj[k].push_back(j);
    }
    for (int q = 0; q < 4; ++q) bwg[q] += nwg[q];
    foff += (size_t)c->nforce * (nrhs[i] + nbeta);
    b->force_off.push_back(foff);
    const int nt = (c->nfa + 7) / 8, nw = nt * (nt + 1) / 2;
    AsmJob a{c->G, c->T, c->d_par, mat_out(c), nt, nw, wg, wg + pair_assembly_wg(c)};
    wg = a.wg1;
    asmj.push_back(a);
    ChainCtl* ctl = reinterpret_cast<ChainCtl*>(c->d_chain);
    for (int k = 0; k < c->nb; ++k) {
      if (b->prec == NEMOH_MIXED) {
        auto* A = reinterpret_cast<C2<float>*>(c->D32[k]);
        auto* P = reinterpret_cast<C2<float>*>(c->Pinv32[k]);
        m32.push_back(LuJob<float>{A, P, c->d_info, c->ld32, c->nfa, 0, 0, 0, PanelJob{0, 0, 0, 0, 0, 0},
                                   lu_form(c).band});
        t32[k].push_back(TrsvJob<float>{A, P, reinterpret_cast<C2<float>*>(c->R32),
                                        reinterpret_cast<C2<float>*>(c->W32), ctl, c->d_info + 1, c->d_rs + k,
                                        c->ld32, c->nfa, nrhs[i], 0, 0});
      } else {
        auto* A = reinterpret_cast<C2<double>*>(c->D[k]);
        auto* P = reinterpret_cast<C2<double>*>(c->Pinv[k]);
        m64.push_back(LuJob<double>{A, P, c->d_info, c->nfa, c->nfa, 0, 0, 0, PanelJob{0, 0, 0, 0, 0, 0},
                                    lu_form(c).band});
        t64[k].push_back(TrsvJob<double>{A, P, reinterpret_cast<C2<double>*>(c->Z[k]),
                                         reinterpret_cast<C2<double>*>(c->W), ctl, c->d_info + 1, nullptr, c->nfa,
                                         c->nfa, nrhs[i], 0, 0});
      }
    }
  }
  // each body's forces at its offset of the batch's force buffer
  std::vector<BodyJob>* jobs_k = bj;
  hipError_t e = hipMalloc((void**)&b->d_force, std::max<size_t>(1, foff) * sizeof(double2));
  for (int k = 0; k < b->nb && e == hipSuccess; ++k) {
    for (size_t i = 0; i < jobs_k[k].size(); ++i) jobs_k[k][i].force = b->d_force + b->force_off[i];
    e = hipMalloc((void**)&b->d_body[k], std::max<size_t>(1, jobs_k[k].size()) * sizeof(BodyJob));
    if (e == hipSuccess)
      e = hipMemcpy(b->d_body[k], jobs_k[k].data(), jobs_k[k].size() * sizeof(BodyJob), hipMemcpyHostToDevice);
  }
  for (int q = 0; q < 4; ++q) b->body_grid[q] = bwg[q];
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_status, std::max<size_t>(1, b->bodies.size()) * sizeof(BatchStatus));
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_omega, std::max<size_t>(1, b->bodies.size()) * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_asm_part, std::max<size_t>(1, asmj.size()) * sizeof(AsmJob));
  b->asm_host = asmj;
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_asm, std::max<size_t>(1, asmj.size()) * sizeof(AsmJob));
  if (e == hipSuccess) e = hipMemcpy(b->d_asm, asmj.data(), asmj.size() * sizeof(AsmJob), hipMemcpyHostToDevice);
  b->asm_count = (int)asmj.size();
  b->asm_grid = wg;
  if (e == hipSuccess) e = b->prec == NEMOH_MIXED ? build_lu_group(m32, &b->lu32) : build_lu_group(m64, &b->lu64);
  for (int k = 0; k < b->nb && e == hipSuccess; ++k)
    e = b->prec == NEMOH_MIXED ? build_trsv_group(t32[k], &b->tr32[k]) : build_trsv_group(t64[k], &b->tr64[k]);
  if (e != hipSuccess) return batch_fail(b, NEMOH_ERR_HIP, std::string("batch tables: ") + hipGetErrorString(e));
  for (nemoh_core* c : b->bodies) {
    b->gens.push_back(c->layout_gen);
    b->shapes.insert(b->shapes.end(), {c->nfa, c->nb, c->nforce, c->nrad});
  }
  b->built_rhs = nrhs.empty() ? 0 : nrhs[0];
  return NEMOH_OK;
}

// The job tables no longer match the bodies: a buffer moved, or the problem
// count, the mesh size or the degrees of freedom changed.
bool batch_stale(const nemoh_batch* b, const std::vector<int>& nrhs) {
  if (b->gens.size() != b->bodies.size() || b->shapes.size() != 4 * b->bodies.size() || nrhs.empty() ||
      b->built_rhs != nrhs[0])
    return true;
  for (size_t i = 0; i < b->bodies.size(); ++i) {
    const nemoh_core* c = b->bodies[i];
    const int* sh = b->shapes.data() + 4 * i;
    if (b->gens[i] != c->layout_gen || sh[0] != c->nfa || sh[1] != c->nb || sh[2] != c->nforce || sh[3] != c->nrad)
      return true;
  }
  return false;
}

// Points every body's handle at the batch stream for the duration of a call.
struct StreamSwap {
  nemoh_batch* b;
  std::vector<hipStream_t> saved;
  explicit StreamSwap(nemoh_batch* bb) : b(bb) {
    for (nemoh_core* c : b->bodies) {
      saved.push_back(c->st);
      c->st = b->st;
    }
  }
  ~StreamSwap() {
    for (size_t i = 0; i < b->bodies.size(); ++i) b->bodies[i]->st = saved[i];
  }
};

}  // namespace

extern "C" {

int nemoh_batch_create(nemoh_batch** out, int nbody, nemoh_core* const* bodies, int precision, int max_rhs) {
  if (!out) return NEMOH_ERR_ARG;
  *out = nullptr;
  if (nbody <= 0 || !bodies || (precision != NEMOH_FP64 && precision != NEMOH_MIXED) || max_rhs < 1 ||
      max_rhs > kMaxRhs)
    return NEMOH_ERR_ARG;
  auto* b = new nemoh_batch();
  b->prec = precision;
  b->max_rhs = max_rhs;
  auto bail = [&](int code, const std::string& msg) {
    create_error = msg;
    nemoh_batch_destroy(b);
    return code;
  };
  for (int i = 0; i < nbody; ++i) {
    nemoh_core* c = bodies[i];
    if (!c) return bail(NEMOH_ERR_ARG, "null body handle");
    if (!c->dofs_set) return bail(NEMOH_ERR_ARG, "body " + std::to_string(i) + ": nemoh_core_set_dofs not called");
    if (c->fp32_assembly) return bail(NEMOH_ERR_ARG, "body " + std::to_string(i) + ": FP32 assembly is not batched");
    if (i > 0 && (c->nsym != bodies[0]->nsym || c->depth != bodies[0]->depth || c->g != bodies[0]->g))
      return bail(NEMOH_ERR_ARG, "bodies of a batch share symmetry, depth and g");
    for (int j = 0; j < i; ++j)
      if (bodies[j] == c) return bail(NEMOH_ERR_ARG, "a handle appears twice in the batch");
    b->bodies.push_back(c);
  }
  b->nsym = bodies[0]->nsym;
  b->nb = bodies[0]->nb;
  std::lock_guard<std::recursive_mutex> lk(device_lock());
  if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess)
    return bail(NEMOH_ERR_HIP, "batch stream creation failed");
  for (nemoh_core* c : b->bodies) {
    c->prec = precision;
    if (ensure_matrices(c) || ensure_rhs(c, max_rhs)) return bail(NEMOH_ERR_HIP, c->err);
  }
  *out = b;
  return NEMOH_OK;
}

// Body i at frequency omega[i]: its forces (nforce_i x (nbeta + nrad_i),
// column-major complex) at forces + offset_i, with offset_i the sum of the
// earlier bodies' sizes; fk likewise (nforce_i x nbeta).  Bodies in both
// depth regimes take one assembly launch per regime.
static int batch_solve(nemoh_batch* b, const double* omega, int nbeta, const double* beta, double* forces,
                       double* fk) {
  const size_t nbody = b->bodies.size();
  for (size_t i = 0; i < nbody; ++i)
    if (!(std::fabs(omega[i]) >= 1.E-4)) return batch_fail(b, NEMOH_ERR_OMEGA, "ABS(WR) < 1.E-4");
  StreamSwap swap(b);
  const double pi = 4.0 * std::atan(1.0);
  std::vector<int> nrhs(nbody), fdv(nbody);
  std::vector<char> own(nbody);
  std::vector<double> wv(nbody);
  const bool mixed = b->prec == NEMOH_MIXED;
  // frequency set-up and body conditions of every body; a body whose pairs
  // fill the GPU on their own takes an assembly launch of its own, queued at
  // once so that it runs while the host sets up the next bodies (the
  // single-body kernel keeps its job in kernel arguments; the grouped one
  // reads it from memory and spills 4 registers: 4 % of a 40-body batch of
  // 4000 panels)
  for (size_t i = 0; i < nbody; ++i) {
    nemoh_core* c = b->bodies[i];
    if (c->prec != b->prec) return batch_fail(b, NEMOH_ERR_ARG, "a body's precision changed outside the batch");
    if (c->fp32_assembly)
      return batch_fail(b, NEMOH_ERR_ARG, "body " + std::to_string(i) + ": FP32 assembly is not batched");
    // the matrices of the body's current mesh: a handle may have been bound
    // to another mesh, or had its matrices freed, since the last call (its
    // buffers then move, and the job tables are rebuilt below)
    if (int r = ensure_matrices(c)) return batch_fail(b, r, "body " + std::to_string(i) + ": " + c->err);
    if (lu_form(c).rank != 128)
      return batch_fail(b, NEMOH_ERR_ARG, "body " + std::to_string(i) + ": batches factor in the rank-128 form");
    nrhs[i] = nbeta + c->nrad;
    if (nrhs[i] > b->max_rhs)
      return batch_fail(b, NEMOH_ERR_ARG, "body " + std::to_string(i) + ": more problems than the batch was made for");
    if (nrhs[i] != nrhs[0]) return batch_fail(b, NEMOH_ERR_ARG, "bodies of a batch solve the same number of problems");
    const double period = 2.0 * pi / omega[i];  // T as the Solver's main program forms it
    nemoh_freq f;
    bool bfd;
    if (int r = nemoh_core_freq_setup(c, period, &f)) return batch_fail(b, r, "body " + std::to_string(i) + ": " + c->err);
    if (int r = build_params(c, period, f.amh, f.nexp, f.ambda, f.ar, f.nq, f.qq, f.cq, &bfd, &wv[i]))
      return batch_fail(b, r, "body " + std::to_string(i) + ": " + c->err);
    if (int r = generate_waves(c, omega[i], nbeta, beta, false))
      return batch_fail(b, r, "body " + std::to_string(i) + ": " + c->err);
    fdv[i] = bfd ? 1 : 0;
    own[i] = pair_assembly_wg(c) >= kOwnAssemblyWg;
    c->w_prev = -1.0;  // the factors on the device are overwritten from here on
    if (own[i]) launch_assemble(c->G, c->T, c->d_par, fdv[i], c->nsym, mat_out(c), nullptr, 0, true, b->st);
    c->stats.cache_hit = 0;
    c->stats.refine_iterations = 0;
    c->stats.refine_residual = 0;
  }
  // (the headings' copies above read host memory: every return waits for the
  // stream)
  if (batch_stale(b, nrhs))
    if (int r = batch_build(b, nrhs)) return r;
  const int nbj = (int)nbody;
  const int* G = b->body_grid;
  b->omega_host = wv;
  (void)hipMemcpyAsync(b->d_omega, b->omega_host.data(), nbody * sizeof(double), hipMemcpyHostToDevice, b->st);
  launch_group_info_reset(b->d_body[0], nbj, b->st);
  // the assembly of the other bodies, one grouped launch per depth regime
  // present (the same pair arithmetic as the single-body kernel), then the
  // free-surface panels, per body
  const int nfd = (int)std::count(fdv.begin(), fdv.end(), 1);
  const bool any_big = std::count(own.begin(), own.end(), 1) > 0;
  if (!any_big && (nfd == 0 || nfd == nbj)) {
    launch_assemble_group(b->d_asm, b->asm_count, b->asm_grid, nfd ? 1 : 0, b->nsym, mixed, b->st);
  } else {
    // the grouped bodies of each regime, packed with new workgroup ranges:
    // finite depth first, then infinite depth
    b->asm_part.clear();
    int wg[2] = {0, 0}, cnt[2] = {0, 0};
    for (int reg = 1; reg >= 0; --reg)
      for (size_t i = 0; i < nbody; ++i) {
        AsmJob a = b->asm_host[i];
        const int nwg = a.wg1 - a.wg0;
        if (fdv[i] != reg || own[i]) continue;
        a.wg0 = wg[1 - reg];
        a.wg1 = a.wg0 + nwg;
        wg[1 - reg] = a.wg1;
        ++cnt[1 - reg];
        b->asm_part.push_back(a);
      }
    if (!b->asm_part.empty())
      (void)hipMemcpyAsync(b->d_asm_part, b->asm_part.data(), b->asm_part.size() * sizeof(AsmJob),
                           hipMemcpyHostToDevice, b->st);
    launch_assemble_group(b->d_asm_part, cnt[0], wg[0], 1, b->nsym, mixed, b->st);
    launch_assemble_group(b->d_asm_part + cnt[0], cnt[1], wg[1], 0, b->nsym, mixed, b->st);
  }
  for (size_t i = 0; i < nbody; ++i) {
    nemoh_core* c = b->bodies[i];
    if (c->nsurf > 0)
      launch_assemble_surface(c->G, c->T, c->d_par, fdv[i], c->nsym, mat_out(c), c->d_surf, c->nsurf, b->st);
  }
  // factorisation of every matrix, one launch per step
  const hipEvent_t* ev = nullptr;
  double ev_flops = 0.0, ev_bytes = 0.0;
  if (b->timing) {
    const size_t need = 2 * (mixed ? update_work(b->lu32, ev_flops, ev_bytes) : update_work(b->lu64, ev_flops, ev_bytes));
    if (b->upd_ev.size() < need) {  // created under the device lock, as the handles' events are
      std::lock_guard<std::recursive_mutex> lk(device_lock());
      while (b->upd_ev.size() < need) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return batch_fail(b, NEMOH_ERR_HIP, "event creation failed");
        b->upd_ev.push_back(e);
      }
    }
    ev = b->upd_ev.data();
  }
  const int nupd = mixed ? launch_lu_group(b->lu32, b->st, ev) : launch_lu_group(b->lu64, b->st, ev);
  const int mode = mixed ? NEMOH_MIXED : NEMOH_FP64;
  for (size_t i = 0; i < nbody; ++i) {  // the factors on the device belong to this frequency
    b->bodies[i]->w_prev = wv[i];
    b->bodies[i]->factored_prec = mode;
  }
  // solves, symmetry block by symmetry block (the handles' right-hand-side
  // buffers serve both blocks), each step one launch over every body
  int sweeps = kQueuedSweeps;  // as solve_block queues them (bodies needing more are redone below)
  for (nemoh_core* c : b->bodies) sweeps = std::min(sweeps, c->max_iter);
  for (int k = 0; k < b->nb; ++k) {
    const BodyJob* J = b->d_body[k];
    launch_group_elementwise(J, nbj, G[0], kGroupRhs, k == 0 ? 1.0 : -1.0, b->st);
    if (!mixed) {
      launch_group_elementwise(J, nbj, G[0], kGroupCopyBZ, 0.0, b->st);
      launch_trsv_group(b->tr64[k], b->bodies[0]->spin_limit, b->st);
    } else {
      launch_group_refine_reset(J, nbj, b->st);
      launch_group_elementwise(J, nbj, G[0], kGroupCopyBR, 0.0, b->st);
      launch_group_residual_max(J, nbj, 1, b->st);
      for (int sweep = 0; sweep < sweeps; ++sweep) {
        launch_group_elementwise(J, nbj, G[0], kGroupToF32, 0.0, b->st);
        launch_trsv_group(b->tr32[k], b->bodies[0]->spin_limit, b->st);
        launch_group_elementwise(J, nbj, G[0], sweep == 0 ? kGroupToF64Set : kGroupToF64Add, 0.0, b->st);
        launch_group_elementwise(J, nbj, G[0], kGroupCopyBRSkip, 0.0, b->st);
        launch_group_gemv(J, nbj, G[1], 0, b->st);
        launch_group_residual_max(J, nbj, 0, b->st);
      }
    }
    launch_group_gemv(J, nbj, G[1], 1, b->st);
  }
  // one look at every body's pivot report, wait status and refinement
  launch_group_status(b->d_body[0], nbj, b->nb, b->d_status, b->st);
  std::vector<BatchStatus> status(nbody);
  (void)hipMemcpyAsync(status.data(), b->d_status, nbody * sizeof(BatchStatus), hipMemcpyDeviceToHost, b->st);
  if (!batch_sync(b)) return batch_fail(b, NEMOH_ERR_HIP, "batch launch failed");
  if (ev) {  // the update launches are complete
    for (int u = 0; u < nupd; ++u) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, ev[2 * u], ev[2 * u + 1]) != hipSuccess)
        return batch_fail(b, NEMOH_ERR_HIP, "update timing failed");
      b->upd_ms += ms;
    }
    b->upd_flops += ev_flops;
    b->upd_bytes += ev_bytes;
    b->upd_launches += nupd;
  }
  // sigma, phi and pressure of every body; a body whose wait gave up, or
  // whose refinement needs more sweeps, is solved again on its own handle
  // from its factors (a cache hit), which overwrites its share
  launch_group_combine(b->d_body[0], nbj, G[0], b->d_omega, b->st);
  for (size_t i = 0; i < nbody; ++i) {
    nemoh_core* c = b->bodies[i];
    const BatchStatus& h = status[i];
    if (h.info[0]) {
      char buf[96];
      std::snprintf(buf, sizeof buf, "body %zu: PIVOT INFERIEUR A 1e-20 (row %d)", i, h.info[0]);
      c->w_prev = -1.0;
      return batch_fail(b, NEMOH_ERR_PIVOT, buf);
    }
    bool clean = h.info[1] == 0;
    if (mixed)
      for (int k = 0; k < c->nb; ++k) clean = clean && h.rs[k].done && !h.rs[k].bad;
    if (clean) {
      if (mixed) (void)refine_outcome(c, h.rs, 0);  // refinement statistics
      c->last_nrhs = nrhs[i];
    } else if (int r = run_pipeline(c, wv[i], nrhs[i])) {
      return batch_fail(b, r, "body " + std::to_string(i) + ": " + c->err);
    }
  }
  // forces and Froude-Krylov forces of every body, one launch each, one copy
  if (forces) launch_group_force(b->d_body[0], nbj, G[2], 0, b->st);
  if (fk && nbeta > 0) launch_group_force(b->d_body[0], nbj, G[3], 1, b->st);
  std::vector<double> hf(2 * std::max<size_t>(1, b->force_off.back()));
  (void)hipMemcpyAsync(hf.data(), b->d_force, b->force_off.back() * sizeof(double2), hipMemcpyDeviceToHost, b->st);
  if (!batch_sync(b)) return batch_fail(b, NEMOH_ERR_HIP, "batch force download failed");
  // device (nforce x cols) row-major -> the caller's column-major, per body
  size_t fo = 0, ko = 0;
  for (size_t i = 0; i < nbody; ++i) {
    const int nint = b->bodies[i]->nforce;
    auto unpack = [&](const double* src, double* dst, int cols) {
      for (int k = 0; k < nint; ++k)
        for (int q = 0; q < cols; ++q) {
          dst[2 * (k + (size_t)nint * q)] = src[2 * ((size_t)k * cols + q)];
          dst[2 * (k + (size_t)nint * q) + 1] = src[2 * ((size_t)k * cols + q) + 1];
        }
    };
    const double* src = hf.data() + 2 * b->force_off[i];
    if (forces && nint > 0) unpack(src, forces + 2 * fo, nrhs[i]);
    if (fk && nint > 0 && nbeta > 0) unpack(src + 2 * (size_t)nint * nrhs[i], fk + 2 * ko, nbeta);
    fo += (size_t)nint * nrhs[i];
    ko += (size_t)nint * nbeta;
  }
  return NEMOH_OK;
}

int nemoh_batch_solve_waves(nemoh_batch* b, double omega, int nbeta, const double* beta, double* forces, double* fk) {
  if (!b) return NEMOH_ERR_ARG;
  const std::vector<double> w(b->bodies.size(), omega);
  return batch_solve(b, w.data(), nbeta, beta, forces, fk);
}

int nemoh_batch_solve_waves_multi(nemoh_batch* b, const double* omega, int nbeta, const double* beta, double* forces,
                                  double* fk) {
  if (!b || !omega) return NEMOH_ERR_ARG;
  return batch_solve(b, omega, nbeta, beta, forces, fk);
}

int nemoh_batch_set_timing(nemoh_batch* b, int on) {
  if (!b) return NEMOH_ERR_ARG;
  b->timing = on != 0;
  b->upd_ms = b->upd_flops = b->upd_bytes = 0.0;
  b->upd_launches = 0;
  return NEMOH_OK;
}

int nemoh_batch_update_time(const nemoh_batch* b, double* ms, double* flops, double* bytes, int* launches) {
  if (!b) return NEMOH_ERR_ARG;
  if (ms) *ms = b->upd_ms;
  if (flops) *flops = b->upd_flops;
  if (bytes) *bytes = b->upd_bytes;
  if (launches) *launches = b->upd_launches;
  return NEMOH_OK;
}

const char* nemoh_batch_last_error(const nemoh_batch* b) {
  if (b) return b->err.c_str();
  return create_error.empty() ? "null batch" : create_error.c_str();
}

void nemoh_batch_destroy(nemoh_batch* b) {
  if (!b) return;
  std::lock_guard<std::recursive_mutex> lk(device_lock());
  if (b->st) (void)hipStreamSynchronize(b->st);
  batch_free_tables(b);
  for (hipEvent_t e : b->upd_ev) (void)hipEventDestroy(e);
  if (b->st) (void)hipStreamDestroy(b->st);
  delete b;
}

}  // extern "C"
