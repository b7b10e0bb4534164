/*
 * CPU restatement of the NEMOH Solver hot path (TEST INFRASTRUCTURE ONLY).
 *
 * This is the oracle: a plain-C, FP64, single-threaded restatement of the
 * reference algorithm, used by tests/ to check the HIP product and by
 * bench.py as the CPU-baseline leg.  Nothing in nemoh_amd/ links or calls
 * it.  Each function cites the reference file:line it restates.
 *
 * Parity pinning (what checks each function; see tests/test_oracle_*.py):
 *   ora_gg            golden GG samples from the FP64 reference build, plus
 *                     scipy exp1 closed form to the approximation's error
 *   ora_crek          golden CREK tables (4 x 328 x 46) from the reference
 *   ora_cint          golden CINT_INFD rules; Gauss-Legendre nodes to 1e-7
 *   ora_x0            golden X0 values; x*tanh(x)=K to the bisection tol
 *   ora_lisc          golden LISC (NEXP, AMBDA, AR) from the reference
 *   ora_prepare_mesh  golden GOMG outputs; area/normal closed forms
 *   ora_pair_infd/fd  golden per-pair dumps of VAV*+VNS* from the reference;
 *                     Rankine part vs adaptive quadrature of 1/r
 *   ora_session_*     golden Forces.dat of whole reference Solver runs;
 *                     GAUSSZ vs numpy.linalg on small matrices
 *   ora_potential     golden COMPUTE_POTENTIAL_DOMAIN values from the
 *                     reference; the Rankine part vs quadrature of 1/r
 *   ora_kochin        golden COMPUTE_KOCHIN values from the reference; the
 *                     one-point panel quadrature the reference keeps beside
 *                     it (COMPUTE_KOCHIN.f90:47-55) on a refined mesh
 */
#ifndef NEMOH_ORACLE_H
#define NEMOH_ORACLE_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORA_NR 328
#define ORA_NZ 46

/* Mesh + environment, the content of COM_VAR.f90:29-64 that the hot path
 * reads.  Arrays are owned by the caller; connectivity is 0-based. */
typedef struct {
  int np, nfa, nsym;                 /* NP, NFA, NSYMY */
  const double *x, *y, *z;           /* [np] */
  const int *m1, *m2, *m3, *m4;      /* [nfa], 0-based */
  const double *xg, *yg, *zg;        /* [nfa] collocation points (GOMG) */
  const double *xn, *yn, *zn;        /* [nfa] normals */
  const double *aire, *tdis;         /* [nfa] */
  double zer, depth, g, rho;
  const double *xr, *xz;             /* [328], [46] */
  const double *apd1x, *apd1z, *apd2x, *apd2z; /* [328*46], (i,j) -> i + 328*j */
} ora_ctx;

/* Per-pair output of VAV*+VNS* (COM_VAR.f90:87-90), in this order:
 * SP1 SM1 SP2 SM2 VSXP1 VSYP1 VSZP1 VSXM1 VSYM1 VSZM1
 * VSXP2 VSYP2 VSZP2 VSXM2 VSYM2 VSZM2 */
#define ORA_PAIR_N 16

/* Frequency-dependent scalars for the wave part. */
typedef struct {
  double t;           /* period T */
  double amh;         /* finite depth: AMH = X0(w^2 h / g) */
  int nexp;           /* finite depth: number of Prony terms */
  double ambda[31], ar[31];
  int nq;             /* theta rule length actually used (COM_VAR NQ) */
  double qq[251], cq[251];
} ora_freq;

double complex ora_gg(double complex z, double complex cex);
void ora_crek(double *xr, double *xz, double *apd1x, double *apd1z,
              double *apd2x, double *apd2z);
void ora_cint(double akk, int *n, double *qq, double *cq);
double ora_x0(double ak);
int ora_lisc(double akh, double amh, int *nexp, double *ambda, double *ar);
int ora_prepare_mesh(int np, int nfa, int nsym, const double *x, const double *y,
                     const double *z, const int *m1, const int *m2, const int *m3,
                     const int *m4, double *xg, double *yg, double *zg, double *xn,
                     double *yn, double *zn, double *aire, double *tdis, double *zer);

/* one (ISP, IFP) pair, 0-based; returns 1 if an undefined branch of the
 * reference was hit (AKZ > -1.5e-6 in the submerged branch), else 0 */
int ora_pair_infd(const ora_ctx *c, const ora_freq *f, int isp, int ifp, double *out);
int ora_pair_fd(const ora_ctx *c, const ora_freq *f, int isp, int ifp, double *out);

/* Solver session: the module state the reference keeps across SOLVE_BVP
 * calls (w_previous, NQ, AInv cache). */
typedef struct ora_session ora_session;
ora_session *ora_session_new(const ora_ctx *c);
void ora_session_free(ora_session *s);
/* SOLVE_BVP for one problem: NVEL[nfa * 2^nsym] -> PRESSURE[nfa * 2^nsym];
 * also returns ZIGB, ZIGS, ZPB, ZPS [nfa] when the pointers are non-NULL.
 * Returns 0, or a nonzero reference error code. */
int ora_solve_bvp(ora_session *s, double period, const double complex *nvel,
                  double complex *pressure, double complex *zigb, double complex *zigs,
                  double complex *zpb, double complex *zps);
int ora_session_undefined_pairs(const ora_session *s);

/* per-frequency set-up (theta rule, X0, LISC) with the module-NQ rule */
int ora_freq_setup(const ora_ctx *c, double period, int *nq_state, ora_freq *f, int *is_fd);
/* all nfa x nfa pairs, out[(i*nfa + j)*16 + k]; returns undefined count */
int ora_pairs_all(const ora_ctx *c, const ora_freq *f, int fd, double *out);
/* F(i, p) -= PRESSURE(c) * NDS(i, c) (NEMOH.f90:137-141) */
void ora_forces(int nint, int ntot, const double *nds, const double complex *pressure,
                double complex *force);

/* GAUSSZ (M_SOLVER.f90:31) on an n x m column-major matrix with leading
 * dimension lda; returns 0 or 1 (pivot below 1e-20). */
int ora_gaussz(double complex *a, int lda, int n, int m);

/* COMPUTE_KOCHIN (COMPUTE_KOCHIN.f90:25-153): Kochin function at angle
 * theta of the sources zigb, zigs [nfa] (zigs read only when nsym = 1) for
 * wave number kwave, reference point (xeff, yeff). */
double complex ora_kochin(const ora_ctx *c, double kwave, double xeff, double yeff, double theta,
                          const double complex *zigb, const double complex *zigs);
/* COMPUTE_POTENTIAL_DOMAIN (COMPUTE_POTENTIAL_DOMAIN.f90:24-82 with VVV,
 * VNV, VNVF and VVG of COMPUTE_GREEN_FREESURFACE.f90): potential at
 * (xc, yc, zc) of the sources zigb, zigs for period t; am0, amh, nexp,
 * ambda, ar as SOLVE_BVP passes them (kwave, AMH; LISC's fit in finite
 * depth).  Returns -1 where the reference stops (AKZ > -1.5e-6). */
int ora_potential(const ora_ctx *c, double t, double am0, double amh, int nexp, const double *ambda,
                  const double *ar, double xc, double yc, double zc, const double complex *zigb,
                  const double complex *zigs, double complex *phi);
int ora_potential_vec(const ora_ctx *c, double t, double am0, double amh, int nexp, const double *ambda,
                      const double *ar, int npt, const double *pts, const double complex *zigb,
                      const double complex *zigs, double complex *phi);
/* ora_kochin at n angles into out[n] (for ctypes). */
void ora_kochin_vec(const ora_ctx *c, double kwave, double xeff, double yeff, int n, const double *theta,
                    const double complex *zigb, const double complex *zigs, double complex *out);

#ifdef __cplusplus
}
#endif
#endif
