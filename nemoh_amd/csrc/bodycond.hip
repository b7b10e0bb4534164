// Body conditions, force weights and Froude-Krylov pressures on the device:
// the preProcessor's share of a frequency solve (preProcessor/Mesh.f90:
// 207-233, BodyConditions.f90:27-137, Integration.f90:27-86, Main.f90:
// 199-243, Common/Environment.f90:86-123), so that a standalone solve needs
// no per-frequency input from the host.
//
// The preProcessor has its own panel geometry (centres and normals of two
// triangles P1P2P4 and P3P4P2), distinct from the Solver's (PREPARE_MESH);
// both are kept.  Products and sums follow the reference's evaluation order
// with contraction off, so apart from the transcendental functions the
// results round as the reference's do.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace nemoh {

// Panel centre XM, unit normal N and area A (preProcessor/Mesh.f90:207-233)
// into pg = [xm(3n) | nrm(3n) | a(n)], component-major.
__global__ __launch_bounds__(256) void panel_props_kernel(Geometry G, double* __restrict__ pg) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = G.n;
  if (i >= n) return;
  double P[4][3];
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    P[l][0] = G.vx[4 * (size_t)i + l];
    P[l][1] = G.vy[4 * (size_t)i + l];
    P[l][2] = G.vz[4 * (size_t)i + l];
  }
  double u[3], v[3], w1[3], w2[3];
  for (int j = 0; j < 3; ++j) {
    u[j] = P[1][j] - P[0][j];
    v[j] = P[3][j] - P[1][j];
  }
  w1[0] = u[1] * v[2] - u[2] * v[1];
  w1[1] = u[2] * v[0] - u[0] * v[2];
  w1[2] = u[0] * v[1] - u[1] * v[0];
  const double a1 = 0.5 * sqrt(w1[0] * w1[0] + w1[1] * w1[1] + w1[2] * w1[2]);
  for (int j = 0; j < 3; ++j) {
    u[j] = P[3][j] - P[2][j];
    v[j] = P[1][j] - P[2][j];
  }
  w2[0] = u[1] * v[2] - u[2] * v[1];
  w2[1] = u[2] * v[0] - u[0] * v[2];
  w2[2] = u[0] * v[1] - u[1] * v[0];
  const double a2 = 0.5 * sqrt(w2[0] * w2[0] + w2[1] * w2[1] + w2[2] * w2[2]);
  const double a = a1 + a2;
  const double third = 1.0 / 3;
  for (int j = 0; j < 3; ++j)
    pg[(size_t)j * n + i] =
        third * (P[0][j] + P[1][j] + P[3][j]) * a1 / a + third * (P[1][j] + P[2][j] + P[3][j]) * a2 / a;
  const double s[3] = {w1[0] + w2[0], w1[1] + w2[1], w1[2] + w2[2]};
  const double ns = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  for (int j = 0; j < 3; ++j) pg[(size_t)(3 + j) * n + i] = s[j] / ns;
  pg[(size_t)6 * n + i] = a;
}

// n . (motion of degree of freedom d) at panel c of the full surface (c >= n:
// the y-image of panel c - n), BodyConditions.f90:39-80 / Integration.f90:
// 38-77; zero on panels of other bodies.
__device__ __forceinline__ double mode_normal(const double* __restrict__ pg, const int* __restrict__ body, int n,
                                              int c, const Dof& d) {
#pragma clang fp contract(off)
  const bool img = c >= n;
  const int i = img ? c - n : c;
  if (body[i] != d.body) return 0.0;
  const double x = pg[i], y = pg[(size_t)n + i], z = pg[(size_t)2 * n + i];
  const double n1 = pg[(size_t)3 * n + i], n2 = pg[(size_t)4 * n + i], n3 = pg[(size_t)5 * n + i];
  double vel[3];
  if (d.kind == 1) {
    vel[0] = d.dir[0];
    vel[1] = d.dir[1];
    vel[2] = d.dir[2];
  } else {
    const double ym = img ? -y : y;
    vel[0] = d.dir[1] * (z - d.axis[2]) - d.dir[2] * (ym - d.axis[1]);
    vel[1] = d.dir[2] * (x - d.axis[0]) - d.dir[0] * (z - d.axis[2]);
    vel[2] = d.dir[0] * (ym - d.axis[1]) - d.dir[1] * (x - d.axis[0]);
  }
  return img ? n1 * vel[0] - n2 * vel[1] + n3 * vel[2] : n1 * vel[0] + n2 * vel[1] + n3 * vel[2];
}

// Radiation right-hand sides: out[c * ld + col0 + q] = (n . motion of dof q, 0).
__global__ __launch_bounds__(256) void radiation_kernel(const double* __restrict__ pg, const int* __restrict__ body,
                                                        const Dof* __restrict__ dofs, int nrad, int n, int ntot,
                                                        double2* __restrict__ out, int ld, int col0) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ntot * nrad) return;
  const int c = e / nrad, q = e % nrad;
  out[(size_t)c * ld + col0 + q] = make_double2(mode_normal(pg, body, n, c, dofs[q]), 0.0);
}

// Force weights NDS(k, c) = (n . motion of dof k) A (Integration.f90:27-86),
// row-major (nforce x ntot).
__global__ __launch_bounds__(256) void nds_kernel(const double* __restrict__ pg, const int* __restrict__ body,
                                                  const Dof* __restrict__ dofs, int nforce, int n, int ntot,
                                                  double* __restrict__ nds) {
#pragma clang fp contract(off)
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ntot * nforce) return;
  const int k = e / ntot, c = e % ntot;
  const int i = c >= n ? c - n : c;
  nds[e] = mode_normal(pg, body, n, c, dofs[k]) * pg[(size_t)6 * n + i];
}

// Incident regular wave at panel c (Compute_Wave, Common/Environment.f90:
// 86-99, with CIH/SIH :102-123): pressure p and NVEL = -(V . n) for heading
// b (BodyConditions.f90:110-133); out[c * ld + b], pinc[c * nbeta + b].
struct WaveArgs {
  double k, w, g, rho, depth, xeff, yeff;
};
__global__ __launch_bounds__(256) void diffraction_kernel(const double* __restrict__ pg, WaveArgs W,
                                                          const double* __restrict__ beta, int nbeta, int n, int ntot,
                                                          double2* __restrict__ out, int ld,
                                                          double2* __restrict__ pinc) {
#pragma clang fp contract(off)
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ntot * nbeta) return;
  const int c = e / nbeta, b = e % nbeta;
  const bool img = c >= n;
  const int i = img ? c - n : c;
  const double x = pg[i], y = img ? -pg[(size_t)n + i] : pg[(size_t)n + i], z = pg[(size_t)2 * n + i];
  const double n1 = pg[(size_t)3 * n + i], n2 = pg[(size_t)4 * n + i], n3 = pg[(size_t)5 * n + i];
  const double cb = cos(beta[b]), sb = sin(beta[b]);
  const double wbar = (x - W.xeff) * cb + (y - W.yeff) * sb;
  double cih, sih;
  const double kh = W.k * W.depth;
  if (kh <= 20.0 && kh > 0.0) {
    const double ch = cosh(kh);
    cih = cosh(W.k * (z + W.depth)) / ch;
    sih = sinh(W.k * (z + W.depth)) / ch;
  } else {
    cih = sih = exp(W.k * z);
  }
  double es, ec;
  sincos(W.k * wbar, &es, &ec);  // CEXP(II*k*wbar)
  const double sp = W.rho * W.g * cih;
  const double sx = W.g / W.w * W.k * cb * cih, sy = W.g / W.w * W.k * sb * cih;
  const double sz = W.g / W.w * W.k * sih;  // Vz = -(i sz) e
  const double vxr = sx * ec, vxi = sx * es, vyr = sy * ec, vyi = sy * es;
  const double vzr = sz * es, vzi = -(sz * ec);
  double nr, ni;
  if (img) {
    nr = vxr * n1 - vyr * n2 + vzr * n3;
    ni = vxi * n1 - vyi * n2 + vzi * n3;
  } else {
    nr = vxr * n1 + vyr * n2 + vzr * n3;
    ni = vxi * n1 + vyi * n2 + vzi * n3;
  }
  out[(size_t)c * ld + b] = make_double2(-nr, -ni);
  pinc[(size_t)c * nbeta + b] = make_double2(sp * ec, sp * es);
}

void launch_panel_props(const Geometry& G, double* pg, hipStream_t st) {
  panel_props_kernel<<<(G.n + 255) / 256, 256, 0, st>>>(G, pg);
}
void launch_radiation(const double* pg, const int* body, const Dof* dofs, int nrad, int n, int ntot, double2* out,
                      int ld, int col0, hipStream_t st) {
  if (nrad > 0)
    radiation_kernel<<<(ntot * nrad + 255) / 256, 256, 0, st>>>(pg, body, dofs, nrad, n, ntot, out, ld, col0);
}
void launch_nds(const double* pg, const int* body, const Dof* dofs, int nforce, int n, int ntot, double* nds,
                hipStream_t st) {
  if (nforce > 0) nds_kernel<<<(ntot * nforce + 255) / 256, 256, 0, st>>>(pg, body, dofs, nforce, n, ntot, nds);
}
void launch_diffraction(const double* pg, double k, double w, double g, double rho, double depth, double xeff,
                        double yeff, const double* beta, int nbeta, int n, int ntot, double2* out, int ld,
                        double2* pinc, hipStream_t st) {
  if (nbeta > 0)
    diffraction_kernel<<<(ntot * nbeta + 255) / 256, 256, 0, st>>>(
        pg, WaveArgs{k, w, g, rho, depth, xeff, yeff}, beta, nbeta, n, ntot, out, ld, pinc);
}

}  // namespace nemoh
