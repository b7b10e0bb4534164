// The reference's case files: Nemoh.cal, body meshes, the preProcessor's
// outputs (mesh/L10.dat, L12.dat, Mesh.tec, Integration.dat, Kochin.dat,
// Freesurface.dat, Normalvelocities.dat, results/FKForce.dat, FKForce.tec,
// index.dat) and the Solver's (results/Forces.dat, Kochin.*.dat,
// freesurface.*.dat, pressure.*.dat), read and written record for record as
// the reference reads and writes them (fortran_io.h gives the text rules).
#pragma once
#include <complex>
#include <string>
#include <vector>

#include "fortran_io.h"

namespace nemoh {
namespace casefiles {

using cplx = std::complex<double>;

// One radiation or force-integration line of Nemoh.cal: ICase (1
// translation, 2 rotation), Direction, Axis; Body and Mode as the
// preProcessor numbers them (preProcessor/Main.f90:43-49, 170-184).
struct DofCase {
  int icase = 0;
  double dir[3] = {0, 0, 0}, axis[3] = {0, 0, 0};
  int body = 0, mode = 0;
};

struct BodyCal {
  std::string mesh_file;
  int npoints = 0, npanels = 0;
  std::vector<DofCase> rad, force;
};

// Nemoh.cal as the preProcessor (Main.f90:93-184, Mesh.f90:120-160,
// Environment.f90:34-42) and the postProcessor (IRF, Plot_WaveElevation)
// read it.
struct NemohCal {
  double rho = 0, g = 0, depth = 0, xeff = 0, yeff = 0;
  std::vector<BodyCal> bodies;
  int nw = 0, nbeta = 0;
  double wmin = 0, wmax = 0, betamin = 0, betamax = 0;
  int switch_potential = 0;
  int ntheta = 0;
  double thetamin = 0, thetamax = 0;
  int nx = 0, ny = 0;
  double lx = 0, ly = 0;
  std::vector<double> w;     // rad/s (Main.f90:113-120)
  std::vector<double> beta;  // radians (Main.f90:121-129)
  std::vector<DofCase> rad, force;  // all bodies, in order
};
NemohCal read_nemoh_cal(const std::string& path);

// The preProcessor's mesh (Mesh.f90:93-232): every body merged, vertices as
// the mesh files give them, its own panel centres, normals and areas.
struct PreMesh {
  int isym = 0, nbodies = 0;
  std::vector<double> x, y, z;
  std::vector<int> p;       // 4 per panel, 1-based, offset per body
  std::vector<int> cpanel;  // body of each panel (1-based)
  std::vector<double> xm, nrm, a;  // 3, 3, 1 per panel
  int np() const { return (int)x.size(); }
  int nfa() const { return (int)cpanel.size(); }
};
// mesh files are opened relative to base_dir (the reference opens them
// relative to its working directory, the case directory when ID.dat is ".")
PreMesh read_body_meshes(const NemohCal& cal, const std::string& nemoh_cal_path, const std::string& base_dir);

std::string l12_text(const PreMesh& m);
std::string l10_text(const PreMesh& m);
std::string mesh_tec_text(const PreMesh& m);
std::string kochin_dat_text(const NemohCal& c);
std::string freesurface_dat_text(const NemohCal& c);
std::string index_dat_text(const NemohCal& c);
// fnds(nint, ntot) row-major
std::string integration_text(int nint, int ntot, const std::vector<double>& fnds);
// nvel(ntot, nprob), problem-major: nvel[q * ntot + c]
std::string normal_velocities_text(const NemohCal& c, int ntot, const std::vector<cplx>& nvel, int switch_fs,
                                   int switch_kochin);
// fk(nw, nbeta, nint): fk[(i * nbeta + j) * nint + k]
std::string fkforce_dat_text(const NemohCal& c, const std::vector<cplx>& fk);
std::string fkforce_tec_text(const NemohCal& c, const std::vector<cplx>& fk);

// The Solver's inputs, read as the Solver reads them.
struct SolverMesh {  // Common/Mesh.f90:88-134 (L10.dat header, L12.dat)
  int isym = 0, nbodies = 0;
  std::vector<double> x, y, z;
  std::vector<int> p;  // 4 per panel, 1-based
  std::vector<int> cpanel;
  int np() const { return (int)x.size(); }
  int nfa() const { return (int)cpanel.size(); }
};
SolverMesh read_solver_mesh(const std::string& dir);

struct Environment {  // INITIALIZATION.f90:44-50
  double rho = 0, g = 0, depth = 0, xeff = 0, yeff = 0;
};
Environment read_environment(const std::string& nemoh_cal_path);

struct BodyConditions {  // Solver/Core/Bodyconditions.f90:65-88
  int nprob = 0;
  std::vector<double> omega, type;
  std::vector<int> sw_potential, sw_fs, sw_kochin;
  std::vector<cplx> nvel;  // problem-major: nvel[q * ntot + c]
};
BodyConditions read_body_conditions(const std::string& path, int ntot);

// Integration.dat (NEMOH.f90:63-69): nds(nint, ntot) row-major
std::vector<double> read_integration(const std::string& path, int ntot, int* nint);
std::vector<double> read_kochin_dat(const std::string& path);  // NEMOH.f90:96-104
struct FreeSurfaceMesh {  // NEMOH.f90:106-118
  int npoints = 0, npanels = 0;
  std::vector<double> x, y;
  std::vector<int> p;  // 4 per panel
};
FreeSurfaceMesh read_freesurface_dat(const std::string& path);

// The Solver's outputs.  force(nint, nprob): force[q * nint + k]
std::string forces_text(const BodyConditions& bc, int nint, const std::vector<cplx>& force);
std::string kochin_text(const std::vector<double>& theta, const cplx* h);  // OUTPUT.f90:25-40
std::string freesurface_text(const FreeSurfaceMesh& fs, const cplx* eta);  // OUTPUT.f90:42-63
std::string pressure_text(const SolverMesh& m, const cplx* pressure);     // OUTPUT.f90:65-101
std::string problem_tag(int number);  // '(I5)' of the problem number

// The postProcessor (postProcessor/Main.f90:21-101) on a case directory
// holding the Solver's results: RadiationCoefficients.tec,
// DiffractionForce.tec, ExcitationForce.tec, CA.dat, CM.dat, Fe.dat and,
// when Nemoh.cal asks for it, IRF.tec.
void postprocess(const std::string& dir);

}  // namespace casefiles
}  // namespace nemoh
