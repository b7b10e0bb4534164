// Library-internal entry points of the solver core, used by the request
// loop (server.cpp); not part of the C ABI.
#pragma once
#include <cstddef>

struct nemoh_core;

namespace nemoh {

// Re-targets a handle at another mesh (arguments as nemoh_core_create, geo =
// XG YG ZG XN YN ZN AIRE TDIS), keeping its stream, wave tables and device
// buffers; buffers are replaced only by larger ones.  The factorisation
// cache and the module NQ are reset as for a new handle.
int core_bind_mesh(nemoh_core* c, int np, int nfa, int nsym, const double* x, const double* y, const double* z,
                   const int* m1, const int* m2, const int* m3, const int* m4, const double* const geo[8], double zer,
                   double g, double rho, double depth);

// The module NQ each frequency of a sequential standalone session over
// periods[0..nfreq) starts from (an infinite-depth frequency reuses the NQ
// of the last finite-depth one), from the handle's current NQ.
int core_nq_sequence(nemoh_core* c, int nfreq, const double* periods, int* nq_before);

// Number of visible HIP devices (0 when there is none or HIP fails).
int core_device_count();

// Sets the module NQ carried into the next standalone frequency set-up.
void core_set_nq_state(nemoh_core* c, int nq);

// Device bytes the handle holds now (every buffer it allocated).
size_t core_held_bytes(const nemoh_core* c);

// Device bytes a handle needs to solve one frequency of a mesh of nfa
// panels (np vertices, symmetry nsym) in `precision` with nrhs problems:
// influence matrices, factors, right-hand-side workspace, geometry and body
// conditions.
size_t core_footprint_bytes(int np, int nfa, int nsym, int precision, int nrhs);

// Frees the handle's matrices, factors and right-hand-side workspace (they
// are allocated again by the next solve); waits for its stream first.
int core_release_matrices(nemoh_core* c);

}  // namespace nemoh
