// Host-side launchers implemented in the .hip translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace cfd {

template <class T> void launch_grad(const cfd_mesh_s* m, const void* phi, void* out, int mask, int wall, hipStream_t s);
template <class T> void launch_lapl(const cfd_mesh_s* m, const void* phi, void* out, int mask, hipStream_t s);
template <class T>
void launch_stab(const cfd_mesh_s* m, const void* phi, void* out, int mask, void* scratch3, hipStream_t s);
template <class T>
void launch_heat(const cfd_mesh_s* m, void* Ta, void* Tb, double alpha, const void* af, double dt, int64_t nsteps,
                 hipStream_t s);
template <class T>
void launch_wave(const cfd_mesh_s* m, void* const a[4], void* const b[4], double K, const void* Kf, double dt,
                 int64_t nsteps, hipStream_t s);

template <class T>
void launch_wave_march(const cfd_mesh_s* m, void* const a[4], void* const b[4], double K, const void* Kf, double dt,
                       int64_t nsteps, int ty, int L, int remap, hipStream_t s);

template <class T>
void launch_ewise(int op, const void* a, const void* b, const void* aux, double s, void* out, int64_t n,
                  hipStream_t st);
// min over cells of dl / (sqrt(gamma p / rho) + |u|) written to dev_out (one T)
template <class T>
void launch_cfl(const void* u, const void* v, const void* w, const void* p, const void* rho, T dl, T gamma,
                int64_t n, void* partials, int nblocks, void* dev_out, hipStream_t s);
template <class T> void launch_nonfinite(const void* x, int64_t n, unsigned long long* dev_count, hipStream_t s);

// pool helpers (runtime.hip)
void* pool_alloc(cfd_ctx_t ctx, int64_t bytes, hipStream_t s);
void pool_free(cfd_ctx_t ctx, void* p, hipStream_t s);

}  // namespace cfd
