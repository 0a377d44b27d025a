// Host-side launchers implemented in the .hip translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace cfd {

// Sensor cells of the trace-recording forward step and the adjoint step,
// passed by value as a kernel argument (row = cell / Z = x * Y + y, z = cell % Z).
constexpr int kMaxSensors = 64;
struct SensorList {
    int n;
    int row[kMaxSensors];
    int x[kMaxSensors];
    int y[kMaxSensors];
    int z[kMaxSensors];
    int64_t cell[kMaxSensors];
};
SensorList make_sensors(const cfd_mesh_s* m, const int64_t* cells, int n);

template <class T>
void launch_wave_trace_step(const cfd_mesh_s* m, const void* const in[4], void* const out[4], double K,
                            const void* Kf, double dt, const SensorList& sens, void* trace, hipStream_t s);
template <class T>
void launch_wave_adj_step(const cfd_mesh_s* m, const void* const in[4], void* const out[4], double K,
                          const void* Kf, double dt, const SensorList& sens, const void* src,
                          const void* const velt[3], void* gK, hipStream_t s);
template <class T> void launch_scatter_add(void* lp, const SensorList& sens, const void* src, hipStream_t s);
template <class T>
void run_wave_traces(const cfd_mesh_s* m, void* const a[4], void* const b[4], double K, const void* Kf, double dt,
                     int64_t nsteps, const SensorList& sens, void* traces, void* ckpt, int64_t every, hipStream_t s);
template <class T>
void run_wave_adjoint(const cfd_mesh_s* m, void* const la[4], void* const lb[4], double K, const void* Kf,
                      double dt, int64_t nsteps, const SensorList& sens, const void* tgrad, const void* ckpt,
                      int64_t every, void* gK, hipStream_t s);

template <class T>
void run_heat_traces(const cfd_mesh_s* m, void* Ta, void* Tb, double alpha, const void* af, double dt,
                     int64_t nsteps, const SensorList& sens, void* traces, void* ckpt, int64_t every, hipStream_t s);
template <class T>
void run_heat_adjoint(const cfd_mesh_s* m, void* la, void* lb, double alpha, const void* af, double dt,
                      int64_t nsteps, const SensorList& sens, const void* tgrad, const void* ckpt, int64_t every,
                      void* gA, hipStream_t s);

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
                       int64_t nsteps, int variant, int ty, int L, int remap, hipStream_t s);

// Row form of the acoustic step (wave_row.hip): z extent a multiple of 4, at
// most 256.  order: bit 0 XCD-contiguous tiles, bit 1 tiles last to first;
// cache: bit 0 non-temporal stores, bit 1 non-temporal loads.  Returns false
// (nothing launched) for meshes or field pointers it cannot take.
bool wave_row_supported(int64_t Z);
template <class T>
bool launch_wave_row(const UniformGeo<T>& g, void* const* in, void* const* out, double K, const void* Kf, double dt,
                     int ty, int L, int order, int cache, hipStream_t s);

// Two heat steps per launch in row form (heat_row.hip), constant alpha; nw
// waves per workgroup (8, 12 or 16).  Returns false (nothing launched) for
// meshes or field pointers it cannot take.
bool heat_row_supported(int64_t Z, int nw);
template <class T>
bool launch_heat_row2(const UniformGeo<T>& g, const void* in, void* out, double alpha, double dt, int nw, int L,
                      hipStream_t s);

// Two steps per pass of the row form (wave_row2.hip), constant K.  Returns
// false (nothing launched) for meshes or field pointers it cannot take.
template <class T>
bool launch_wave_row2(const UniformGeo<T>& g, void* const* in, void* const* out, double K, double dt, int ty, int L,
                      int order, bool nt, hipStream_t s);

// Up to K heat steps per launch in LDS (heat_multi.hip): B x B x min(Z, 64-2K)
// output boxes, B 4 or 8, K 2 or 4 (B 8 with K 4 in FP32 only), constant alpha.
bool heat_multi_supported(int B, int K);
template <class T>
void launch_heat_multi(const UniformGeo<T>& g, const void* in, void* out, double alpha, double dt, int B, int K,
                       int ns, hipStream_t s);

template <class T>
void launch_euler(cfd_mesh_s* m, void* const a[5], void* const b[5], double gamma, double cfl, int64_t nsteps,
                  void* dt_hist, hipStream_t s);
template <class T>
void launch_ueuler(cfd_mesh_s* m, void* const a[5], void* const b[5], double gamma, double cfl, int64_t nsteps,
                   void* dt_hist, hipStream_t s);

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
