// ORACLE — test infrastructure only (see cfd_oracle.hpp).
// extern "C" surface of the CPU oracle so pytest and the CPU-baseline leg of
// bench.py can drive it through ctypes.  Scalars arrive as double and are
// converted to the mesh's scalar type once, at entry.
#include <chrono>
#include <cstdint>
#include <cstring>
#include <string>

#include <omp.h>

#include "cfd_oracle.hpp"

namespace {
struct OMesh {
    int dtype;  // 0 = float, 1 = double
    cfdo::HexMesh<float> f;
    cfdo::HexMesh<double> d;
};

template <class T> cfdo::HexMesh<T>& pick(OMesh* m);
template <> cfdo::HexMesh<float>& pick<float>(OMesh* m) { return m->f; }
template <> cfdo::HexMesh<double>& pick<double>(OMesh* m) { return m->d; }

template <class T>
void build(OMesh* om, int64_t X, int64_t Y, int64_t Z, double dx, double dy, double dz, const double* perturb) {
    auto& m = pick<T>(om);
    std::vector<T> pt;
    if (perturb) {
        const int64_t V = (X + 1) * (Y + 1) * (Z + 1);
        pt.resize(3 * V);
        for (int64_t k = 0; k < 3 * V; ++k) pt[k] = T(perturb[k]);
    }
    cfdo::build_topology<T>(m, X, Y, Z, T(dx), T(dy), T(dz), perturb ? pt.data() : nullptr);
    cfdo::compute_geometry<T>(m);
    cfdo::compute_fv_geometry<T>(m);
}

template <class V> int copy_vec(const std::vector<V>& v, void* out) {
    std::memcpy(out, v.data(), v.size() * sizeof(V));
    return 0;
}

template <class T> int table(OMesh* om, const std::string& nm, void* out) {
    auto& m = pick<T>(om);
    if (nm == "ids") return copy_vec(m.ids, out);
    if (nm == "ids_bf") return copy_vec(m.ids_bf, out);
    if (nm == "nx") return copy_vec(m.nx, out);
    if (nm == "ny") return copy_vec(m.ny, out);
    if (nm == "nz") return copy_vec(m.nz, out);
    if (nm == "nax") return copy_vec(m.nax, out);
    if (nm == "nay") return copy_vec(m.nay, out);
    if (nm == "naz") return copy_vec(m.naz, out);
    if (nm == "area") return copy_vec(m.area6, out);
    if (nm == "alpha") return copy_vec(m.alpha_d, out);
    if (nm == "len") return copy_vec(m.len6, out);
    if (nm == "vol") return copy_vec(m.vol, out);
    if (nm == "is_bnd") return copy_vec(m.is_bnd, out);
    if (nm == "ccx") return copy_vec(m.ccx, out);
    if (nm == "ccy") return copy_vec(m.ccy, out);
    if (nm == "ccz") return copy_vec(m.ccz, out);
    if (nm == "vx") return copy_vec(m.vx, out);
    if (nm == "vy") return copy_vec(m.vy, out);
    if (nm == "vz") return copy_vec(m.vz, out);
    if (nm == "face_verts") return copy_vec(m.face_verts, out);
    if (nm == "face_cells") return copy_vec(m.face_cells, out);
    if (nm == "cell_faces") return copy_vec(m.cell_faces, out);
    if (nm == "cell_verts") return copy_vec(m.cell_verts, out);
    if (nm == "fnx") return copy_vec(m.fnx, out);
    if (nm == "fny") return copy_vec(m.fny, out);
    if (nm == "fnz") return copy_vec(m.fnz, out);
    if (nm == "farea") return copy_vec(m.farea, out);
    if (nm == "svx") return copy_vec(m.svx, out);
    if (nm == "svy") return copy_vec(m.svy, out);
    if (nm == "svz") return copy_vec(m.svz, out);
    if (nm == "sva") return copy_vec(m.sva, out);
    if (nm == "svol") return copy_vec(m.svol, out);
    return -1;
}
}  // namespace

extern "C" {

void* oracle_mesh_create(int64_t X, int64_t Y, int64_t Z, double dx, double dy, double dz, int dtype,
                         const double* perturb) {
    auto* om = new OMesh();
    om->dtype = dtype;
    if (dtype == 0) build<float>(om, X, Y, Z, dx, dy, dz, perturb);
    else build<double>(om, X, Y, Z, dx, dy, dz, perturb);
    return om;
}

void* oracle_mesh_create_conn(int64_t N, int64_t F, int64_t V, const double* verts, const int64_t* face_verts,
                              const int64_t* face_cells, const int64_t* cell_faces, const int64_t* cell_verts,
                              int dtype) {
    auto* om = new OMesh();
    om->dtype = dtype;
    if (dtype == 0) {
        cfdo::build_from_connectivity<float>(om->f, N, F, V, verts, face_verts, face_cells, cell_faces, cell_verts);
        cfdo::compute_geometry<float>(om->f);
        cfdo::compute_fv_geometry<float>(om->f);
    } else {
        cfdo::build_from_connectivity<double>(om->d, N, F, V, verts, face_verts, face_cells, cell_faces, cell_verts);
        cfdo::compute_geometry<double>(om->d);
        cfdo::compute_fv_geometry<double>(om->d);
    }
    return om;
}

void oracle_mesh_destroy(void* h) { delete static_cast<OMesh*>(h); }

void oracle_mesh_sizes(void* h, int64_t* out) {
    auto* om = static_cast<OMesh*>(h);
    if (om->dtype == 0) { out[0] = om->f.N; out[1] = om->f.F; out[2] = om->f.V; }
    else { out[0] = om->d.N; out[1] = om->d.F; out[2] = om->d.V; }
}

int oracle_mesh_table(void* h, const char* name, void* out) {
    auto* om = static_cast<OMesh*>(h);
    return om->dtype == 0 ? table<float>(om, name, out) : table<double>(om, name, out);
}

#define ORACLE_DISPATCH(om, CALL_F, CALL_D) \
    do { if ((om)->dtype == 0) { CALL_F; } else { CALL_D; } } while (0)

void oracle_grad(void* h, const void* phi, int mask, int wall_axis, void* out) {
    auto* om = static_cast<OMesh*>(h);
    ORACLE_DISPATCH(om, cfdo::grad(om->f, (const float*)phi, mask, wall_axis, (float*)out),
                    cfdo::grad(om->d, (const double*)phi, mask, wall_axis, (double*)out));
}

void oracle_lapl(void* h, const void* phi, int mask, void* out) {
    auto* om = static_cast<OMesh*>(h);
    ORACLE_DISPATCH(om, cfdo::lapl(om->f, (const float*)phi, mask, (float*)out),
                    cfdo::lapl(om->d, (const double*)phi, mask, (double*)out));
}

void oracle_stab(void* h, const void* phi, int mask, void* out) {
    auto* om = static_cast<OMesh*>(h);
    ORACLE_DISPATCH(om, cfdo::stab(om->f, (const float*)phi, mask, (float*)out),
                    cfdo::stab(om->d, (const double*)phi, mask, (double*)out));
}

void oracle_heat_run(void* h, void* Tf, double alpha, const void* alpha_field, double dt, int64_t steps) {
    auto* om = static_cast<OMesh*>(h);
    ORACLE_DISPATCH(om,
                    cfdo::heat_run(om->f, (float*)Tf, float(alpha), (const float*)alpha_field, float(dt), steps),
                    cfdo::heat_run(om->d, (double*)Tf, alpha, (const double*)alpha_field, dt, steps));
}

void oracle_wave_run(void* h, void* p, void* u, void* v, void* w, double K, const void* K_field, double dt,
                     int64_t steps) {
    auto* om = static_cast<OMesh*>(h);
    ORACLE_DISPATCH(om,
                    cfdo::wave_run(om->f, (float*)p, (float*)u, (float*)v, (float*)w, float(K),
                                   (const float*)K_field, float(dt), steps),
                    cfdo::wave_run(om->d, (double*)p, (double*)u, (double*)v, (double*)w, K,
                                   (const double*)K_field, dt, steps));
}

void oracle_wave_run_fused(void* h, void* p, void* u, void* v, void* w, double K, const void* K_field, double dt,
                           int64_t steps) {
    auto* om = static_cast<OMesh*>(h);
    ORACLE_DISPATCH(om,
                    cfdo::wave_run_fused(om->f, (float*)p, (float*)u, (float*)v, (float*)w, float(K),
                                         (const float*)K_field, float(dt), steps),
                    cfdo::wave_run_fused(om->d, (double*)p, (double*)u, (double*)v, (double*)w, K,
                                         (const double*)K_field, dt, steps));
}

// OpenMP threads of the oracle's loops (the CPU baseline's 1-thread and
// all-thread runs); returns the previous setting.
int oracle_set_threads(int n) {
    const int prev = omp_get_max_threads();
    if (n > 0) omp_set_num_threads(n);
    return prev;
}

double oracle_cfl3d(void* h, double CFL, const void* u, const void* v, const void* w, const void* p,
                    const void* rho) {
    auto* om = static_cast<OMesh*>(h);
    if (om->dtype == 0)
        return cfdo::cfl3d(float(CFL), om->f, (const float*)u, (const float*)v, (const float*)w, (const float*)p,
                           (const float*)rho);
    return cfdo::cfl3d(CFL, om->d, (const double*)u, (const double*)v, (const double*)w, (const double*)p,
                       (const double*)rho);
}

double oracle_cfl1d(void* h, double CFL, const void* u, const void* p, const void* rho) {
    auto* om = static_cast<OMesh*>(h);
    if (om->dtype == 0)
        return cfdo::cfl1d(float(CFL), om->f, (const float*)u, (const float*)p, (const float*)rho);
    return cfdo::cfl1d(CFL, om->d, (const double*)u, (const double*)p, (const double*)rho);
}

// `steps` Rusanov steps of the Euler system; U = 5 contiguous N-arrays (rho,
// rho u, rho v, rho w, E), updated in place; dt_hist (nullable) gets each dt.
void oracle_euler_run(void* h, void* U, double gamma, double CFL, int64_t steps, void* dt_hist) {
    auto* om = static_cast<OMesh*>(h);
    if (om->dtype == 0) {
        float* b = (float*)U;
        const int64_t N = om->f.N;
        float* const S[5] = {b, b + N, b + 2 * N, b + 3 * N, b + 4 * N};
        cfdo::euler_run(om->f, S, float(gamma), float(CFL), steps, (float*)dt_hist);
    } else {
        double* b = (double*)U;
        const int64_t N = om->d.N;
        double* const S[5] = {b, b + N, b + 2 * N, b + 3 * N, b + 4 * N};
        cfdo::euler_run(om->d, S, gamma, CFL, steps, (double*)dt_hist);
    }
}

double oracle_euler_dt(void* h, const void* U, double gamma, double CFL) {
    auto* om = static_cast<OMesh*>(h);
    if (om->dtype == 0) {
        const float* b = (const float*)U;
        const int64_t N = om->f.N;
        const float* const S[5] = {b, b + N, b + 2 * N, b + 3 * N, b + 4 * N};
        return cfdo::euler_dt(om->f, S, float(gamma), float(CFL));
    }
    const double* b = (const double*)U;
    const int64_t N = om->d.N;
    const double* const S[5] = {b, b + N, b + 2 * N, b + 3 * N, b + 4 * N};
    return cfdo::euler_dt(om->d, S, gamma, CFL);
}

// D2tVar rule: 1 when step t (0-based) leapfrogs under `rule` (0 fused, 1
// per-operator with history, 2 per-operator without history).
int oracle_d2t_leapfrogs(int64_t t, int rule) { return cfdo::d2t_leapfrogs(t, rule) ? 1 : 0; }

// One step of {d2t(u), '=', k * (sum of d2d<axis>(u) over axis_mask)}; `prev`
// is read only when `leapfrog` is set.
void oracle_d2t_step(void* h, const void* u, const void* prev, double k, int axis_mask, double dt, int leapfrog,
                     void* out) {
    auto* om = static_cast<OMesh*>(h);
    ORACLE_DISPATCH(om,
                    cfdo::d2t_step(om->f, (const float*)u, (const float*)prev, float(k), axis_mask, float(dt),
                                   leapfrog != 0, (float*)out),
                    cfdo::d2t_step(om->d, (const double*)u, (const double*)prev, k, axis_mask, dt, leapfrog != 0,
                                   (double*)out));
}

}  // extern "C"
