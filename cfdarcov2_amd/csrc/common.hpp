// Internal runtime types shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cfdarco.h"

namespace cfd {

// Error carried from the implementation to the C-ABI wrapper, which turns it
// into a status code plus a thread-local message.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define CFD_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw ::cfd::Error(CFD_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define CFD_REQUIRE(cond, msg)                                   \
    do {                                                         \
        if (!(cond)) throw ::cfd::Error(CFD_E_INVALID, (msg));  \
    } while (0)

// Runs `body`, mapping exceptions to a status code (no exception crosses the ABI).
template <class F> int guarded(F&& body) {
    try {
        body();
        return CFD_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return CFD_E_NOMEM;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return CFD_E_INVALID;
    }
}

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

}  // namespace cfd

// Library context: one device, one stream-ordered memory pool.
struct cfd_ctx_s {
    int device = 0;
    hipMemPool_t pool = nullptr;
    std::mutex mu;
    int64_t limit = 0;          // 0 = uncapped
    int64_t in_use = 0;
    int64_t peak = 0;
    std::vector<std::pair<void*, int64_t>> live;  // outstanding pool blocks
    int num_cu = 0;
};

namespace cfd {

// Constant geometry of an exactly uniform hex mesh, per axis a = 0 (x), 1 (y), 2 (z).
template <class T> struct UniformGeo {
    int64_t X, Y, Z;
    T area[3];       // face area of the a-normal faces
    T alpha[3];      // alpha_d on interior a-faces
    T alpha_b[3];    // alpha_d on boundary a-faces
    T len[3];        // cell centre to a-face centre distance
    T vol;           // cell volume
    T dl;            // min(dx, dy, dz)
    T inv_vol;       // 1 / vol
    int vol_pow2;    // vol is an exact power of two: x / vol == x * inv_vol bit for bit
};

// Face-ordered SoA geometry for a general (non-uniform) structured hex mesh.
// Face families in the reference's local face order: z-faces (fc 0,1), y-faces
// (fc 2,3), x-faces (fc 4,5).  Family index of the face between cells along:
//   z: x*Y*(Z+1) + y*(Z+1) + zf       y: x*(Y+1)*Z + yf*Z + z       x: xf*Y*Z + y*Z + z
// Per face: the axis-direction normal (reference _normals), area, alpha_d.
// Per cell: volume and the six centre-to-face distances (stab only).
template <class T> struct GeneralGeo {
    int64_t X, Y, Z;
    const T* fn[3][3];   // [family z,y,x][component x,y,z]
    const T* farea[3];
    const T* falpha[3];
    const T* vol;
    const T* len;        // 6 x N, fc-major
    T dl;
};

// Finite-volume geometry of the Euler stepper: per face family the +axis unit
// normal and the area of the face's vector area, and the cell volumes built from
// the vector areas (built on first use).
template <class T> struct FvGeo {
    int64_t X, Y, Z;
    const T* n[3][3];    // [family z,y,x][component x,y,z]
    const T* A[3];
    const T* vol;
};

// Connectivity of an unstructured hex mesh (cfd_mesh_create_unstructured), in
// the caller's numbering.
struct UConn {
    int64_t N = 0, F = 0, V = 0;
    std::vector<double> verts;        // V x 3
    std::vector<int64_t> face_verts;  // F x 4
    std::vector<int64_t> face_cells;  // F x 2, -1 when absent
    std::vector<int64_t> cell_faces;  // N x 6 (local order z-, z+, y-, y+, x-, x+)
    std::vector<int64_t> cell_verts;  // N x 8
};

// Device geometry of an unstructured hex mesh, cells and faces in the mesh's
// (reordered) numbering.  Per (cell, local face) arrays are fc-major (6 x N).
template <class T> struct UnstructGeo {
    int64_t N, F;
    const int32_t* nbr;    // neighbour across the face, -1 on the boundary
    const int32_t* face;   // face id
    const int8_t* osign;   // outward normal of the face = osign * its normal
    const T* fn[3];        // per face: normal (first three vertices)
    const T* area;
    const T* alpha;        // alpha_d (the same from both cells)
    const T* vol;          // per cell
    const T* len;          // per (fc, cell), fc-major: centre-to-face distance
    // the Rusanov stepper's geometry, built on its first use (null before):
    // per face the +axis unit vector-area normal and area, per cell the volume
    const T* sn[3] = {nullptr, nullptr, nullptr};
    const T* sa = nullptr;
    const T* svol = nullptr;
};

}  // namespace cfd

struct cfd_mesh_s {
    cfd_ctx_t ctx = nullptr;
    int dtype = CFD_F32;
    int64_t X = 0, Y = 0, Z = 0, N = 0, F = 0, V = 0;
    bool uniform = false;
    double dx = 0, dy = 0, dz = 0;
    cfd::UniformGeo<float> ug32{};
    cfd::UniformGeo<double> ug64{};
    cfd::GeneralGeo<float> gg32{};
    cfd::GeneralGeo<double> gg64{};
    void* dev_block = nullptr;     // one allocation holding all device geometry
    int64_t device_bytes = 0;
    // host copies of the reference-layout tables are rebuilt on demand
    std::vector<double> vertex_offsets;
    bool has_offsets = false;
    // fused acoustic step on uniform meshes: variant 0 = per-thread kernel,
    // 1 = 2.5D marching tiles, 7-9 = row form, 11-12 = two-step row form (the
    // variants of cfd_mesh_set_wave_shape); uniform meshes whose z extent
    // suits the row form start on it
    int wave_variant = 1;
    int wave_ty = 16;
    int wave_L = 32;
    int wave_remap = 1;
    // heat step on uniform meshes: variant 0 = one thread per cell, one step
    // per launch; 1 = two steps per launch in row form (`heat_nw` waves per
    // workgroup, `heat_L` planes per chunk); 2 = up to `heat_L` steps per
    // launch in LDS boxes `heat_nw` cells wide
    int heat_variant = 0;
    int heat_nw = 8;
    int heat_L = 16;
    // cfd_heat_run's captured block of steps (a HIP graph of the launches of
    // kHeatGraphSteps steps from Ta back to Ta), the arguments it was captured
    // with, the library stream it was captured on and the stream of its last
    // launch
    hipGraphExec_t heat_exec = nullptr;
    hipStream_t capture = nullptr;
    hipStream_t heat_last = nullptr;
    const void* hg_Ta = nullptr;
    const void* hg_Tb = nullptr;
    const void* hg_af = nullptr;
    double hg_alpha = 0, hg_dt = 0;
    int hg_shape[3] = {-1, 0, 0};
    // Euler stepper: variant 0 = one thread per cell, 1 = 2.5D march with
    // `euler_ty` rows x 64 columns x `euler_L` planes per workgroup
    int euler_variant = 1;
    int euler_ty = 4;
    int euler_L = 16;
    // adjoint of the acoustic step (cfd_wave_traces / cfd_wave_adjoint):
    // form 0 = one thread per cell, 1 = row form with `adj_ty` output waves
    // and `adj_L` planes per workgroup (where the mesh allows it; constant K
    // for the transposed step)
    int adj_form = 1;
    int adj_ty = 4;
    int adj_L = 8;
    // unstructured meshes: connectivity in the caller's numbering, the cell
    // and face orders the device arrays use (perm[new] = old, iperm[old] =
    // new) and the device geometry (in dev_block)
    bool unstructured = false;
    cfd::UConn uconn;
    std::vector<int64_t> cell_perm, cell_iperm, face_perm;
    cfd::UnstructGeo<float> uu32{};
    cfd::UnstructGeo<double> uu64{};
    // Euler stepper geometry (device block allocated on first use)
    cfd::FvGeo<float> fv32{};
    cfd::FvGeo<double> fv64{};
    void* fv_block = nullptr;
    int64_t fv_bytes = 0;
};
