// C-ABI entry points: context, stream-ordered pool, hex mesh, operator and
// fused-step dispatch.  Every function maps failures to a status code plus a
// thread-local message (cfd_last_error); nothing throws across the ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "hexmesh.hpp"
#include "kernels.hpp"

namespace cfd {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

static int elem_size(int dtype) { return dtype == CFD_F64 ? 8 : 4; }
static void check_dtype(int dtype) {
    CFD_REQUIRE(dtype == CFD_F32 || dtype == CFD_F64, "dtype must be CFD_F32 or CFD_F64");
}

// ---------------------------------------------------------------------------
// stream-ordered pool (hipMemPool with a release threshold of "never", so freed
// blocks are reused by later allocations on the same device)
// ---------------------------------------------------------------------------
void* pool_alloc(cfd_ctx_t ctx, int64_t bytes, hipStream_t s) {
    CFD_REQUIRE(ctx != nullptr, "null context");
    CFD_REQUIRE(bytes >= 0, "negative allocation size");
    if (bytes == 0) bytes = 1;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (ctx->limit > 0 && ctx->in_use + bytes > ctx->limit)
            throw Error(CFD_E_NOMEM, "pool limit reached: " + std::to_string(ctx->in_use) + " + " +
                                         std::to_string(bytes) + " > " + std::to_string(ctx->limit) + " bytes");
    }
    void* p = nullptr;
    CFD_HIP(hipSetDevice(ctx->device));
    hipError_t e = hipMallocFromPoolAsync(&p, size_t(bytes), ctx->pool, s);
    if (e != hipSuccess || p == nullptr) {
        (void)hipGetLastError();
        throw Error(CFD_E_NOMEM, std::string("device pool allocation of ") + std::to_string(bytes) +
                                     " bytes failed: " + hipGetErrorString(e));
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->live.emplace_back(p, bytes);
    ctx->in_use += bytes;
    ctx->peak = std::max(ctx->peak, ctx->in_use);
    return p;
}

void pool_free(cfd_ctx_t ctx, void* p, hipStream_t s) {
    if (p == nullptr) return;
    int64_t bytes = -1;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        for (size_t i = 0; i < ctx->live.size(); ++i)
            if (ctx->live[i].first == p) {
                bytes = ctx->live[i].second;
                ctx->live[i] = ctx->live.back();
                ctx->live.pop_back();
                break;
            }
    }
    CFD_REQUIRE(bytes >= 0, "cfd_free: pointer was not allocated by this context");
    CFD_HIP(hipFreeAsync(p, s));
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->in_use -= bytes;
}

// ---------------------------------------------------------------------------
// mesh construction
// ---------------------------------------------------------------------------
template <class T> static bool all_equal(const std::vector<T>& v, T ref) {
    for (const T& x : v)
        if (!(x == ref)) return false;
    return true;
}

// Determines whether the geometry is exactly uniform and, if so, fills the
// constants; returns false otherwise.
template <class T> static bool detect_uniform(const HexGeometry<T>& g, UniformGeo<T>& u) {
    const int64_t X = g.X, Y = g.Y, Z = g.Z;
    u.X = X; u.Y = Y; u.Z = Z;
    // axis-normal of each family must be exactly +e_axis
    for (int fam = 0; fam < 3; ++fam) {
        const int axis = 2 - fam;
        const int fc = 2 * fam;  // a local face of this family (lower)
        for (int64_t f = 0; f < g.nf[fam]; ++f) {
            const P3<T> ax = HexGeometry<T>::axis_normal(g.fnorm(fam, f), fc);
            const T comp[3] = {ax.x, ax.y, ax.z};
            for (int d = 0; d < 3; ++d)
                if (!(comp[d] == (d == axis ? T(1) : T(0)))) return false;
        }
        if (!all_equal(g.farea[fam], g.farea[fam][0])) return false;
        u.area[axis] = g.farea[fam][0];
    }
    if (!all_equal(g.vol, g.vol[0])) return false;
    u.vol = g.vol[0];
    // alpha on interior and boundary faces of each family
    for (int fam = 0; fam < 3; ++fam) {
        const int axis = 2 - fam;
        const int64_t ext[3] = {X, Y, Z};
        bool have_i = false, have_b = false;
        T ai = 0, ab = 0;
        for (int64_t x = 0; x < X + (axis == 0); ++x)
            for (int64_t y = 0; y < Y + (axis == 1); ++y)
                for (int64_t z = 0; z < Z + (axis == 2); ++z) {
                    const int64_t k = axis == 0 ? x : axis == 1 ? y : z;
                    const int64_t f = fam == 0 ? g.zface(x, y, z) : fam == 1 ? g.yface(x, y, z) : g.xface(x, y, z);
                    const bool bnd = (k == 0) || (k == ext[axis]);
                    const T a = g.falpha[fam][f];
                    if (bnd) {
                        if (!have_b) { ab = a; have_b = true; }
                        else if (!(a == ab)) return false;
                    } else {
                        if (!have_i) { ai = a; have_i = true; }
                        else if (!(a == ai)) return false;
                    }
                }
        u.alpha_b[axis] = ab;
        u.alpha[axis] = have_i ? ai : ab;
    }
    // centre-to-face distance per axis (both faces of the pair)
    for (int axis = 0; axis < 3; ++axis) {
        const int fc_lo = 2 * (2 - axis), fc_hi = fc_lo + 1;
        const T ref = g.len[int64_t(fc_lo) * g.N];
        for (int64_t i = 0; i < g.N; ++i)
            if (!(g.len[int64_t(fc_lo) * g.N + i] == ref) || !(g.len[int64_t(fc_hi) * g.N + i] == ref)) return false;
        u.len[axis] = ref;
    }
    return true;
}

template <class T> static T min3(T a, T b, T c) { return std::min(std::min(a, b), c); }

template <class T>
static void upload_general(cfd_mesh_s* m, const HexGeometry<T>& g, GeneralGeo<T>& gg, hipStream_t s) {
    // one block: per family [ax(3), area, alpha] * nf, then vol (N), len (6N)
    int64_t total = 0;
    for (int fam = 0; fam < 3; ++fam) total += 5 * g.nf[fam];
    total += 7 * g.N;
    const int64_t bytes = total * int64_t(sizeof(T));
    m->dev_block = pool_alloc(m->ctx, bytes, s);
    m->device_bytes = bytes;
    std::vector<T> host(static_cast<size_t>(total));
    T* dev = (T*)m->dev_block;
    int64_t off = 0;
    gg.X = g.X; gg.Y = g.Y; gg.Z = g.Z;
    for (int fam = 0; fam < 3; ++fam) {
        const int fc = 2 * fam;
        const int64_t nf = g.nf[fam];
        for (int64_t f = 0; f < nf; ++f) {
            const P3<T> ax = HexGeometry<T>::axis_normal(g.fnorm(fam, f), fc);
            host[off + 0 * nf + f] = ax.x;
            host[off + 1 * nf + f] = ax.y;
            host[off + 2 * nf + f] = ax.z;
            host[off + 3 * nf + f] = g.farea[fam][f];
            host[off + 4 * nf + f] = g.falpha[fam][f];
        }
        for (int d = 0; d < 3; ++d) gg.fn[fam][d] = dev + off + d * nf;
        gg.farea[fam] = dev + off + 3 * nf;
        gg.falpha[fam] = dev + off + 4 * nf;
        off += 5 * nf;
    }
    std::copy(g.vol.begin(), g.vol.end(), host.begin() + off);
    gg.vol = dev + off;
    off += g.N;
    std::copy(g.len.begin(), g.len.end(), host.begin() + off);
    gg.len = dev + off;
    off += 6 * g.N;
    CFD_HIP(hipMemcpyAsync(dev, host.data(), size_t(bytes), hipMemcpyHostToDevice, s));
    CFD_HIP(hipStreamSynchronize(s));
}

template <class T> static void create_mesh(cfd_mesh_s* m, const cfd_hex_desc& d) {
    HexGeometry<T> g;
    build_hex_geometry<T>(g, d.nx, d.ny, d.nz, T(d.dx), T(d.dy), T(d.dz), d.vertex_offsets);
    UniformGeo<T> u{};
    const bool uni = !d.force_general && detect_uniform(g, u);
    const T dl = min3(T(d.dx), T(d.dy), T(d.dz));
    if (uni) {
        u.dl = dl;
        // dividing by a power of two is exact, so it may become a multiplication
        int ex = 0;
        u.vol_pow2 = (u.vol > T(0) && std::frexp(u.vol, &ex) == T(0.5)) ? 1 : 0;
        u.inv_vol = T(1) / u.vol;
        m->uniform = true;
        // two heat steps per launch on large meshes (256^3: 1.5-1.7x the
        // per-thread kernel); small ones stay on the per-thread kernel
        if (d.nx * d.ny * d.nz >= (int64_t(1) << 23) && heat_row_supported(d.nz, 12)) {
            m->heat_variant = 1;
            m->heat_nw = 12;
            m->heat_L = 16;
        }
        if (wave_row_supported(d.nz)) {  // row forms, the fastest measured at 256^3
            if constexpr (sizeof(T) == 4) {  // two steps per pass
                m->wave_variant = 11;
                m->wave_ty = 8;
                m->wave_L = 32;
                m->wave_remap = 3;
            } else {
                m->wave_variant = 7;
                m->wave_ty = 4;
                m->wave_L = 32;
                m->wave_remap = 1;
            }
        }
        if constexpr (sizeof(T) == 4) m->ug32 = u; else m->ug64 = u;
    } else {
        m->uniform = false;
        GeneralGeo<T> gg{};
        gg.dl = dl;
        upload_general(m, g, gg, nullptr);
        if constexpr (sizeof(T) == 4) m->gg32 = gg; else m->gg64 = gg;
    }
}

// Reference-layout table reconstruction (Mesh3D::compute, src/mesh3d.cpp:202-315).
template <class T>
static void hex_table(int64_t X_, int64_t Y_, int64_t Z_, double dx, double dy, double dz, const double* offsets,
                      int which, void* out) {
    HexGeometry<T> g;
    build_hex_geometry<T>(g, X_, Y_, Z_, T(dx), T(dy), T(dz), offsets);
    const int64_t N = g.N, X = g.X, Y = g.Y, Z = g.Z;
    if (which >= CFD_T_FV_NX && which <= CFD_T_FV_VOL) {
        std::vector<T> fv;
        int64_t off_fam[3], off_vol;
        build_fv_geometry<T>(g, fv, off_fam, off_vol);
        T* ot = (T*)out;
        if (which == CFD_T_FV_VOL) {
            std::memcpy(out, fv.data() + off_vol, size_t(N) * sizeof(T));
            return;
        }
        const int comp = which - CFD_T_FV_NX;  // 0..2 normal components, 3 area
        for (int64_t x = 0; x < X; ++x)
            for (int64_t y = 0; y < Y; ++y)
                for (int64_t z = 0; z < Z; ++z) {
                    const int64_t i = Z * Y * x + Z * y + z;
                    for (int fc = 0; fc < 6; ++fc) {
                        int fam; int64_t f;
                        g.cell_face(x, y, z, fc, fam, f);
                        ot[int64_t(fc) * N + i] = fv[size_t(off_fam[fam] + comp * g.nf[fam] + f)];
                    }
                }
        return;
    }
    if (which == CFD_T_VOL || which == CFD_T_CCX || which == CFD_T_CCY || which == CFD_T_CCZ) {
        const std::vector<T>& src = which == CFD_T_VOL ? g.vol : which == CFD_T_CCX ? g.ccx
                                  : which == CFD_T_CCY ? g.ccy : g.ccz;
        std::memcpy(out, src.data(), size_t(N) * sizeof(T));
        return;
    }
    int64_t* oi = (int64_t*)out;
    T* ot = (T*)out;
    for (int64_t x = 0; x < X; ++x)
        for (int64_t y = 0; y < Y; ++y)
            for (int64_t z = 0; z < Z; ++z) {
                const int64_t i = Z * Y * x + Z * y + z;
                const int64_t crd[3] = {x, y, z}, ext[3] = {X, Y, Z}, str[3] = {Y * Z, Z, 1};
                bool bnd = false;
                for (int fc = 0; fc < 6; ++fc) {
                    const int axis = 2 - (fc >> 1);
                    const bool upper = fc & 1;
                    const bool b = upper ? crd[axis] == ext[axis] - 1 : crd[axis] == 0;
                    bnd = bnd || b;
                    const int64_t nb = b ? i : (upper ? i + str[axis] : i - str[axis]);
                    int fam; int64_t f;
                    g.cell_face(x, y, z, fc, fam, f);
                    const P3<T> ax = HexGeometry<T>::axis_normal(g.fnorm(fam, f), fc);
                    const P3<T> na = (fc & 1) ? ax : P3<T>{-ax.x, -ax.y, -ax.z};
                    const int64_t k = int64_t(fc) * N + i;
                    switch (which) {
                        case CFD_T_IDS: oi[k] = nb; break;
                        case CFD_T_IDS_BF: oi[k] = b ? -1 : nb; break;
                        case CFD_T_NX: ot[k] = ax.x; break;
                        case CFD_T_NY: ot[k] = ax.y; break;
                        case CFD_T_NZ: ot[k] = ax.z; break;
                        case CFD_T_NAX: ot[k] = na.x; break;
                        case CFD_T_NAY: ot[k] = na.y; break;
                        case CFD_T_NAZ: ot[k] = na.z; break;
                        case CFD_T_AREA: ot[k] = g.farea[fam][f]; break;
                        case CFD_T_ALPHA: ot[k] = g.falpha[fam][f]; break;
                        case CFD_T_LEN: ot[k] = g.len[k]; break;
                        default: break;
                    }
                }
                if (which == CFD_T_IS_BND) ot[i] = bnd ? T(1) : T(0);
            }
    CFD_REQUIRE(which >= CFD_T_IDS && which <= CFD_T_CCZ, "unknown mesh table id");
}

static bool valid_table(int t) { return t >= CFD_T_IDS && t <= CFD_T_FV_VOL; }

// unstructured meshes (umesh.hip)
template <class T> void create_unstructured(cfd_mesh_s* m, int reorder);
template <class T>
void unstructured_table(const UConn& c, const int64_t* perm, const int64_t* iperm, int which, void* out);
template <class T> void launch_ugrad(const cfd_mesh_s* m, const void* phi, void* out, int mask, int wall, hipStream_t s);
template <class T> void launch_ulapl(const cfd_mesh_s* m, const void* phi, void* out, int mask, hipStream_t s);
template <class T>
void launch_ustab(const cfd_mesh_s* m, const void* phi, void* out, int mask, void* work, hipStream_t s);
template <class T>
void launch_uwave(const cfd_mesh_s* m, void* const a[4], void* const b[4], double K, const void* Kf, double dt,
                  int64_t nsteps, hipStream_t s);
template <class T>
void launch_uheat(const cfd_mesh_s* m, void* Ta, void* Tb, double alpha, const void* af, double dt, int64_t nsteps,
                  hipStream_t s);

static void require_lattice(const cfd_mesh_s* m, const char* what) {
    if (m->unstructured)
        throw Error(CFD_E_UNSUPPORTED, std::string(what) +
                                           ": unstructured meshes take cfd_grad, cfd_lapl, cfd_stab, cfd_heat_run, "
                                           "cfd_wave_run and cfd_euler_run only");
}

static UConn uconn_of(const cfd_umesh_desc* d) {
    CFD_REQUIRE(d != nullptr, "null unstructured mesh description");
    CFD_REQUIRE(d->n_cells > 0 && d->n_faces > 0 && d->n_vertices > 0, "unstructured mesh: empty mesh");
    CFD_REQUIRE(d->vertices && d->face_verts && d->face_cells && d->cell_faces && d->cell_verts,
                "unstructured mesh: null connectivity array");
    UConn c;
    c.N = d->n_cells; c.F = d->n_faces; c.V = d->n_vertices;
    c.verts.assign(d->vertices, d->vertices + 3 * c.V);
    c.face_verts.assign(d->face_verts, d->face_verts + 4 * c.F);
    c.face_cells.assign(d->face_cells, d->face_cells + 2 * c.F);
    c.cell_faces.assign(d->cell_faces, d->cell_faces + 6 * c.N);
    c.cell_verts.assign(d->cell_verts, d->cell_verts + 8 * c.N);
    return c;
}

// ---------------------------------------------------------------------------
// Fixed workspace of cfd_stab_noalloc: the first stage's three gradient fields
// of the two-launch stabilisation on general meshes (3 N entries of the mesh
// dtype).  Each GPU that loads the library holds its own copy until the
// process ends.  The array's name is a host-side stand-in, never passed to a
// kernel or a copy: its device address is looked up once per device, in
// cfd_init with that device current, and only that address is used.
// ---------------------------------------------------------------------------
constexpr int64_t kNoAllocBytes = int64_t(48) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_addr[kMaxDevices] = {};

// Resolves and records `scratch`'s address on `device` (the current device).
static void resolve_scratch(int device) {
    CFD_REQUIRE(device >= 0 && device < kMaxDevices, "device index beyond the workspace table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_addr[device]) return;
    void* ptr = nullptr;
    CFD_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    CFD_REQUIRE(ptr != nullptr, "hipGetSymbolAddress returned a null workspace address");
    g_scratch_addr[device] = ptr;
}

static void* scratch_addr(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    void* p = (device >= 0 && device < kMaxDevices) ? g_scratch_addr[device] : nullptr;
    CFD_REQUIRE(p != nullptr, "workspace address not resolved for this device (cfd_init not called on it)");
    return p;
}

}  // namespace cfd

using namespace cfd;

// Steps per captured heat block: even, so a block leaves the state where it
// found it (Ta), and long enough that one graph launch replaces many kernel
// launches.
constexpr int64_t kHeatGraphSteps = 32;

namespace {

void drop_heat_graph(cfd_mesh_s* m) {
    if (!m->heat_exec) return;
    // a launch of the old graph may still be running
    if (m->heat_last) CFD_HIP(hipStreamSynchronize(m->heat_last));
    (void)hipGraphExecDestroy(m->heat_exec);
    m->heat_exec = nullptr;
    m->heat_last = nullptr;
}

// Captures kHeatGraphSteps steps of launch_heat on the mesh's own stream.
void build_heat_graph(cfd_mesh_s* m, void* Ta, void* Tb, double alpha, const void* af, double dt) {
    drop_heat_graph(m);
    if (!m->capture) CFD_HIP(hipStreamCreateWithFlags(&m->capture, hipStreamNonBlocking));
    hipGraph_t graph = nullptr;
    CFD_HIP(hipStreamBeginCapture(m->capture, hipStreamCaptureModeThreadLocal));
    try {
        if (m->dtype == CFD_F32) cfd::launch_heat<float>(m, Ta, Tb, alpha, af, dt, kHeatGraphSteps, m->capture);
        else cfd::launch_heat<double>(m, Ta, Tb, alpha, af, dt, kHeatGraphSteps, m->capture);
    } catch (...) {
        (void)hipStreamEndCapture(m->capture, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    CFD_HIP(hipStreamEndCapture(m->capture, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    CFD_HIP(e);
    m->heat_exec = exec;
    m->hg_Ta = Ta;
    m->hg_Tb = Tb;
    m->hg_af = af;
    m->hg_alpha = alpha;
    m->hg_dt = dt;
    m->hg_shape[0] = m->heat_variant;
    m->hg_shape[1] = m->heat_nw;
    m->hg_shape[2] = m->heat_L;
}

bool heat_graph_matches(const cfd_mesh_s* m, const void* Ta, const void* Tb, double alpha, const void* af,
                        double dt) {
    return m->heat_exec && m->hg_Ta == Ta && m->hg_Tb == Tb && m->hg_af == af && m->hg_alpha == alpha &&
           m->hg_dt == dt && m->hg_shape[0] == m->heat_variant && m->hg_shape[1] == m->heat_nw &&
           m->hg_shape[2] == m->heat_L;
}

bool stream_capturing(hipStream_t s) {
    if (!s) return false;  // the null stream cannot be captured
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    CFD_HIP(hipStreamIsCapturing(s, &st));
    return st != hipStreamCaptureStatusNone;
}

}  // namespace

namespace {

// The kernel shapes cfd_mesh_autotune times (the forms and tiles of
// cfd_mesh_set_wave_shape / cfd_mesh_set_heat_shape; shapes a mesh does not
// allow are skipped).
std::vector<std::array<int, 4>> tune_candidates(int kind) {
    std::vector<std::array<int, 4>> c;
    if (kind == CFD_TUNE_WAVE) {
        c.push_back({0, 16, 32, 1});
        for (int ty : {4, 8, 16, 32})
            for (int L : {4, 8, 16, 32, 64})
                for (int r : {0, 1}) c.push_back({1, ty, L, r});
        for (int v : {7, 8, 9})
            for (int ty : {4, 8, 14})
                for (int L : {4, 8, 16, 32, 64})
                    for (int r : {0, 1, 2, 3}) c.push_back({v, ty, L, r});
        for (int v : {11, 12})
            for (int ty : {2, 4, 8, 10})
                for (int L : {8, 16, 32, 64})
                    for (int r : {0, 1, 2, 3}) c.push_back({v, ty, L, r});
    } else {
        c.push_back({0, 8, 16, 0});
        for (int w : {8, 12, 16})
            for (int L : {1, 2, 4, 8, 16}) c.push_back({1, w, L, 0});
        for (int B : {4, 8})
            for (int K : {2, 4}) c.push_back({2, B, K, 0});
    }
    return c;
}

}  // namespace

extern "C" {

const char* cfd_last_error(void) { return g_last_error.c_str(); }
int cfd_version(void) { return 1; }

int cfd_init(int device, cfd_ctx_t* out) {
    return guarded([&] {
        CFD_REQUIRE(out != nullptr, "cfd_init: null output");
        int n = 0;
        CFD_HIP(hipGetDeviceCount(&n));
        CFD_REQUIRE(device >= 0 && device < n, "cfd_init: device index out of range");
        CFD_HIP(hipSetDevice(device));
        auto ctx = std::make_unique<cfd_ctx_s>();
        ctx->device = device;
        hipMemPoolProps props{};
        props.allocType = hipMemAllocationTypePinned;
        props.handleTypes = hipMemHandleTypeNone;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = device;
        CFD_HIP(hipMemPoolCreate(&ctx->pool, &props));
        uint64_t thr = UINT64_MAX;
        CFD_HIP(hipMemPoolSetAttribute(ctx->pool, hipMemPoolAttrReleaseThreshold, &thr));
        hipDeviceProp_t prop;
        CFD_HIP(hipGetDeviceProperties(&prop, device));
        ctx->num_cu = prop.multiProcessorCount;
        resolve_scratch(device);
        *out = ctx.release();
    });
}

int cfd_destroy(cfd_ctx_t ctx) {
    return guarded([&] {
        if (!ctx) return;
        CFD_HIP(hipSetDevice(ctx->device));
        CFD_HIP(hipDeviceSynchronize());
        for (auto& b : ctx->live) (void)hipFreeAsync(b.first, nullptr);
        ctx->live.clear();
        CFD_HIP(hipDeviceSynchronize());
        CFD_HIP(hipMemPoolDestroy(ctx->pool));
        delete ctx;
    });
}

int cfd_alloc(cfd_ctx_t ctx, int64_t bytes, void** dptr, void* stream) {
    return guarded([&] {
        CFD_REQUIRE(dptr != nullptr, "cfd_alloc: null output");
        *dptr = pool_alloc(ctx, bytes, as_stream(stream));
    });
}

int cfd_free(cfd_ctx_t ctx, void* dptr, void* stream) {
    return guarded([&] {
        CFD_REQUIRE(ctx != nullptr, "null context");
        pool_free(ctx, dptr, as_stream(stream));
    });
}

int cfd_pool_set_limit(cfd_ctx_t ctx, int64_t bytes) {
    return guarded([&] {
        CFD_REQUIRE(ctx != nullptr, "null context");
        CFD_REQUIRE(bytes >= 0, "negative pool limit");
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->limit = bytes;
    });
}

int cfd_pool_stats(cfd_ctx_t ctx, int64_t* in_use, int64_t* peak) {
    return guarded([&] {
        CFD_REQUIRE(ctx != nullptr, "null context");
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (in_use) *in_use = ctx->in_use;
        if (peak) *peak = ctx->peak;
    });
}

int cfd_memcpy(cfd_ctx_t ctx, void* dst, const void* src, int64_t bytes, int kind, void* stream) {
    return guarded([&] {
        CFD_REQUIRE(ctx != nullptr, "null context");
        CFD_REQUIRE(bytes >= 0, "negative copy size");
        if (bytes == 0) return;
        CFD_REQUIRE(dst != nullptr && src != nullptr, "cfd_memcpy: null pointer");
        hipMemcpyKind k = kind == CFD_H2D ? hipMemcpyHostToDevice
                        : kind == CFD_D2H ? hipMemcpyDeviceToHost
                        : kind == CFD_D2D ? hipMemcpyDeviceToDevice : hipMemcpyDefault;
        CFD_REQUIRE(kind == CFD_H2D || kind == CFD_D2H || kind == CFD_D2D, "cfd_memcpy: bad kind");
        CFD_HIP(hipMemcpyAsync(dst, src, size_t(bytes), k, as_stream(stream)));
        if (kind != CFD_D2D) CFD_HIP(hipStreamSynchronize(as_stream(stream)));
    });
}

int cfd_stream_sync(void* stream) {
    return guarded([&] { CFD_HIP(hipStreamSynchronize(as_stream(stream))); });
}

int cfd_mesh_create_hex(cfd_ctx_t ctx, const cfd_hex_desc* desc, cfd_mesh_t* out) {
    return guarded([&] {
        CFD_REQUIRE(ctx != nullptr && desc != nullptr && out != nullptr, "cfd_mesh_create_hex: null argument");
        check_dtype(desc->dtype);
        CFD_REQUIRE(desc->nx > 0 && desc->ny > 0 && desc->nz > 0, "mesh extents must be positive");
        CFD_REQUIRE(desc->ny <= 65535 * 4 && desc->nx <= 65535, "mesh extent exceeds launch grid limits");
        CFD_REQUIRE(desc->dx > 0 && desc->dy > 0 && desc->dz > 0, "mesh spacing must be positive");
        CFD_HIP(hipSetDevice(ctx->device));
        auto m = std::make_unique<cfd_mesh_s>();
        m->ctx = ctx;
        m->dtype = desc->dtype;
        m->X = desc->nx; m->Y = desc->ny; m->Z = desc->nz;
        m->N = m->X * m->Y * m->Z;
        m->V = (m->X + 1) * (m->Y + 1) * (m->Z + 1);
        m->F = m->X * m->Y * (m->Z + 1) + m->Z * m->Y * (m->X + 1) + m->X * m->Z * (m->Y + 1);
        m->dx = desc->dx; m->dy = desc->dy; m->dz = desc->dz;
        if (desc->vertex_offsets) {
            m->has_offsets = true;
            m->vertex_offsets.assign(desc->vertex_offsets, desc->vertex_offsets + 3 * m->V);
        }
        if (m->dtype == CFD_F32) create_mesh<float>(m.get(), *desc);
        else create_mesh<double>(m.get(), *desc);
        *out = m.release();
    });
}

int cfd_mesh_destroy(cfd_mesh_t mesh) {
    return guarded([&] {
        if (!mesh) return;
        drop_heat_graph(mesh);
        if (mesh->capture) (void)hipStreamDestroy(mesh->capture);
        if (mesh->dev_block) pool_free(mesh->ctx, mesh->dev_block, nullptr);
        if (mesh->fv_block) pool_free(mesh->ctx, mesh->fv_block, nullptr);
        if (mesh->dev_block || mesh->fv_block) CFD_HIP(hipStreamSynchronize(nullptr));
        delete mesh;
    });
}

int cfd_mesh_get_info(cfd_mesh_t m, cfd_mesh_info* info) {
    return guarded([&] {
        CFD_REQUIRE(m != nullptr && info != nullptr, "null argument");
        info->n_cells = m->N; info->n_faces = m->F; info->n_vertices = m->V;
        info->nx = m->X; info->ny = m->Y; info->nz = m->Z;
        info->dtype = m->dtype;
        info->uniform = m->uniform ? 1 : 0;
        info->device_bytes = m->device_bytes;
    });
}

int cfd_mesh_table(cfd_mesh_t m, int table, void* host_out) {
    return guarded([&] {
        if (m && m->unstructured) {
            CFD_REQUIRE(host_out != nullptr, "cfd_mesh_table: null output");
            CFD_REQUIRE(table >= CFD_T_IDS && table <= CFD_T_CCZ,
                        "unstructured meshes have the reference tables 0 to 15");
            if (m->dtype == CFD_F32)
                unstructured_table<float>(m->uconn, m->cell_perm.data(), m->cell_iperm.data(), table, host_out);
            else
                unstructured_table<double>(m->uconn, m->cell_perm.data(), m->cell_iperm.data(), table, host_out);
            return;
        }
        CFD_REQUIRE(m != nullptr && host_out != nullptr, "null argument");
        CFD_REQUIRE(valid_table(table), "unknown mesh table id");
        const double* off = m->has_offsets ? m->vertex_offsets.data() : nullptr;
        if (m->dtype == CFD_F32) hex_table<float>(m->X, m->Y, m->Z, m->dx, m->dy, m->dz, off, table, host_out);
        else hex_table<double>(m->X, m->Y, m->Z, m->dx, m->dy, m->dz, off, table, host_out);
    });
}

int cfd_hex_table(const cfd_hex_desc* d, int table, void* host_out) {
    return guarded([&] {
        CFD_REQUIRE(d != nullptr && host_out != nullptr, "null argument");
        check_dtype(d->dtype);
        CFD_REQUIRE(d->nx > 0 && d->ny > 0 && d->nz > 0, "mesh extents must be positive");
        CFD_REQUIRE(valid_table(table), "unknown mesh table id");
        if (d->dtype == CFD_F32)
            hex_table<float>(d->nx, d->ny, d->nz, d->dx, d->dy, d->dz, d->vertex_offsets, table, host_out);
        else
            hex_table<double>(d->nx, d->ny, d->nz, d->dx, d->dy, d->dz, d->vertex_offsets, table, host_out);
    });
}

#define DISPATCH(m, F, ...) ((m)->dtype == CFD_F32 ? F<float>(__VA_ARGS__) : F<double>(__VA_ARGS__))

int cfd_mesh_create_unstructured(cfd_ctx_t ctx, const cfd_umesh_desc* desc, cfd_mesh_t* out) {
    return guarded([&] {
        CFD_REQUIRE(ctx != nullptr && desc != nullptr && out != nullptr,
                    "cfd_mesh_create_unstructured: null argument");
        check_dtype(desc->dtype);
        CFD_HIP(hipSetDevice(ctx->device));
        auto m = std::make_unique<cfd_mesh_s>();
        m->ctx = ctx;
        m->dtype = desc->dtype;
        m->unstructured = true;
        m->uniform = false;
        m->uconn = uconn_of(desc);
        m->N = desc->n_cells; m->F = desc->n_faces; m->V = desc->n_vertices;
        if (m->dtype == CFD_F32) create_unstructured<float>(m.get(), desc->reorder);
        else create_unstructured<double>(m.get(), desc->reorder);
        *out = m.release();
    });
}

int cfd_mesh_permutation(cfd_mesh_t m, int64_t* cell_perm, int64_t* face_perm) {
    return guarded([&] {
        CFD_REQUIRE(m != nullptr, "null mesh");
        if (!m->unstructured) {  // lattice meshes keep the reference numbering
            for (int64_t i = 0; cell_perm && i < m->N; ++i) cell_perm[i] = i;
            for (int64_t f = 0; face_perm && f < m->F; ++f) face_perm[f] = f;
            return;
        }
        if (cell_perm) std::copy(m->cell_perm.begin(), m->cell_perm.end(), cell_perm);
        if (face_perm) std::copy(m->face_perm.begin(), m->face_perm.end(), face_perm);
    });
}

int cfd_unstructured_table(const cfd_umesh_desc* d, int table, void* host_out) {
    return guarded([&] {
        CFD_REQUIRE(host_out != nullptr, "cfd_unstructured_table: null output");
        CFD_REQUIRE(table >= CFD_T_IDS && table <= CFD_T_CCZ, "unstructured meshes have the reference tables 0 to 15");
        check_dtype(d ? d->dtype : -1);
        const UConn c = uconn_of(d);
        if (d->dtype == CFD_F32) unstructured_table<float>(c, nullptr, nullptr, table, host_out);
        else unstructured_table<double>(c, nullptr, nullptr, table, host_out);
    });
}

int cfd_grad(cfd_mesh_t m, const void* phi, void* out, int mask, int wall, void* stream) {
    return guarded([&] {
        CFD_REQUIRE(m && phi && out, "cfd_grad: null argument");
        CFD_REQUIRE(mask >= 0 && mask <= 7 && wall >= -1 && wall <= 2, "cfd_grad: bad axis_mask or wall_axis");
        if (m->unstructured) DISPATCH(m, launch_ugrad, m, phi, out, mask, wall, as_stream(stream));
        else DISPATCH(m, launch_grad, m, phi, out, mask, wall, as_stream(stream));
    });
}

int cfd_lapl(cfd_mesh_t m, const void* phi, void* out, int mask, void* stream) {
    return guarded([&] {
        CFD_REQUIRE(m && phi && out, "cfd_lapl: null argument");
        CFD_REQUIRE(mask >= 0 && mask <= 7, "cfd_lapl: bad axis_mask");
        if (m->unstructured) DISPATCH(m, launch_ulapl, m, phi, out, mask, as_stream(stream));
        else DISPATCH(m, launch_lapl, m, phi, out, mask, as_stream(stream));
    });
}

int cfd_stab(cfd_mesh_t m, const void* phi, void* out, int mask, void* stream) {
    return guarded([&] {
        CFD_REQUIRE(m && phi && out, "cfd_stab: null argument");
        CFD_REQUIRE(mask >= 0 && mask <= 7, "cfd_stab: bad axis_mask");
        hipStream_t s = as_stream(stream);
        // the first stage's gradients go through 3N pool entries on general
        // and unstructured meshes; uniform meshes recompute them in one pass
        void* work = m->uniform && !m->unstructured ? nullptr : pool_alloc(m->ctx, 3 * m->N * elem_size(m->dtype), s);
        try {
            if (m->unstructured) DISPATCH(m, launch_ustab, m, phi, out, mask, work, s);
            else DISPATCH(m, launch_stab, m, phi, out, mask, work, s);
        } catch (...) {
            if (work) pool_free(m->ctx, work, s);
            throw;
        }
        if (work) pool_free(m->ctx, work, s);
    });
}

int cfd_stab_noalloc(cfd_mesh_t m, const void* phi, void* out, int mask, void* stream) {
    return guarded([&] {
        CFD_REQUIRE(m && phi && out, "cfd_stab_noalloc: null argument");
        require_lattice(m, "cfd_stab_noalloc");
        CFD_REQUIRE(mask >= 0 && mask <= 7, "cfd_stab_noalloc: bad axis_mask");
        void* work = nullptr;
        if (!m->uniform) {
            // the kernels write all 3 N entries before reading any, so the
            // previous call's contents need no clearing
            const int64_t need = 3 * m->N * elem_size(m->dtype);
            if (need > kNoAllocBytes)
                throw Error(CFD_E_INVALID, "cfd_stab_noalloc: a general mesh of " + std::to_string(m->N) +
                                               " cells needs " + std::to_string(need) +
                                               " workspace bytes, beyond the fixed " +
                                               std::to_string(kNoAllocBytes) +
                                               " (at most 2097152 FP64 or 4194304 FP32 cells)");
            work = scratch_addr(m->ctx->device);
        }
        DISPATCH(m, launch_stab, m, phi, out, mask, work, as_stream(stream));
    });
}

int cfd_ewise(cfd_ctx_t ctx, int op, int dtype, const void* a, const void* b, const void* aux, double s,
              void* out, int64_t n, void* stream) {
    return guarded([&] {
        CFD_REQUIRE(ctx != nullptr && out != nullptr, "cfd_ewise: null argument");
        check_dtype(dtype);
        CFD_REQUIRE(op >= CFD_EW_ADD && op <= CFD_EW_ABS, "cfd_ewise: unknown op");
        CFD_REQUIRE(n >= 0, "cfd_ewise: negative length");
        const bool needs_a = op != CFD_EW_FILL;
        const bool needs_b = op == CFD_EW_ADD || op == CFD_EW_SUB || op == CFD_EW_MUL || op == CFD_EW_DIV ||
                             op == CFD_EW_AXPY || op == CFD_EW_LEAP;
        CFD_REQUIRE(!needs_a || a != nullptr, "cfd_ewise: operand a is null");
        CFD_REQUIRE(!needs_b || b != nullptr, "cfd_ewise: operand b is null");
        CFD_REQUIRE(op != CFD_EW_LEAP || aux != nullptr, "cfd_ewise: LEAP needs aux");
        if (dtype == CFD_F32) launch_ewise<float>(op, a, b, aux, s, out, n, as_stream(stream));
        else launch_ewise<double>(op, a, b, aux, s, out, n, as_stream(stream));
    });
}

int cfd_cfl_min(cfd_ctx_t ctx, int dtype, const void* u, const void* v, const void* w, const void* p,
                const void* rho, int64_t n, double dl, double cfl, double gamma, double* dt_out, void* stream) {
    return guarded([&] {
        CFD_REQUIRE(ctx && u && p && rho && dt_out, "cfd_cfl_min: null argument");
        check_dtype(dtype);
        CFD_REQUIRE(n > 0, "cfd_cfl_min: empty field");
        hipStream_t s = as_stream(stream);
        const int nb = int(std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), 2048));
        const int es = elem_size(dtype);
        void* work = pool_alloc(ctx, int64_t(nb + 1) * es, s);
        char* base = (char*)work;
        double host = 0;
        try {
            if (dtype == CFD_F32) {
                launch_cfl<float>(u, v, w, p, rho, float(dl), float(gamma), n, base, nb, base + int64_t(nb) * es, s);
                float r = 0;
                CFD_HIP(hipMemcpyAsync(&r, base + int64_t(nb) * es, 4, hipMemcpyDeviceToHost, s));
                CFD_HIP(hipStreamSynchronize(s));
                host = double(float(cfl) * r);
            } else {
                launch_cfl<double>(u, v, w, p, rho, dl, gamma, n, base, nb, base + int64_t(nb) * es, s);
                double r = 0;
                CFD_HIP(hipMemcpyAsync(&r, base + int64_t(nb) * es, 8, hipMemcpyDeviceToHost, s));
                CFD_HIP(hipStreamSynchronize(s));
                host = cfl * r;
            }
        } catch (...) {
            pool_free(ctx, work, s);
            throw;
        }
        pool_free(ctx, work, s);
        *dt_out = host;
    });
}

int cfd_mesh_geometry(cfd_mesh_t m, cfd_geometry* out) {
    return guarded([&] {
        CFD_REQUIRE(m && out, "cfd_mesh_geometry: null argument");
        require_lattice(m, "cfd_mesh_geometry");
        cfd_geometry g{};
        g.uniform = m->uniform ? 1 : 0;
        g.dtype = m->dtype;
        g.nx = m->X; g.ny = m->Y; g.nz = m->Z;
        auto fill = [&](const auto& ug, const auto& gg) {
            for (int a = 0; a < 3; ++a) {
                g.area[a] = double(ug.area[a]);
                g.alpha[a] = double(ug.alpha[a]);
                g.alpha_b[a] = double(ug.alpha_b[a]);
                g.len[a] = double(ug.len[a]);
            }
            g.vol = double(ug.vol);
            for (int f = 0; f < 3; ++f) {
                for (int d = 0; d < 3; ++d) g.face_normal[f][d] = gg.fn[f][d];
                g.face_area[f] = gg.farea[f];
                g.face_alpha[f] = gg.falpha[f];
            }
            g.cell_vol = gg.vol;
            g.cell_len = gg.len;
        };
        if (m->dtype == CFD_F32) fill(m->ug32, m->gg32);
        else fill(m->ug64, m->gg64);
        *out = g;
    });
}

int cfd_count_nonfinite(cfd_ctx_t ctx, int dtype, const void* x, int64_t n, int64_t* bad, void* stream) {
    return guarded([&] {
        CFD_REQUIRE(ctx && x && bad, "cfd_count_nonfinite: null argument");
        check_dtype(dtype);
        hipStream_t s = as_stream(stream);
        void* work = pool_alloc(ctx, 1025 * 8, s);
        unsigned long long total = 0;
        try {
            if (dtype == CFD_F32) launch_nonfinite<float>(x, n, (unsigned long long*)work, s);
            else launch_nonfinite<double>(x, n, (unsigned long long*)work, s);
            CFD_HIP(hipMemcpyAsync(&total, (char*)work + 1024 * 8, 8, hipMemcpyDeviceToHost, s));
            CFD_HIP(hipStreamSynchronize(s));
        } catch (...) {
            pool_free(ctx, work, s);
            throw;
        }
        pool_free(ctx, work, s);
        *bad = int64_t(total);
    });
}


int cfd_heat_run(cfd_mesh_t m, void* Ta, void* Tb, double alpha, const void* af, double dt, int64_t nsteps,
                 void* stream) {
    return guarded([&] {
        CFD_REQUIRE(m && Ta && Tb, "cfd_heat_run: null argument");
        CFD_REQUIRE(Ta != Tb, "cfd_heat_run: ping-pong buffers must differ");
        CFD_REQUIRE(nsteps >= 0, "cfd_heat_run: negative step count");
        hipStream_t s = as_stream(stream);
        if (m->unstructured) {
            DISPATCH(m, launch_uheat, m, Ta, Tb, alpha, af, dt, nsteps, s);
            return;
        }
        // Runs of at least one block replay a captured block of steps: one
        // graph launch instead of one kernel launch per step, which is most
        // of a step's time on small meshes.  A caller that is itself
        // capturing the stream gets the plain launches in its graph;
        // CFDARCO_HEAT_GRAPH=0 turns the blocks off.
        const char* env = std::getenv("CFDARCO_HEAT_GRAPH");
        const bool blocks = nsteps >= kHeatGraphSteps && !(env && env[0] == '0') && !stream_capturing(s);
        if (!blocks) {
            DISPATCH(m, launch_heat, m, Ta, Tb, alpha, af, dt, nsteps, s);
            return;
        }
        if (!heat_graph_matches(m, Ta, Tb, alpha, af, dt)) build_heat_graph(m, Ta, Tb, alpha, af, dt);
        const int64_t reps = nsteps / kHeatGraphSteps;
        for (int64_t r = 0; r < reps; ++r) CFD_HIP(hipGraphLaunch(m->heat_exec, s));
        m->heat_last = s;
        // the blocks leave the state in Ta; the remaining steps start there
        const int64_t rest = nsteps - reps * kHeatGraphSteps;
        if (rest) DISPATCH(m, launch_heat, m, Ta, Tb, alpha, af, dt, rest, s);
    });
}

int cfd_wave_run(cfd_mesh_t m, void* const a[4], void* const b[4], double K, const void* Kf, double dt,
                 int64_t nsteps, void* stream) {
    return guarded([&] {
        CFD_REQUIRE(m && a && b, "cfd_wave_run: null argument");
        for (int i = 0; i < 4; ++i) {
            CFD_REQUIRE(a[i] && b[i], "cfd_wave_run: null state pointer");
            for (int j = 0; j < 4; ++j) CFD_REQUIRE(a[i] != b[j], "cfd_wave_run: state_a and state_b overlap");
        }
        CFD_REQUIRE(nsteps >= 0, "cfd_wave_run: negative step count");
        if (m->unstructured) DISPATCH(m, launch_uwave, m, a, b, K, Kf, dt, nsteps, as_stream(stream));
        else DISPATCH(m, launch_wave, m, a, b, K, Kf, dt, nsteps, as_stream(stream));
    });
}

int cfd_euler_run(cfd_mesh_t m, void* const a[5], void* const b[5], double gamma, double cfl, int64_t nsteps,
                  void* dt_hist, void* stream) {
    return guarded([&] {
        CFD_REQUIRE(m && a && b, "cfd_euler_run: null argument");
        for (int i = 0; i < 5; ++i) {
            CFD_REQUIRE(a[i] && b[i], "cfd_euler_run: null state pointer");
            for (int j = 0; j < 5; ++j) CFD_REQUIRE(a[i] != b[j], "cfd_euler_run: state_a and state_b overlap");
        }
        CFD_REQUIRE(gamma > 1.0 && std::isfinite(gamma), "cfd_euler_run: gamma must be finite and > 1");
        CFD_REQUIRE(cfl > 0.0 && std::isfinite(cfl), "cfd_euler_run: cfl must be finite and positive");
        CFD_REQUIRE(nsteps >= 0, "cfd_euler_run: negative step count");
        if (m->unstructured) DISPATCH(m, launch_ueuler, m, a, b, gamma, cfl, nsteps, dt_hist, as_stream(stream));
        else DISPATCH(m, launch_euler, m, a, b, gamma, cfl, nsteps, dt_hist, as_stream(stream));
    });
}

static void check_state4(void* const a[4], void* const b[4], const char* what) {
    CFD_REQUIRE(a && b, std::string(what) + ": null argument");
    for (int i = 0; i < 4; ++i) {
        CFD_REQUIRE(a[i] && b[i], std::string(what) + ": null state pointer");
        for (int j = 0; j < 4; ++j) CFD_REQUIRE(a[i] != b[j], std::string(what) + ": state_a and state_b overlap");
    }
}

int cfd_wave_traces(cfd_mesh_t m, void* const a[4], void* const b[4], double K, const void* Kf, double dt,
                    int64_t nsteps, const int64_t* sensor_cells, int nsensors, void* traces, void* checkpoints,
                    int64_t every, void* stream) {
    return guarded([&] {
        CFD_REQUIRE(m != nullptr, "cfd_wave_traces: null mesh");
        require_lattice(m, "cfd_wave_traces");
        check_state4(a, b, "cfd_wave_traces");
        if (!m->uniform) throw Error(CFD_E_UNSUPPORTED, "cfd_wave_traces: uniform meshes only");
        CFD_REQUIRE(nsteps >= 0, "cfd_wave_traces: negative step count");
        CFD_REQUIRE(!checkpoints || every >= 1, "cfd_wave_traces: checkpoint interval must be >= 1");
        const SensorList sens = make_sensors(m, sensor_cells, nsensors);
        DISPATCH(m, run_wave_traces, m, a, b, K, Kf, dt, nsteps, sens, traces, checkpoints, every,
                 as_stream(stream));
    });
}

int cfd_wave_adjoint(cfd_mesh_t m, void* const adj_a[4], void* const adj_b[4], double K, const void* Kf, double dt,
                     int64_t nsteps, const int64_t* sensor_cells, int nsensors, const void* trace_grad,
                     const void* checkpoints, int64_t every, void* grad_K, void* stream) {
    return guarded([&] {
        CFD_REQUIRE(m != nullptr, "cfd_wave_adjoint: null mesh");
        require_lattice(m, "cfd_wave_adjoint");
        check_state4(adj_a, adj_b, "cfd_wave_adjoint");
        if (!m->uniform) throw Error(CFD_E_UNSUPPORTED, "cfd_wave_adjoint: uniform meshes only");
        CFD_REQUIRE(nsteps >= 0, "cfd_wave_adjoint: negative step count");
        CFD_REQUIRE(!grad_K || (checkpoints && every >= 1),
                    "cfd_wave_adjoint: grad_K needs the forward checkpoints and their interval");
        const SensorList sens = make_sensors(m, sensor_cells, nsensors);
        DISPATCH(m, run_wave_adjoint, m, adj_a, adj_b, K, Kf, dt, nsteps, sens, trace_grad, checkpoints, every,
                 grad_K, as_stream(stream));
    });
}

int cfd_heat_traces(cfd_mesh_t m, void* Ta, void* Tb, double alpha, const void* af, double dt, int64_t nsteps,
                    const int64_t* sensor_cells, int nsensors, void* traces, void* checkpoints, int64_t every,
                    void* stream) {
    return guarded([&] {
        CFD_REQUIRE(m && Ta && Tb, "cfd_heat_traces: null argument");
        require_lattice(m, "cfd_heat_traces");
        CFD_REQUIRE(Ta != Tb, "cfd_heat_traces: ping-pong buffers must differ");
        if (!m->uniform) throw Error(CFD_E_UNSUPPORTED, "cfd_heat_traces: uniform meshes only");
        CFD_REQUIRE(nsteps >= 0, "cfd_heat_traces: negative step count");
        CFD_REQUIRE(!checkpoints || every >= 1, "cfd_heat_traces: checkpoint interval must be >= 1");
        const SensorList sens = make_sensors(m, sensor_cells, nsensors);
        DISPATCH(m, run_heat_traces, m, Ta, Tb, alpha, af, dt, nsteps, sens, traces, checkpoints, every,
                 as_stream(stream));
    });
}

int cfd_heat_adjoint(cfd_mesh_t m, void* adj_a, void* adj_b, double alpha, const void* af, double dt,
                     int64_t nsteps, const int64_t* sensor_cells, int nsensors, const void* trace_grad,
                     const void* checkpoints, int64_t every, void* grad_alpha, void* stream) {
    return guarded([&] {
        CFD_REQUIRE(m && adj_a && adj_b, "cfd_heat_adjoint: null argument");
        require_lattice(m, "cfd_heat_adjoint");
        CFD_REQUIRE(adj_a != adj_b, "cfd_heat_adjoint: ping-pong buffers must differ");
        if (!m->uniform) throw Error(CFD_E_UNSUPPORTED, "cfd_heat_adjoint: uniform meshes only");
        CFD_REQUIRE(nsteps >= 0, "cfd_heat_adjoint: negative step count");
        CFD_REQUIRE(!grad_alpha || (checkpoints && every >= 1),
                    "cfd_heat_adjoint: grad_alpha needs the forward checkpoints and their interval");
        const SensorList sens = make_sensors(m, sensor_cells, nsensors);
        DISPATCH(m, run_heat_adjoint, m, adj_a, adj_b, alpha, af, dt, nsteps, sens, trace_grad, checkpoints, every,
                 grad_alpha, as_stream(stream));
    });
}

int cfd_mesh_set_euler_shape(cfd_mesh_t m, int variant, int tile_rows, int chunk) {
    return guarded([&] {
        CFD_REQUIRE(m != nullptr, "null mesh");
        require_lattice(m, "cfd_mesh_set_euler_shape");
        CFD_REQUIRE(variant == 0 || variant == 1, "euler variant must be 0 or 1");
        CFD_REQUIRE(tile_rows == 2 || tile_rows == 4 || (tile_rows == 8 && m->dtype == CFD_F32),
                    "euler tile_rows must be 2, 4 or 8 (2 or 4 in FP64: the FP64 8-row build spills registers)");
        CFD_REQUIRE(chunk >= 1 && chunk <= 65535, "euler chunk must be in [1, 65535]");
        m->euler_variant = variant;
        m->euler_ty = tile_rows;
        m->euler_L = chunk;
    });
}

int cfd_mesh_set_heat_shape(cfd_mesh_t m, int variant, int waves, int chunk) {
    return guarded([&] {
        CFD_REQUIRE(m != nullptr, "null mesh");
        require_lattice(m, "cfd_mesh_set_heat_shape");
        CFD_REQUIRE(variant >= 0 && variant <= 2, "heat variant must be 0, 1 or 2");
        if (variant == 2) {
            CFD_REQUIRE(heat_multi_supported(waves, chunk),
                        "multi-step heat boxes: 4 or 8 cells wide (waves), 2 or 4 steps per launch (chunk)");
            CFD_REQUIRE(m->uniform, "the multi-step heat form needs an exactly uniform mesh");
            m->heat_variant = variant;
            m->heat_nw = waves;
            m->heat_L = chunk;
            return;
        }
        CFD_REQUIRE(waves == 8 || waves == 12 || waves == 16, "heat waves must be 8, 12 or 16");
        CFD_REQUIRE(chunk >= 1 && chunk <= 65535, "heat chunk must be in [1, 65535]");
        if (variant == 1 && !(m->uniform && heat_row_supported(m->Z, waves)))
            throw Error(CFD_E_UNSUPPORTED,
                        "the two-step heat form needs an exactly uniform mesh whose z extent is a multiple of 4 "
                        "and at most 256");
        m->heat_variant = variant;
        m->heat_nw = waves;
        m->heat_L = chunk;
    });
}


int cfd_mesh_autotune(cfd_mesh_t m, int kind, int steps, int* shape_out, double* us_per_step_out, void* stream) {
    return guarded([&] {
        CFD_REQUIRE(m != nullptr && shape_out != nullptr, "cfd_mesh_autotune: null argument");
        require_lattice(m, "cfd_mesh_autotune");
        CFD_REQUIRE(kind == CFD_TUNE_WAVE || kind == CFD_TUNE_HEAT, "cfd_mesh_autotune: unknown kind");
        CFD_REQUIRE(steps >= 2 && steps <= 100000, "cfd_mesh_autotune: steps must be in [2, 100000]");
        CFD_HIP(hipSetDevice(m->ctx->device));
        hipStream_t s = as_stream(stream);
        const size_t field = size_t(m->N) * (m->dtype == CFD_F32 ? sizeof(float) : sizeof(double));
        const int nf = kind == CFD_TUNE_WAVE ? 8 : 2;
        // scratch fields (zero: the kernels do the same work for any data)
        char* buf = (char*)pool_alloc(m->ctx, int64_t(field) * nf, s);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto release = [&] {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            pool_free(m->ctx, buf, s);
        };
        try {
            CFD_HIP(hipMemsetAsync(buf, 0, field * nf, s));
            CFD_HIP(hipEventCreate(&e0));
            CFD_HIP(hipEventCreate(&e1));
            void* a[4];
            void* b[4];
            for (int f = 0; f < 4; ++f) {
                a[f] = buf + size_t(f) * field;
                b[f] = buf + size_t(4 + f) * field;
            }
            auto run = [&](int64_t n) {
                if (kind == CFD_TUNE_WAVE) DISPATCH(m, launch_wave, m, a, b, -1.0, nullptr, 1e-3, n, s);
                else DISPATCH(m, launch_heat, m, a[0], a[1], 1.0, nullptr, 1e-3, n, s);
            };
            std::array<int, 4> best{};
            float best_ms = std::numeric_limits<float>::infinity();
            for (const auto& c : tune_candidates(kind)) {
                const int rc = kind == CFD_TUNE_WAVE ? cfd_mesh_set_wave_shape(m, c[0], c[1], c[2], c[3])
                                                     : cfd_mesh_set_heat_shape(m, c[0], c[1], c[2]);
                if (rc != CFD_OK) continue;  // a shape this mesh does not allow
                run(2);
                float ms = std::numeric_limits<float>::infinity();
                for (int rep = 0; rep < 2; ++rep) {
                    CFD_HIP(hipEventRecord(e0, s));
                    run(steps);
                    CFD_HIP(hipEventRecord(e1, s));
                    CFD_HIP(hipEventSynchronize(e1));
                    float t = 0;
                    CFD_HIP(hipEventElapsedTime(&t, e0, e1));
                    ms = std::min(ms, t / float(steps));
                }
                if (ms < best_ms) {
                    best_ms = ms;
                    best = c;
                }
            }
            CFD_REQUIRE(best_ms < std::numeric_limits<float>::infinity(), "cfd_mesh_autotune: no shape runs");
            const int rc = kind == CFD_TUNE_WAVE ? cfd_mesh_set_wave_shape(m, best[0], best[1], best[2], best[3])
                                                 : cfd_mesh_set_heat_shape(m, best[0], best[1], best[2]);
            CFD_REQUIRE(rc == CFD_OK, "cfd_mesh_autotune: the chosen shape was refused");
            for (int k = 0; k < 4; ++k) shape_out[k] = best[k];
            if (us_per_step_out) *us_per_step_out = double(best_ms) * 1e3;
        } catch (...) {
            release();
            throw;
        }
        release();
    });
}

int cfd_mesh_set_adjoint_shape(cfd_mesh_t m, int form, int tile_rows, int chunk) {
    return guarded([&] {
        CFD_REQUIRE(m != nullptr, "null mesh");
        require_lattice(m, "cfd_mesh_set_adjoint_shape");
        CFD_REQUIRE(form == 0 || form == 1, "adjoint form must be 0 or 1");
        CFD_REQUIRE(chunk >= 1 && chunk <= 65535, "adjoint chunk must be in [1, 65535]");
        if (form == 1 && !(m->uniform && wave_row_supported(m->Z)))
            throw Error(CFD_E_UNSUPPORTED,
                        "the adjoint row form needs an exactly uniform mesh whose z extent is a multiple of 4 and "
                        "at most 256");
        if (m->dtype == CFD_F32) {
            CFD_REQUIRE(tile_rows == 4 || tile_rows == 8, "adjoint row tiles have 4 or 8 output waves");
        } else {
            // the FP64 8-row builds spill registers (see launch_wave_row_aux)
            CFD_REQUIRE(tile_rows == 4, "FP64 adjoint row tiles have 4 output waves");
        }
        m->adj_form = form;
        m->adj_ty = tile_rows;
        m->adj_L = chunk;
    });
}

int cfd_mesh_set_wave_shape(cfd_mesh_t m, int variant, int tile_rows, int chunk, int xcd_remap) {
    return guarded([&] {
        CFD_REQUIRE(m != nullptr, "null mesh");
        require_lattice(m, "cfd_mesh_set_wave_shape");
        CFD_REQUIRE(variant == 0 || variant == 1 || (variant >= 7 && variant <= 9) || variant == 11 || variant == 12,
                    "wave variant must be 0, 1, 7, 8, 9, 11 or 12");
        const bool row_form = variant >= 7;
        if (variant >= 11) {
            if (m->dtype == CFD_F32)
                CFD_REQUIRE(tile_rows == 4 || tile_rows == 8 || tile_rows == 10,
                            "two-step row-form (11, 12) tile_rows must be 4, 8 or 10 in FP32");
            else
                CFD_REQUIRE(tile_rows == 2, "two-step row-form (11, 12) tile_rows must be 2 in FP64");
            if (!wave_row_supported(m->Z))
                throw Error(CFD_E_UNSUPPORTED, "the row form needs a z extent that is a multiple of 4 and at most 256");
        } else if (row_form) {
            CFD_REQUIRE(tile_rows == 4 || tile_rows == 8 || (tile_rows == 14 && m->dtype == CFD_F32),
                        "row-form (7 to 9) tile_rows must be 4 or 8 (or 14 in FP32)");
            if (!wave_row_supported(m->Z))
                throw Error(CFD_E_UNSUPPORTED, "the row form (variants 7 to 9) needs a z extent that is a multiple of 4 and at most 256");
        } else {
            CFD_REQUIRE(tile_rows == 4 || tile_rows == 8 || tile_rows == 16 || tile_rows == 32,
                        "tile_rows must be 4, 8, 16 or 32");
        }
        CFD_REQUIRE(chunk >= 1 && chunk <= 65535, "chunk must be in [1, 65535]");
        m->wave_variant = variant;
        m->wave_ty = tile_rows;
        m->wave_L = chunk;
        // bit 0: XCD-contiguous tiles; bit 1 (row form only): odd steps walk
        // the tiles in reverse
        m->wave_remap = row_form ? (xcd_remap & 3) : (xcd_remap ? 1 : 0);
    });
}

}  // extern "C"
