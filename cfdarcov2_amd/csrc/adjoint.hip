// Adjoint of the fused acoustic step on uniform meshes (no reference
// counterpart), for gradients of sensor-trace misfits with respect to the
// initial pressure and a per-cell K = -c^2 field.
//
// Forward step S (the fused step of the wave kernels, same operation order):
//   u' = u - dt Gx p,  v' = v - dt Gy p,  w' = w - dt Gz p
//   p' = p + dt K (Dx u' + Dy v' + Dz w')
// G_a is the Green-Gauss gradient with the reference boundary rule
// (neighbour = self) and D_a the same with zero normal velocity on the walls.
// On a uniform mesh the two are negative transposes of each other
// (summation by parts: interior rows are h (q_{i+1} - q_{i-1}), the boundary
// rows of G are h (q_1 - q_0) and h (q_{n-1} - q_{n-2}), those of D are
// h (q_0 + q_1) and -h (q_{n-2} + q_{n-1}), h = A / 2V), so G^T = -D and
// D^T = -G, and the transposed step S^T is again a fused two-stage stencil:
//   mu_a = lam_a - G_a(q),  q = (dt K) lam_p
//   nu_p = lam_p + dt (Dx mu_x + Dy mu_y + Dz mu_z)
// The K gradient adds lam_p(t) dt (D . vel(s_t)) per step, with the forward
// velocities of step t recomputed from checkpoints.
//
// Sensor traces: the forward kernel writes p' at up to 64 sensor cells, the
// adjoint kernel adds the trace gradients of the previous step to nu_p there.
// The sensor list is a kernel argument; each wave tests it against its own row
// of 64 z-cells with scalar instructions, so cells without sensors pay nothing
// but that scan.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "row.hpp"

namespace cfd {

// row form (adj_row.hip)
template <class T> bool adj_row_usable(const cfd_mesh_s* m, const void* const* fields, int n);
template <class T>
void launch_wave_row_aux(const cfd_mesh_s* m, int mode, const void* const in[4], void* const out[4], double K,
                         const void* Kf, double dt, const SensorList* sens, const row::RowAux<T>& aux, int rev,
                         hipStream_t s);

// A device copy of the sensor list for the row-form kernels, which read it
// from memory the caches keep (the kernel-argument copy of 1.5 KB costs each
// workgroup a slow first read), followed by the single-step row tiles' per-tile
// sensor masks.  One pool block; the copy is waited for, so the host staging
// vector may go.
template <class T> std::vector<uint64_t> adj_row_tile_sensors(const cfd_mesh_s* m, const SensorList& sens);
struct DevSensors {
    const SensorList* list = nullptr;
    const uint64_t* tiles = nullptr;
    void* block = nullptr;
};
template <class T> static DevSensors upload_sensors(const cfd_mesh_s* m, const SensorList& sens, hipStream_t s) {
    const std::vector<uint64_t> tiles = adj_row_tile_sensors<T>(m, sens);
    const size_t lb = (sizeof(SensorList) + 255) / 256 * 256;
    std::vector<char> host(lb + tiles.size() * sizeof(uint64_t), 0);
    std::memcpy(host.data(), &sens, sizeof(SensorList));
    std::memcpy(host.data() + lb, tiles.data(), tiles.size() * sizeof(uint64_t));
    DevSensors d;
    d.block = pool_alloc(m->ctx, int64_t(host.size()), s);
    hipError_t e = hipMemcpyAsync(d.block, host.data(), host.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        pool_free(m->ctx, d.block, s);
        CFD_HIP(e);
    }
    d.list = static_cast<const SensorList*>(d.block);
    d.tiles = reinterpret_cast<const uint64_t*>(static_cast<char*>(d.block) + lb);
    return d;
}

namespace {

template <class T> __device__ inline T a_lin(T c, T nb) { return (c + nb) * T(0.5); }
// Green-Gauss face pair ((-f_lo) A + f_hi A) / V; with PW the volume is an
// exact power of two and the division is the (bit-identical) product with 1/V.
template <bool PW, class T> __device__ inline T a_pair(T f_lo, T f_hi, T A, const UniformGeo<T>& g) {
    const T t_lo = (-f_lo) * A;
    const T t_hi = f_hi * A;
    if constexpr (PW) return (t_lo + t_hi) * g.inv_vol;
    else return (t_lo + t_hi) / g.vol;
}

struct TCell {
    int64_t x, y, z, id;
    bool valid;
};

template <class T> __device__ inline TCell tcell(const UniformGeo<T>& g) {
    TCell c;
    c.z = int64_t(blockIdx.x) * 64 + threadIdx.x;
    c.y = int64_t(blockIdx.y) * 4 + threadIdx.y;
    c.x = blockIdx.z;
    c.valid = c.z < g.Z && c.y < g.Y;
    c.id = (c.x * g.Y + c.y) * g.Z + c.z;
    return c;
}

// Sensor j of the wave holding this thread's row, or -1.  `row` and `z0` are
// wave-uniform; the loop runs on the scalar unit.
__device__ inline int sensor_of(const SensorList& s, int row, int z0, int z) {
    int hit = -1;
    for (int j = 0; j < s.n; ++j) {
        if (s.row[j] == row && s.z[j] >= z0 && s.z[j] < z0 + 64) {
            if (s.z[j] == z) hit = j;
        }
    }
    return hit;
}

__device__ inline int wave_row(int64_t x, int64_t Y) {
    return __builtin_amdgcn_readfirstlane(int(x * Y) + int(blockIdx.y) * 4 + int(threadIdx.y));
}

// velocity update of cell j along axis a from p (reference boundary rule)
template <bool PW, class T>
__device__ inline T fwd_vel(const UniformGeo<T>& g, const T* __restrict__ p, const T* __restrict__ q, int64_t j,
                            int64_t stride, bool lo, bool hi, int a, T dt) {
    const T pc = p[j];
    const T pl = lo ? pc : p[j - stride];
    const T ph = hi ? pc : p[j + stride];
    const T G = a_pair<PW>(a_lin(pc, pl), a_lin(pc, ph), g.area[a], g);
    return dt * (-G) + q[j];
}

}  // namespace

// Forward step with optional sensor recording: bitwise the same arithmetic as
// the per-thread fused acoustic kernel.
template <class T, bool PW>
__global__ void __launch_bounds__(256) wave_trace_k(UniformGeo<T> g, const T* __restrict__ p,
                                                    const T* __restrict__ u, const T* __restrict__ v,
                                                    const T* __restrict__ w, T* __restrict__ po,
                                                    T* __restrict__ uo, T* __restrict__ vo, T* __restrict__ wo, T K,
                                                    const T* __restrict__ Kf, T dt, SensorList sens,
                                                    T* __restrict__ trace) {
    const TCell c = tcell(g);
    const int row = wave_row(c.x, g.Y);
    if (!c.valid) return;
    const T* vel[3] = {u, v, w};
    T* velo[3] = {uo, vo, wo};
    const int64_t crd[3] = {c.x, c.y, c.z};
    const int64_t ext[3] = {g.X, g.Y, g.Z};
    const int64_t str[3] = {g.Y * g.Z, g.Z, 1};
    T S = T(0);
    for (int a = 0; a < 3; ++a) {
        const bool lo = crd[a] == 0, hi = crd[a] == ext[a] - 1;
        const T qc = fwd_vel<PW>(g, p, vel[a], c.id, str[a], lo, hi, a, dt);
        velo[a][c.id] = qc;
        T f_lo = T(0), f_hi = T(0);
        if (!lo) f_lo = a_lin(qc, fwd_vel<PW>(g, p, vel[a], c.id - str[a], str[a], crd[a] - 1 == 0, false, a, dt));
        if (!hi) f_hi = a_lin(qc, fwd_vel<PW>(g, p, vel[a], c.id + str[a], str[a], false, crd[a] + 1 == ext[a] - 1, a, dt));
        const T G = a_pair<PW>(f_lo, f_hi, g.area[a], g);
        S = (a == 0) ? G : S + G;
    }
    const T k = Kf ? Kf[c.id] : K;
    const T pn = dt * (k * S) + p[c.id];
    po[c.id] = pn;
    if (trace) {
        const int j = sensor_of(sens, row, int(blockIdx.x) * 64, int(c.z));
        if (j >= 0) trace[j] = pn;
    }
}

// adjoint velocity of cell j along axis a: mu = lam_a - G_a(q), q = (dt k) lam_p
template <bool PW, class T>
__device__ inline T adj_vel(const UniformGeo<T>& g, const T* __restrict__ lp, const T* __restrict__ la, T K,
                            const T* __restrict__ Kf, int64_t j, int64_t stride, bool lo, bool hi, int a, T dt) {
    auto q = [&](int64_t i) { return (dt * (Kf ? Kf[i] : K)) * lp[i]; };
    const T qc = q(j);
    const T ql = lo ? qc : q(j - stride);
    const T qh = hi ? qc : q(j + stride);
    const T G = a_pair<PW>(a_lin(qc, ql), a_lin(qc, qh), g.area[a], g);
    return la[j] - G;
}

// One transposed step lam_t -> lam_{t-1}.  src (nullable): trace gradients of
// step t-1, added at the sensors.  vel_t/gK (nullable): the forward
// velocities of step t and the K-gradient accumulator.
template <class T, bool PW>
__global__ void __launch_bounds__(256) wave_adj_k(UniformGeo<T> g, const T* __restrict__ lp,
                                                  const T* __restrict__ lu, const T* __restrict__ lv,
                                                  const T* __restrict__ lw, T* __restrict__ op,
                                                  T* __restrict__ ou, T* __restrict__ ov, T* __restrict__ ow, T K,
                                                  const T* __restrict__ Kf, T dt, SensorList sens,
                                                  const T* __restrict__ src, const T* __restrict__ ut,
                                                  const T* __restrict__ vt, const T* __restrict__ wt,
                                                  T* __restrict__ gK) {
    const TCell c = tcell(g);
    const int row = wave_row(c.x, g.Y);
    if (!c.valid) return;
    const T* lam[3] = {lu, lv, lw};
    T* outv[3] = {ou, ov, ow};
    const T* velt[3] = {ut, vt, wt};
    const int64_t crd[3] = {c.x, c.y, c.z};
    const int64_t ext[3] = {g.X, g.Y, g.Z};
    const int64_t str[3] = {g.Y * g.Z, g.Z, 1};
    T S = T(0), St = T(0);
    for (int a = 0; a < 3; ++a) {
        const bool lo = crd[a] == 0, hi = crd[a] == ext[a] - 1;
        const T mc = adj_vel<PW>(g, lp, lam[a], K, Kf, c.id, str[a], lo, hi, a, dt);
        outv[a][c.id] = mc;
        T f_lo = T(0), f_hi = T(0);
        if (!lo) f_lo = a_lin(mc, adj_vel<PW>(g, lp, lam[a], K, Kf, c.id - str[a], str[a], crd[a] - 1 == 0, false, a, dt));
        if (!hi)
            f_hi = a_lin(mc, adj_vel<PW>(g, lp, lam[a], K, Kf, c.id + str[a], str[a], false, crd[a] + 1 == ext[a] - 1, a, dt));
        const T G = a_pair<PW>(f_lo, f_hi, g.area[a], g);
        S = (a == 0) ? G : S + G;
        if (gK) {  // forward divergence of step t, as the forward computes it
            const T vc = velt[a][c.id];
            const T tl = lo ? T(0) : a_lin(vc, velt[a][c.id - str[a]]);
            const T th = hi ? T(0) : a_lin(vc, velt[a][c.id + str[a]]);
            const T Gt = a_pair<PW>(tl, th, g.area[a], g);
            St = (a == 0) ? Gt : St + Gt;
        }
    }
    const T lpc = lp[c.id];
    if (gK) gK[c.id] = gK[c.id] + (lpc * dt) * St;
    T np = dt * S + lpc;
    if (src) {
        const int j = sensor_of(sens, row, int(blockIdx.x) * 64, int(c.z));
        if (j >= 0) np = np + src[j];
    }
    op[c.id] = np;
}

// lam_p[sensor j] += src[j]
template <class T> __global__ void scatter_add_k(T* __restrict__ lp, SensorList sens, const T* __restrict__ src) {
    const int j = threadIdx.x;
    if (j < sens.n) {
        const int64_t cell = sens.cell[j];
        lp[cell] = lp[cell] + src[j];
    }
}

namespace {
const dim3 kBlk(64, 4, 1);
template <class T> dim3 tgrid(const UniformGeo<T>& g) {
    return dim3(unsigned((g.Z + 63) / 64), unsigned((g.Y + 3) / 4), unsigned(g.X));
}
template <class T> const UniformGeo<T>& ug(const cfd_mesh_s* m) {
    if constexpr (sizeof(T) == 4) return m->ug32; else return m->ug64;
}
}  // namespace

SensorList make_sensors(const cfd_mesh_s* m, const int64_t* cells, int n) {
    CFD_REQUIRE(n >= 0 && n <= kMaxSensors, "sensor count must be in [0, 64]");
    CFD_REQUIRE(n == 0 || cells != nullptr, "null sensor list");
    CFD_REQUIRE(m->X * m->Y <= int64_t(0x7fffffff), "mesh too large for sensor rows");
    SensorList s{};
    s.n = n;
    for (int j = 0; j < n; ++j) {
        CFD_REQUIRE(cells[j] >= 0 && cells[j] < m->N, "sensor cell out of range");
        for (int i = 0; i < j; ++i) CFD_REQUIRE(cells[i] != cells[j], "duplicate sensor cell");
        s.cell[j] = cells[j];
        s.row[j] = int(cells[j] / m->Z);
        s.x[j] = int(cells[j] / (m->Z * m->Y));
        s.y[j] = int((cells[j] / m->Z) % m->Y);
        s.z[j] = int(cells[j] % m->Z);
    }
    return s;
}

template <class T>
void launch_wave_trace_step(const cfd_mesh_s* m, const void* const in[4], void* const out[4], double K,
                            const void* Kf, double dt, const SensorList& sens, void* trace, hipStream_t s) {
    const UniformGeo<T>& g = ug<T>(m);
    auto* kern = g.vol_pow2 ? wave_trace_k<T, true> : wave_trace_k<T, false>;
    kern<<<tgrid(g), kBlk, 0, s>>>(g, (const T*)in[0], (const T*)in[1], (const T*)in[2], (const T*)in[3],
                                   (T*)out[0], (T*)out[1], (T*)out[2], (T*)out[3], T(K), (const T*)Kf, T(dt), sens,
                                   (T*)trace);
}

template <class T>
void launch_wave_adj_step(const cfd_mesh_s* m, const void* const in[4], void* const out[4], double K,
                          const void* Kf, double dt, const SensorList& sens, const void* src,
                          const void* const velt[3], void* gK, hipStream_t s) {
    const UniformGeo<T>& g = ug<T>(m);
    const T* vt[3] = {nullptr, nullptr, nullptr};
    if (gK)
        for (int a = 0; a < 3; ++a) vt[a] = (const T*)velt[a];
    auto* kern = g.vol_pow2 ? wave_adj_k<T, true> : wave_adj_k<T, false>;
    kern<<<tgrid(g), kBlk, 0, s>>>(g, (const T*)in[0], (const T*)in[1], (const T*)in[2], (const T*)in[3],
                                   (T*)out[0], (T*)out[1], (T*)out[2], (T*)out[3], T(K), (const T*)Kf, T(dt), sens,
                                   (const T*)src, vt[0], vt[1], vt[2], (T*)gK);
}

template <class T> void launch_scatter_add(void* lp, const SensorList& sens, const void* src, hipStream_t s) {
    if (sens.n == 0) return;
    scatter_add_k<T><<<1, 64, 0, s>>>((T*)lp, sens, (const T*)src);
}

// ---------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------

// nsteps forward steps from state_a, ping-pong with state_b; after step t
// (1-based) p at the sensors goes to traces[(t-1) * n]; before steps
// 0, every, 2 every, ... the state {p,u,v,w} is copied to checkpoint k.
template <class T>
void run_wave_traces(const cfd_mesh_s* m, void* const a[4], void* const b[4], double K, const void* Kf, double dt,
                     int64_t nsteps, const SensorList& sens, void* traces, void* ckpt, int64_t every,
                     hipStream_t s) {
    const size_t field = size_t(m->N) * sizeof(T);
    const void* fields[9] = {a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3], Kf};
    const bool rowf = adj_row_usable<T>(m, fields, 9);
    auto trace_row = [&](int64_t t) -> T* { return traces ? (T*)traces + t * sens.n : nullptr; };
    const DevSensors ds = rowf ? upload_sensors<T>(m, sens, s) : DevSensors{};
    const SensorList* dsens = ds.list;
    try {
        for (int64_t t = 0; t < nsteps; ++t) {
            void* const* in = (t & 1) ? b : a;
            void* const* out = (t & 1) ? a : b;
            if (ckpt && t % every == 0) {
                char* dst = (char*)ckpt + size_t(t / every) * 4 * field;
                for (int f = 0; f < 4; ++f)
                    CFD_HIP(hipMemcpyAsync(dst + f * field, in[f], field, hipMemcpyDeviceToDevice, s));
            }
            T* tr = trace_row(t);
            if (rowf) {
                const row::RowAux<T> aux{T(0), tr, nullptr, nullptr, nullptr, nullptr, ds.tiles};
                launch_wave_row_aux<T>(m, 1, (const void* const*)in, out, K, Kf, dt, dsens, aux, int(t & 1), s);
            } else {
                launch_wave_trace_step<T>(m, (const void* const*)in, out, K, Kf, dt, sens, tr, s);
            }
        }
        CFD_HIP(hipGetLastError());
    } catch (...) {
        pool_free(m->ctx, ds.block, s);
        throw;
    }
    pool_free(m->ctx, ds.block, s);
}

// Reverse loop in row form (constant K).  With grad_K, each checkpoint
// segment is recomputed forward in row form, keeping every step's divergence S
// (one field per step) for the K-gradient term (lam_p dt) S_t, and the state
// in two ping-pong buffers.
template <class T>
static void run_wave_adjoint_row(const cfd_mesh_s* m, void* const la[4], void* const lb[4], double K, double dt,
                                 int64_t nsteps, const SensorList& sens, const void* tgrad, const void* ckpt,
                                 int64_t every, void* gK, hipStream_t s) {
    const size_t field = size_t(m->N) * sizeof(T);
    const T dK = T(T(dt) * T(K));
    void* ws = nullptr;     // S of each step of one segment: every fields
    void* st = nullptr;     // recomputed state: 2 x {p, u, v, w}
    const DevSensors ds = upload_sensors<T>(m, sens, s);
    const SensorList* dsens = ds.list;
    try {
        if (gK) {
            ws = pool_alloc(m->ctx, int64_t(every) * int64_t(field), s);
            st = pool_alloc(m->ctx, 8 * int64_t(field), s);
        }
        int64_t done = 0;
        // the trace gradients added after transposed step t (those of step t-1)
        auto src_of = [&](int64_t t) -> const T* {
            return (tgrad && t >= 2) ? (const T*)tgrad + (t - 2) * sens.n : nullptr;
        };
        auto adj_step = [&](int64_t t, const void* s_in) {
            void* const* in = (done & 1) ? lb : la;
            void* const* out = (done & 1) ? la : lb;
            const T* src = src_of(t);
            const row::RowAux<T> aux{dK, nullptr, nullptr, src, (const T*)s_in, s_in ? (T*)gK : nullptr, ds.tiles};
            launch_wave_row_aux<T>(m, 2, (const void* const*)in, out, K, nullptr, dt, dsens, aux, int(done & 1), s);
            ++done;
        };
        int64_t t = nsteps;
        while (t >= 1) {
            if (!gK) {
                adj_step(t, nullptr);
                --t;
                continue;
            }
            const int64_t seg = (t - 1) / every, t0 = seg * every, len = t - t0;
            const char* ck = (const char*)ckpt + size_t(seg) * 4 * field;
            for (int64_t k = 0; k < len; ++k) {  // state t0+k -> t0+k+1
                const void* in[4];
                void* out[4];
                for (int f = 0; f < 4; ++f) {
                    in[f] = k == 0 ? (const void*)(ck + f * field) : (const void*)((char*)st + (size_t(k & 1) * 4 + f) * field);
                    out[f] = (char*)st + (size_t((k + 1) & 1) * 4 + f) * field;
                }
                const row::RowAux<T> aux{T(0), nullptr, (T*)((char*)ws + size_t(k) * field), nullptr, nullptr,
                                         nullptr};
                launch_wave_row_aux<T>(m, 1, in, out, K, nullptr, dt, dsens, aux, int(k & 1), s);
            }
            for (int64_t tt = t; tt > t0; --tt) adj_step(tt, (char*)ws + size_t(tt - t0 - 1) * field);
            t = t0;
        }
        CFD_HIP(hipGetLastError());
    } catch (...) {
        pool_free(m->ctx, ws, s);
        pool_free(m->ctx, st, s);
        pool_free(m->ctx, ds.block, s);
        throw;
    }
    pool_free(m->ctx, ws, s);
    pool_free(m->ctx, st, s);
    pool_free(m->ctx, ds.block, s);
}

// Reverse loop.  adj_a holds the adjoint of the final state (without the last
// step's trace gradients); on return the adjoint of the initial state is in
// adj_a when nsteps is even, adj_b when odd.  grad_K (nullable) accumulates
// dL/dK per cell and needs the checkpoints: each segment of `every` steps is
// recomputed from its checkpoint, keeping the segment's velocities.
template <class T>
void run_wave_adjoint(const cfd_mesh_s* m, void* const la[4], void* const lb[4], double K, const void* Kf,
                      double dt, int64_t nsteps, const SensorList& sens, const void* tgrad, const void* ckpt,
                      int64_t every, void* gK, hipStream_t s) {
    const int64_t N = m->N;
    const size_t field = size_t(N) * sizeof(T);
    if (tgrad && nsteps >= 1) launch_scatter_add<T>(la[0], sens, (const T*)tgrad + (nsteps - 1) * sens.n, s);
    {
        const void* fields[10] = {la[0], la[1], la[2], la[3], lb[0], lb[1], lb[2], lb[3], gK, ckpt};
        if (!Kf && adj_row_usable<T>(m, fields, 10)) {
            run_wave_adjoint_row<T>(m, la, lb, K, dt, nsteps, sens, tgrad, ckpt, every, gK, s);
            return;
        }
    }
    void* ws = nullptr;       // recomputed velocities of one segment: every x 3 fields
    void* pbuf = nullptr;     // recomputed pressures (ping-pong)
    if (gK) {
        ws = pool_alloc(m->ctx, int64_t(every) * 3 * int64_t(field), s);
        pbuf = pool_alloc(m->ctx, 2 * int64_t(field), s);
    }
    try {
        int64_t done = 0;
        auto adj_step = [&](int64_t t, const void* const velt[3]) {
            void* const* in = (done & 1) ? lb : la;
            void* const* out = (done & 1) ? la : lb;
            const void* src = (tgrad && t >= 2) ? (const void*)((const T*)tgrad + (t - 2) * sens.n) : nullptr;
            launch_wave_adj_step<T>(m, (const void* const*)in, out, K, Kf, dt, sens, src, velt, gK, s);
            ++done;
        };
        int64_t t = nsteps;
        while (t >= 1) {
            if (!gK) {
                const void* none[3] = {nullptr, nullptr, nullptr};
                adj_step(t, none);
                --t;
                continue;
            }
            const int64_t seg = (t - 1) / every, t0 = seg * every, len = t - t0;
            const char* ck = (const char*)ckpt + size_t(seg) * 4 * field;
            for (int64_t k = 0; k < len; ++k) {  // state t0+k -> t0+k+1
                const void* in[4];
                void* out[4];
                if (k == 0) {
                    for (int f = 0; f < 4; ++f) in[f] = ck + f * field;
                } else {
                    in[0] = (char*)pbuf + size_t(k & 1) * field;
                    for (int f = 0; f < 3; ++f) in[1 + f] = (char*)ws + (size_t(k - 1) * 3 + f) * field;
                }
                out[0] = (char*)pbuf + size_t((k + 1) & 1) * field;
                for (int f = 0; f < 3; ++f) out[1 + f] = (char*)ws + (size_t(k) * 3 + f) * field;
                launch_wave_trace_step<T>(m, in, out, K, Kf, dt, sens, nullptr, s);
            }
            for (int64_t tt = t; tt > t0; --tt) {
                const int64_t k = tt - t0 - 1;
                const void* velt[3];
                for (int f = 0; f < 3; ++f) velt[f] = (char*)ws + (size_t(k) * 3 + f) * field;
                adj_step(tt, velt);
            }
            t = t0;
        }
        CFD_HIP(hipGetLastError());
    } catch (...) {
        pool_free(m->ctx, ws, s);
        pool_free(m->ctx, pbuf, s);
        throw;
    }
    pool_free(m->ctx, ws, s);
    pool_free(m->ctx, pbuf, s);
}

#define CFD_ADJ_INSTANTIATE(T)                                                                                  \
    template void launch_wave_trace_step<T>(const cfd_mesh_s*, const void* const[4], void* const[4], double,   \
                                            const void*, double, const SensorList&, void*, hipStream_t);        \
    template void launch_wave_adj_step<T>(const cfd_mesh_s*, const void* const[4], void* const[4], double,     \
                                          const void*, double, const SensorList&, const void*,                  \
                                          const void* const[3], void*, hipStream_t);                            \
    template void launch_scatter_add<T>(void*, const SensorList&, const void*, hipStream_t);                   \
    template void run_wave_traces<T>(const cfd_mesh_s*, void* const[4], void* const[4], double, const void*,   \
                                     double, int64_t, const SensorList&, void*, void*, int64_t, hipStream_t);   \
    template void run_wave_adjoint<T>(const cfd_mesh_s*, void* const[4], void* const[4], double, const void*,  \
                                      double, int64_t, const SensorList&, const void*, const void*, int64_t,    \
                                      void*, hipStream_t);
CFD_ADJ_INSTANTIATE(float)
CFD_ADJ_INSTANTIATE(double)

}  // namespace cfd
