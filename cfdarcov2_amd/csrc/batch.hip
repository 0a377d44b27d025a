// Batched ensemble of independent heat and acoustic cases (FP32, uniform
// meshes; no reference counterpart).
//
// One launch advances every case by one step: the grid is the concatenation of
// all cases' march tiles (tile rows x 64 columns x `chunk` planes), each
// workgroup finds its case by a binary search over the cases' first-tile
// offsets (scalar loads of a small descriptor table) and runs the same tile
// code as the single-mesh kernels, so every case's fields equal a plain
// single-case run bit for bit.  A run of n steps is captured once into a HIP
// graph (n kernel nodes, no memset nodes) and replayed with one launch.
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "march.hpp"
#include "row.hpp"
#include "row2.hpp"

namespace cfd {

struct BatchCaseDev {
    UniformGeo<float> g;
    float coef;          // heat: alpha; wave: K
    float dt;
    int kind;            // CFD_CASE_HEAT / CFD_CASE_WAVE
    int nzt, nyt;
    int64_t tile0;       // first tile of the case in the batched grid
    float* a[4];
    float* b[4];
};

template <int TYT, int R> union BatchLds {
    MarchLds<float, TYT, R> w;
    HeatLds<float, TYT, R> h;
};

template <int TYT, int R>
__global__ void __launch_bounds__(64 * TYT) batch_step_k(const BatchCaseDev* __restrict__ cases, int ncases,
                                                         int parity, int L) {
    __shared__ BatchLds<TYT, R> lds;
    const int64_t tile = blockIdx.x;
    int lo = 0, hi = ncases - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cases[mid].tile0 <= tile) lo = mid;
        else hi = mid - 1;
    }
    const BatchCaseDev& cs = cases[lo];
    const int64_t t = tile - cs.tile0;
    const int64_t zt = t % cs.nzt;
    const int64_t yt = (t / cs.nzt) % cs.nyt;
    const int64_t xc = t / (int64_t(cs.nzt) * cs.nyt);
    float* const* in = parity ? cs.b : cs.a;
    float* const* out = parity ? cs.a : cs.b;
    if (cs.kind == CFD_CASE_HEAT) {
        if (cs.g.vol_pow2) heat_march_tile<float, TYT, R, true>(cs.g, in[0], out[0], cs.coef, cs.dt, L, zt, yt, xc, lds.h);
        else heat_march_tile<float, TYT, R, false>(cs.g, in[0], out[0], cs.coef, cs.dt, L, zt, yt, xc, lds.h);
    } else if (cs.g.vol_pow2) {
        wave_march_tile<float, TYT, R, false, true>(cs.g, in[0], in[1], in[2], in[3], out[0], out[1], out[2],
                                                       out[3], cs.coef, nullptr, cs.dt, L, zt, yt, xc, lds.w);
    } else {
        wave_march_tile<float, TYT, R, false, false>(cs.g, in[0], in[1], in[2], in[3], out[0], out[1], out[2],
                                                        out[3], cs.coef, nullptr, cs.dt, L, zt, yt, xc, lds.w);
    }
}

// Row form of the batched step (form 1): every case's z extent is a multiple
// of 4 and at most 256, so a case's tiles are (y tile, x chunk) pairs of the
// row-form kernels (one wave per z row, TY output waves and two halo waves).
template <int TY>
__global__ void __launch_bounds__(64 * (TY + 2)) batch_row_k(const BatchCaseDev* __restrict__ cases, int ncases,
                                                             int parity, int L) {
    using V4 = row::Row4<float>::type;
    __shared__ V4 sp[TY + 2][64];
    __shared__ V4 sv[TY + 2][64];
    const int64_t tile = blockIdx.x;
    int lo = 0, hi = ncases - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cases[mid].tile0 <= tile) lo = mid;
        else hi = mid - 1;
    }
    const BatchCaseDev& cs = cases[lo];
    const int64_t t = tile - cs.tile0;
    const int yt = int(t % cs.nyt);
    const int64_t xc = t / cs.nyt;
    const int wv = __builtin_amdgcn_readfirstlane(int(threadIdx.y));
    const int64_t X = cs.g.X;
    const int64_t x0 = xc * L;
    const int64_t x1 = (x0 + L < X) ? x0 + L : X;
    const int y0 = yt * TY * row::rows_per_wave(cs.g.Z);
    float* const* in = parity ? cs.b : cs.a;
    float* const* out = parity ? cs.a : cs.b;
    const bool pw = cs.g.vol_pow2 != 0;
    if (cs.kind == CFD_CASE_HEAT) {
        const bool inner = x0 >= 1 && x1 + 1 <= X;
        if (pw) {
            if (inner) row::heat_row_tile<TY, true, false, true>(cs.g, in[0], out[0], cs.coef, cs.dt, x0, x1, y0, wv, sp);
            else row::heat_row_tile<TY, true, true, true>(cs.g, in[0], out[0], cs.coef, cs.dt, x0, x1, y0, wv, sp);
        } else {
            if (inner) row::heat_row_tile<TY, false, false, true>(cs.g, in[0], out[0], cs.coef, cs.dt, x0, x1, y0, wv, sp);
            else row::heat_row_tile<TY, false, true, true>(cs.g, in[0], out[0], cs.coef, cs.dt, x0, x1, y0, wv, sp);
        }
    } else {
        const bool inner = x0 >= 1 && x1 + 2 <= X;
#define CFD_ROW_TILE(PW, XB)                                                                                     \
    row::row_tile<float, TY, PW, false, XB, true>(cs.g, in[0], in[1], in[2], in[3], out[0], out[1], out[2], out[3], \
                                                  cs.coef, nullptr, cs.dt, x0, x1, y0, wv, false, sp, sv)
        if (pw) {
            if (inner) CFD_ROW_TILE(true, false);
            else CFD_ROW_TILE(true, true);
        } else {
            if (inner) CFD_ROW_TILE(false, false);
            else CFD_ROW_TILE(false, true);
        }
#undef CFD_ROW_TILE
    }
}

// Two steps per launch (form 2): acoustic cases run the several-rows-per-wave
// two-step tile, heat cases the two-step heat tile, both NW waves wide.
template <int NW>
__global__ void __launch_bounds__(64 * NW) batch_row2_k(const BatchCaseDev* __restrict__ cases, int ncases,
                                                        int parity, int L) {
    using V4 = row::Row4<float>::type;
    __shared__ V4 s0[NW][64];
    __shared__ V4 s1[NW][64];
    __shared__ V4 s2[NW][64];
    __shared__ V4 s3[NW][64];
    const int64_t tile = blockIdx.x;
    int lo = 0, hi = ncases - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cases[mid].tile0 <= tile) lo = mid;
        else hi = mid - 1;
    }
    const BatchCaseDev& cs = cases[lo];
    const int64_t t = tile - cs.tile0;
    const int yt = int(t % cs.nyt);
    const int64_t xc = t / cs.nyt;
    const int wv = __builtin_amdgcn_readfirstlane(int(threadIdx.y));
    const int64_t X = cs.g.X;
    const int64_t x0 = xc * L;
    const int64_t x1 = (x0 + L < X) ? x0 + L : X;
    const int band = NW * row::rows_per_wave(cs.g.Z);
    float* const* in = parity ? cs.b : cs.a;
    float* const* out = parity ? cs.a : cs.b;
    const bool pw = cs.g.vol_pow2 != 0;
    if (cs.kind == CFD_CASE_HEAT) {
        const int y0 = yt * (band - 4);
        const bool inner = x0 >= 2 && x1 + 2 <= X;
#define CFD_HEAT2(PW, XB) \
    row::heat_row2_tile<float, NW, PW, XB, true>(cs.g, in[0], out[0], cs.coef, cs.dt, x0, x1, y0, wv, s0, s1)
        if (pw) {
            if (inner) CFD_HEAT2(true, false);
            else CFD_HEAT2(true, true);
        } else {
            if (inner) CFD_HEAT2(false, false);
            else CFD_HEAT2(false, true);
        }
#undef CFD_HEAT2
    } else {
        const int y0 = yt * (band - 8);
        const bool inner = x0 >= 4 && x1 + 4 <= X;
#define CFD_WAVE2(PW, XB)                                                                                  \
    row2::tile2m<float, NW, PW, XB, true>(cs.g, in[0], in[1], in[2], in[3], out[0], out[1], out[2], out[3],    \
                                          cs.coef, cs.dt, x0, x1, y0, wv, false, s0, s1, s2, s3)
        if (pw) {
            if (inner) CFD_WAVE2(true, false);
            else CFD_WAVE2(true, true);
        } else {
            if (inner) CFD_WAVE2(false, false);
            else CFD_WAVE2(false, true);
        }
#undef CFD_WAVE2
    }
}

}  // namespace cfd

using namespace cfd;

struct cfd_batch_s {
    cfd_ctx_t ctx = nullptr;
    std::vector<BatchCaseDev> host;
    BatchCaseDev* dev = nullptr;
    int tile_rows = 4;
    int chunk = 8;
    int form = 0;  // 0: 2.5D march tiles, 1: row form, 2: row form, two steps per launch
    int64_t total_tiles = 0;
    // form 2: single-step row-form layout (4-row tiles) for the steps left after
    // the pairs of two-step launches
    std::vector<BatchCaseDev> host1;
    BatchCaseDev* dev1 = nullptr;
    int64_t total_tiles1 = 0;
    bool dirty = true;
    hipStream_t capture = nullptr;
    // stream of the last cfd_batch_run: destroy waits on it before freeing the
    // case descriptors its kernels read
    hipStream_t last_run = nullptr;
    bool ran = false;
    // instantiated graphs keyed by (steps, tile rows, chunk)
    std::map<std::tuple<int64_t, int, int>, hipGraphExec_t> graphs;
};

namespace cfd {

// rows per tile of case c in `form` with tile parameter T (tile rows, or waves
// for form 2)
static int64_t tile_rows_of(const BatchCaseDev& c, int form, int T) {
    const int64_t rw = row::rows_per_wave(c.g.Z);
    if (form == 1) return int64_t(T) * rw;
    if (form == 2) return int64_t(T) * rw - (c.kind == CFD_CASE_HEAT ? 4 : 8);
    return T;
}

// the two-step form needs output rows in every case's tiles
static void check_form2(const std::vector<BatchCaseDev>& cases, int T) {
    for (const auto& c : cases)
        if (tile_rows_of(c, 2, T) < 1)
            throw Error(CFD_E_UNSUPPORTED, "two-step batch tiles of this width leave a case without output rows");
}

static int64_t layout_into(std::vector<BatchCaseDev>& cases, int form, int T, int L) {
    int64_t off = 0;
    for (auto& c : cases) {
        c.nzt = form >= 1 ? 1 : int((c.g.Z + 63) / 64);  // a row-form tile spans the z row
        const int64_t rows = tile_rows_of(c, form, T);
        CFD_REQUIRE(rows >= 1, "batch tile without output rows");
        c.nyt = int((c.g.Y + rows - 1) / rows);
        const int64_t nxc = (c.g.X + L - 1) / L;
        c.tile0 = off;
        off += int64_t(c.nzt) * c.nyt * nxc;
    }
    CFD_REQUIRE(off < (int64_t(1) << 31), "batched grid too large");
    return off;
}

static void batch_layout(cfd_batch_s* b, hipStream_t s) {
    b->total_tiles = layout_into(b->host, b->form, b->tile_rows, b->chunk);
    const size_t bytes = b->host.size() * sizeof(BatchCaseDev);
    CFD_HIP(hipMemcpyAsync(b->dev, b->host.data(), bytes, hipMemcpyHostToDevice, s));
    if (b->form == 2) {
        b->host1 = b->host;
        b->total_tiles1 = layout_into(b->host1, 1, 4, b->chunk);
        CFD_HIP(hipMemcpyAsync(b->dev1, b->host1.data(), bytes, hipMemcpyHostToDevice, s));
    }
    CFD_HIP(hipStreamSynchronize(s));
    b->dirty = false;
}

static void batch_launch(const cfd_batch_s* b, int parity, hipStream_t s) {
    const int n = int(b->host.size());
    const dim3 grid(unsigned(b->total_tiles));
    if (b->form == 2) {  // two steps per launch
        switch (b->tile_rows) {
            case 8: batch_row2_k<8><<<grid, dim3(64, 8), 0, s>>>(b->dev, n, parity, b->chunk); break;
            case 10: batch_row2_k<10><<<grid, dim3(64, 10), 0, s>>>(b->dev, n, parity, b->chunk); break;
            case 12: batch_row2_k<12><<<grid, dim3(64, 12), 0, s>>>(b->dev, n, parity, b->chunk); break;
            default: throw Error(CFD_E_INVALID, "two-step batch tiles have 8, 10 or 12 waves");
        }
        return;
    }
    if (b->form == 1) {
        switch (b->tile_rows) {
            case 4: batch_row_k<4><<<grid, dim3(64, 6), 0, s>>>(b->dev, n, parity, b->chunk); break;
            case 8: batch_row_k<8><<<grid, dim3(64, 10), 0, s>>>(b->dev, n, parity, b->chunk); break;
            default: throw Error(CFD_E_INVALID, "row-form batch tile rows must be 4 or 8");
        }
        return;
    }
    switch (b->tile_rows) {
        case 4: batch_step_k<4, 1><<<grid, dim3(64, 4), 0, s>>>(b->dev, n, parity, b->chunk); break;
        case 8: batch_step_k<4, 2><<<grid, dim3(64, 4), 0, s>>>(b->dev, n, parity, b->chunk); break;
        case 16: batch_step_k<4, 4><<<grid, dim3(64, 4), 0, s>>>(b->dev, n, parity, b->chunk); break;
        default: throw Error(CFD_E_INVALID, "batch tile rows must be 4, 8 or 16");
    }
}

// nsteps steps from state_a: in form 2, pairs of two-step launches (a -> b ->
// a) and the remaining steps one per launch on the single-step layout
static void batch_steps(const cfd_batch_s* b, int64_t nsteps, hipStream_t s) {
    int64_t k0 = 0;
    if (b->form == 2) {
        for (; k0 + 4 <= nsteps; k0 += 4) {
            batch_launch(b, 0, s);
            batch_launch(b, 1, s);
        }
        const int n = int(b->host1.size());
        const dim3 grid(unsigned(b->total_tiles1));
        for (int64_t k = k0; k < nsteps; ++k)
            batch_row_k<4><<<grid, dim3(64, 6), 0, s>>>(b->dev1, n, int(k & 1), b->chunk);
        return;
    }
    for (int64_t k = 0; k < nsteps; ++k) batch_launch(b, int(k & 1), s);
}

static void drop_graphs(cfd_batch_s* b) {
    for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
}

}  // namespace cfd

extern "C" {

int cfd_batch_create(cfd_ctx_t ctx, const cfd_batch_case* cases, int ncases, cfd_batch_t* out) {
    return guarded([&] {
        CFD_REQUIRE(ctx && cases && out, "cfd_batch_create: null argument");
        CFD_REQUIRE(ncases >= 1, "cfd_batch_create: need at least one case");
        auto b = std::make_unique<cfd_batch_s>();
        b->ctx = ctx;
        for (int i = 0; i < ncases; ++i) {
            const cfd_batch_case& c = cases[i];
            CFD_REQUIRE(c.mesh != nullptr, "cfd_batch_create: null mesh");
            if (c.mesh->dtype != CFD_F32 || !c.mesh->uniform)
                throw Error(CFD_E_UNSUPPORTED, "cfd_batch_create: cases must be FP32 uniform meshes");
            CFD_REQUIRE(c.kind == CFD_CASE_HEAT || c.kind == CFD_CASE_WAVE, "cfd_batch_create: unknown case kind");
            const int nf = c.kind == CFD_CASE_HEAT ? 1 : 4;
            for (int f = 0; f < nf; ++f) {
                CFD_REQUIRE(c.state_a[f] && c.state_b[f], "cfd_batch_create: null state pointer");
                for (int k = 0; k < nf; ++k)
                    CFD_REQUIRE(c.state_a[f] != c.state_b[k], "cfd_batch_create: state_a and state_b overlap");
            }
            BatchCaseDev d{};
            d.g = c.mesh->ug32;
            d.coef = float(c.coef);
            d.dt = float(c.dt);
            d.kind = c.kind;
            for (int f = 0; f < 4; ++f) {
                d.a[f] = (float*)(f < nf ? c.state_a[f] : nullptr);
                d.b[f] = (float*)(f < nf ? c.state_b[f] : nullptr);
            }
            b->host.push_back(d);
        }
        b->dev = (BatchCaseDev*)pool_alloc(ctx, int64_t(ncases) * int64_t(sizeof(BatchCaseDev)), nullptr);
        try {
            batch_layout(b.get(), nullptr);
            // the descriptors are written on the null stream and read on the
            // caller's stream, which may be a non-blocking one
            CFD_HIP(hipStreamSynchronize(nullptr));
            CFD_HIP(hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking));
        } catch (...) {
            pool_free(ctx, b->dev, nullptr);
            throw;
        }
        *out = b.release();
    });
}

int cfd_batch_set_form(cfd_batch_t b, int form) {
    return guarded([&] {
        CFD_REQUIRE(b != nullptr, "null batch");
        CFD_REQUIRE(form >= 0 && form <= 2,
                    "batch form must be 0 (march tiles), 1 (row form) or 2 (row form, two steps per launch)");
        if (form >= 1) {
            for (const auto& c : b->host) {
                if (!wave_row_supported(c.g.Z))
                    throw Error(CFD_E_UNSUPPORTED,
                                "the row form needs every case's z extent to be a multiple of 4 and at most 256");
                for (int f = 0; f < 4; ++f)
                    if ((reinterpret_cast<uintptr_t>(c.a[f]) & 15) || (reinterpret_cast<uintptr_t>(c.b[f]) & 15))
                        throw Error(CFD_E_UNSUPPORTED, "the row form needs 16-byte aligned state fields");
            }
            if (form == 1 && b->tile_rows != 4 && b->tile_rows != 8) b->tile_rows = 4;
            if (form == 2) {
                if (b->tile_rows != 8 && b->tile_rows != 10 && b->tile_rows != 12) b->tile_rows = 12;
                check_form2(b->host, b->tile_rows);
                if (!b->dev1) {
                    b->dev1 = (BatchCaseDev*)pool_alloc(b->ctx, int64_t(b->host.size()) * int64_t(sizeof(BatchCaseDev)),
                                                        nullptr);
                    // allocated on the null stream, first written on the run's
                    CFD_HIP(hipStreamSynchronize(nullptr));
                }
            }
        }
        if (form != b->form) {
            b->form = form;
            b->dirty = true;
        }
    });
}

int cfd_batch_set_shape(cfd_batch_t b, int tile_rows, int chunk) {
    return guarded([&] {
        CFD_REQUIRE(b != nullptr, "null batch");
        if (b->form == 2) {
            CFD_REQUIRE(tile_rows == 8 || tile_rows == 10 || tile_rows == 12,
                        "two-step batch tiles have 8, 10 or 12 waves");
            check_form2(b->host, tile_rows);
        } else
            CFD_REQUIRE(tile_rows == 4 || tile_rows == 8 || (tile_rows == 16 && b->form == 0),
                        "batch tile rows must be 4, 8 or 16 (4 or 8 in the row form)");
        CFD_REQUIRE(chunk >= 1 && chunk <= 4096, "batch chunk must be in [1, 4096]");
        if (tile_rows != b->tile_rows || chunk != b->chunk) {
            b->tile_rows = tile_rows;
            b->chunk = chunk;
            b->dirty = true;
        }
    });
}

int cfd_batch_run(cfd_batch_t b, int64_t nsteps, int use_graph, void* stream) {
    return guarded([&] {
        CFD_REQUIRE(b != nullptr, "null batch");
        CFD_REQUIRE(nsteps >= 0, "cfd_batch_run: negative step count");
        hipStream_t s = as_stream(stream);
        if (b->ran && b->last_run != s) CFD_HIP(hipStreamSynchronize(b->last_run));
        b->last_run = s;
        b->ran = true;
        if (b->dirty) {  // new tile shape: new layout, and graphs of the old one are stale
            CFD_HIP(hipStreamSynchronize(s));
            drop_graphs(b);
            batch_layout(b, s);
        }
        if (!use_graph) {
            batch_steps(b, nsteps, s);
            CFD_HIP(hipGetLastError());
            return;
        }
        // a replay starts from state_a again, so the state must end there
        CFD_REQUIRE(nsteps % 2 == 0, "cfd_batch_run: graph runs need an even step count");
        if (nsteps == 0) return;
        const auto key = std::make_tuple(nsteps, b->tile_rows * 4 + b->form, b->chunk);
        auto it = b->graphs.find(key);
        if (it == b->graphs.end()) {
            hipGraph_t graph = nullptr;
            CFD_HIP(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal));
            try {
                batch_steps(b, nsteps, b->capture);
            } catch (...) {
                (void)hipStreamEndCapture(b->capture, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            CFD_HIP(hipStreamEndCapture(b->capture, &graph));
            hipGraphExec_t exec = nullptr;
            hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            CFD_HIP(e);
            it = b->graphs.emplace(key, exec).first;
        }
        CFD_HIP(hipGraphLaunch(it->second, s));
    });
}

int cfd_batch_destroy(cfd_batch_t b) {
    return guarded([&] {
        if (!b) return;
        // in-flight kernels of the last run still read the descriptors
        if (b->ran) CFD_HIP(hipStreamSynchronize(b->last_run));
        drop_graphs(b);
        if (b->capture) (void)hipStreamDestroy(b->capture);
        pool_free(b->ctx, b->dev, nullptr);
        if (b->dev1) pool_free(b->ctx, b->dev1, nullptr);
        CFD_HIP(hipStreamSynchronize(nullptr));
        delete b;
    });
}

}  // extern "C"
