// Process state and device helpers of the operator API: CFDArcoGlobalInit
// (src/cfdarcho_main_3d.cpp:276-347 of the reference), device memory for
// CudaDataMatrix, and the operator free functions of cuda_operators.hpp, all
// on the HIP library's C-ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <vector>
#include <ctime>
#include <iostream>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>

#include "cfdarcho_main_3d.hpp"
#include "cuda_data_matrix.hpp"
#include "cuda_operators.hpp"
#include "mesh3d.hpp"
#include "pool_allocator.hpp"

int CFDArcoGlobalInit::world_size = 1;
int CFDArcoGlobalInit::world_rank = 0;
bool CFDArcoGlobalInit::skip_history = false;
bool CFDArcoGlobalInit::cuda_enabled = false;
bool CFDArcoGlobalInit::hip_enabled = false;
bool CFDArcoGlobalInit::store_stepping = false;
fs::path CFDArcoGlobalInit::store_dir = {};
cfd_ctx_t CFDArcoGlobalInit::ctx_ = nullptr;
bool Allocator::allocator_alive = false;

namespace cfdarco {

void check(int rc, const char* what) {
    if (rc != CFD_OK) throw std::runtime_error(std::string(what) + ": " + cfd_last_error());
}

namespace {

// Freed device blocks kept for reuse, by exact size.  Every allocation of the
// operator API, and every kernel and copy that touches one, is ordered on the
// null stream, so a block released here can go to the next request of the
// same size: the stream runs the new owner's work after the old owner's.  A
// pool round trip (stream-ordered malloc and free) costs about 9 us of host
// time, as much as the elementwise kernel it feeds on an 8M-cell field.
struct FreeBlocks {
    static constexpr size_t kMaxBlocks = 64;
    static constexpr size_t kMaxBytes = size_t(8) << 30;
    std::mutex mu;
    std::multimap<size_t, void*> blocks;
    size_t held = 0;
};

FreeBlocks& free_blocks() {
    static FreeBlocks* c = new FreeBlocks;  // never destroyed: blocks may be released during exit
    return *c;
}

// returns every cached block to the pool (before an allocation the pool's
// byte cap would refuse)
void release_free_blocks(cfd_ctx_t ctx) {
    FreeBlocks& c = free_blocks();
    std::multimap<size_t, void*> all;
    {
        std::lock_guard<std::mutex> lk(c.mu);
        all.swap(c.blocks);
        c.held = 0;
    }
    for (auto& b : all) (void)cfd_free(ctx, b.second, nullptr);
}

}  // namespace

void* device_alloc(size_t bytes) {
    cfd_ctx_t ctx = CFDArcoGlobalInit::context();
    FreeBlocks& c = free_blocks();
    {
        std::lock_guard<std::mutex> lk(c.mu);
        auto it = c.blocks.find(bytes);
        if (it != c.blocks.end()) {
            void* p = it->second;
            c.blocks.erase(it);
            c.held -= bytes;
            return p;
        }
    }
    void* p = nullptr;
    int rc = cfd_alloc(ctx, int64_t(bytes), &p, nullptr);
    if (rc == CFD_E_NOMEM) {
        release_free_blocks(ctx);
        rc = cfd_alloc(ctx, int64_t(bytes), &p, nullptr);
    }
    check(rc, "device allocation");
    return p;
}

void device_free(void* p, size_t bytes) {
    cfd_ctx_t ctx = CFDArcoGlobalInit::context_or_null();
    if (!ctx) return;  // a free after finalize() is a no-op
    FreeBlocks& c = free_blocks();
    {
        std::lock_guard<std::mutex> lk(c.mu);
        if (c.blocks.size() < FreeBlocks::kMaxBlocks && c.held + bytes <= FreeBlocks::kMaxBytes) {
            c.blocks.emplace(bytes, p);
            c.held += bytes;
            return;
        }
    }
    (void)cfd_free(ctx, p, nullptr);
}

void copy_h2d(void* dst, const void* src, size_t bytes) {
    check(cfd_memcpy(CFDArcoGlobalInit::context(), dst, src, int64_t(bytes), CFD_H2D, nullptr), "host-to-device copy");
}
void copy_d2h(void* dst, const void* src, size_t bytes) {
    check(cfd_memcpy(CFDArcoGlobalInit::context(), dst, src, int64_t(bytes), CFD_D2H, nullptr), "device-to-host copy");
}
void copy_d2d(void* dst, const void* src, size_t bytes) {
    check(cfd_memcpy(CFDArcoGlobalInit::context(), dst, src, int64_t(bytes), CFD_D2D, nullptr), "device copy");
}

void fill_device(void* dst, const void* value, size_t elem_bytes, size_t n) {
    if (elem_bytes == 4) {
        float v;
        std::memcpy(&v, value, 4);
        check(cfd_ewise(CFDArcoGlobalInit::context(), CFD_EW_FILL, CFD_F32, nullptr, nullptr, nullptr, double(v), dst,
                        int64_t(n), nullptr),
              "fill");
    } else if (elem_bytes == 8) {
        // 8-byte element types (double, size_t, ptrdiff_t) go through a host vector
        std::vector<uint64_t> host(n);
        uint64_t v;
        std::memcpy(&v, value, 8);
        std::fill(host.begin(), host.end(), v);
        copy_h2d(dst, host.data(), n * 8);
    } else {
        throw std::runtime_error("fill_device: unsupported element size");
    }
}

}  // namespace cfdarco

// Page-locked host blocks for the Eigen subset's large vectors (declared weak
// in cfd_eigen_subset.hpp): field copies between the device and these run at
// DMA speed (256 MB: 4.7 ms against 15.6 ms device-to-host and 9.3 ms
// host-to-device through pageable memory).  Only once the device is enabled,
// and at most kPinnedCap bytes outstanding; beyond that, or on any failure,
// the caller falls back to ordinary memory.
namespace {
constexpr size_t kPinnedCap = size_t(8) << 30;
std::atomic<size_t> pinned_held{0};
}  // namespace

extern "C" void* cfdarco_pinned_alloc(size_t bytes) {
    if (CFDArcoGlobalInit::context_or_null() == nullptr) return nullptr;
    if (pinned_held.fetch_add(bytes) + bytes > kPinnedCap) {
        pinned_held.fetch_sub(bytes);
        return nullptr;
    }
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess || p == nullptr) {
        (void)hipGetLastError();
        pinned_held.fetch_sub(bytes);
        return nullptr;
    }
    return p;
}

extern "C" void cfdarco_pinned_free(void* p, size_t bytes) {
    if (p == nullptr) return;
    (void)hipHostFree(p);
    pinned_held.fetch_sub(bytes);
}

void CFDArcoGlobalInit::initialize(int, char**, bool skip_history_, const fs::path& store_path) {
    world_size = 1;
    world_rank = 0;
    skip_history = skip_history_;
    std::time_t t = std::time(nullptr);
    std::tm* now = std::localtime(&t);
    const std::string stamp = std::to_string(now->tm_year + 1900) + "_" + std::to_string(now->tm_mon + 1) + "_" +
                              std::to_string(now->tm_mday) + "_" + std::to_string(now->tm_hour) + "_" +
                              std::to_string(now->tm_min) + "_" + std::to_string(now->tm_sec);
    store_dir = store_path / ("run_" + stamp);
}

void CFDArcoGlobalInit::enable_cuda(int, bool use_cuda, bool use_hip) {
    if (!use_cuda && !use_hip) throw std::runtime_error{"!use_cuda and !use_hip"};
    if (ctx_ == nullptr) {
        const char* env = std::getenv("CFDARCO_DEVICE");
        const int dev = env ? std::atoi(env) : 0;
        cfdarco::check(cfd_init(dev, &ctx_), "cfd_init");
    }
    cuda_enabled = use_cuda;
    hip_enabled = use_hip;
    Allocator::allocator_alive = true;
}

cfd_ctx_t CFDArcoGlobalInit::context() {
    if (ctx_ == nullptr)
        throw std::runtime_error("device path not enabled: call CFDArcoGlobalInit::enable_cuda (-c / -hip)");
    return ctx_;
}

cfd_ctx_t CFDArcoGlobalInit::context_or_null() { return ctx_; }

// Waits for the device.  The context and its pool stay alive until the
// process ends, because Variables and meshes still own pool blocks (the
// reference resets its pool here and leaks those handles).
void CFDArcoGlobalInit::finalize() {
    if (ctx_ != nullptr) {
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) std::cerr << "finalize: " << hipGetErrorString(e) << std::endl;
    }
}

// ---------------------------------------------------------------------------
// cuda_operators.hpp
// ---------------------------------------------------------------------------
namespace {

CudaDataMatrixD ewise2(int op, const CudaDataMatrixD& a, const CudaDataMatrixD& b) {
    if (a._size != b._size) throw std::runtime_error("elementwise operands differ in size");
    CudaDataMatrixD res{a._size};
    cfdarco::check(cfd_ewise(CFDArcoGlobalInit::context(), op, CFD_F32, a.data.get(), b.data.get(), nullptr, 0.0,
                             res.data.get(), int64_t(a._size), nullptr),
                   "elementwise op");
    return res;
}

CudaDataMatrixD ewise_s(int op, const CudaDataMatrixD& a, float s) {
    CudaDataMatrixD res{a._size};
    cfdarco::check(cfd_ewise(CFDArcoGlobalInit::context(), op, CFD_F32, a.data.get(), nullptr, nullptr, double(s),
                             res.data.get(), int64_t(a._size), nullptr),
                   "elementwise op");
    return res;
}

int mask_of(bool x, bool y, bool z) { return (x ? 1 : 0) | (y ? 2 : 0) | (z ? 4 : 0); }

cfd_mesh_t dev_mesh(CudaMesh3D* mesh) {
    if (mesh == nullptr || mesh->device_mesh() == nullptr)
        throw std::runtime_error("mesh is not on the device: call CudaMesh3D::host_to_device()");
    return mesh->device_mesh();
}

}  // namespace

CudaDataMatrixD add_mtrx(const CudaDataMatrixD& a, const CudaDataMatrixD& b) { return ewise2(CFD_EW_ADD, a, b); }
CudaDataMatrixD sub_mtrx(const CudaDataMatrixD& a, const CudaDataMatrixD& b) { return ewise2(CFD_EW_SUB, a, b); }
CudaDataMatrixD mul_mtrx(const CudaDataMatrixD& a, const CudaDataMatrixD& b) { return ewise2(CFD_EW_MUL, a, b); }
CudaDataMatrixD mul_mtrx(const CudaDataMatrixD& a, float b) { return ewise_s(CFD_EW_MULS, a, b); }
CudaDataMatrixD div_mtrx(const CudaDataMatrixD& a, const CudaDataMatrixD& b) { return ewise2(CFD_EW_DIV, a, b); }
CudaDataMatrixD neg_mtrx(const CudaDataMatrixD& a) { return ewise_s(CFD_EW_NEG, a, 0.f); }
CudaDataMatrixD div_const(const CudaDataMatrixD& a, float b) { return ewise_s(CFD_EW_DIVS, a, b); }
CudaDataMatrixD exp_mtrx(const CudaDataMatrixD& a) { return ewise_s(CFD_EW_EXP, a, 0.f); }
CudaDataMatrixD abs_mtrx(const CudaDataMatrixD& a) { return ewise_s(CFD_EW_ABS, a, 0.f); }

CudaDataMatrixD eval_grad(CudaMesh3D* mesh, const CudaDataMatrixD& a, bool x, bool y, bool z) {
    CudaDataMatrixD res{a._size};
    cfdarco::check(cfd_grad(dev_mesh(mesh), a.data.get(), res.data.get(), mask_of(x, y, z), -1, nullptr), "eval_grad");
    return res;
}

CudaDataMatrixD eval_grad2(CudaMesh3D* mesh, const CudaDataMatrixD& a, bool x, bool y, bool z) {
    CudaDataMatrixD res{a._size};
    cfdarco::check(cfd_lapl(dev_mesh(mesh), a.data.get(), res.data.get(), mask_of(x, y, z), nullptr), "eval_grad2");
    return res;
}

CudaDataMatrixD eval_stab(CudaMesh3D* mesh, const CudaDataMatrixD& a, bool x, bool y, bool z) {
    CudaDataMatrixD res{a._size};
    cfdarco::check(cfd_stab(dev_mesh(mesh), a.data.get(), res.data.get(), mask_of(x, y, z), nullptr), "eval_stab");
    return res;
}

static float cfl_common(float dl, float gamma, const CudaDataMatrixD& p, const CudaDataMatrixD& rho,
                        const CudaDataMatrixD& u, const float* v, const float* w) {
    double out = 0;
    cfdarco::check(cfd_cfl_min(CFDArcoGlobalInit::context(), CFD_F32, u.data.get(), v, w, p.data.get(),
                               rho.data.get(), int64_t(p._size), double(dl), 1.0, double(gamma), &out, nullptr),
                   "cfl_cu");
    return float(out);
}

float cfl_cu(float dl, float gamma, const CudaDataMatrixD& p, const CudaDataMatrixD& rho, const CudaDataMatrixD& u,
             const CudaDataMatrixD& v, const CudaDataMatrixD& w) {
    return cfl_common(dl, gamma, p, rho, u, v.data.get(), w.data.get());
}

float cfl_cu(float dl, float gamma, const CudaDataMatrixD& p, const CudaDataMatrixD& rho, const CudaDataMatrixD& u) {
    return cfl_common(dl, gamma, p, rho, u, nullptr, nullptr);
}

void sync_device() {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) throw std::runtime_error(std::string("sync_device: ") + hipGetErrorString(e));
}
