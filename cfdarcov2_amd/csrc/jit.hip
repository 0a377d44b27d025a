// hipRTC programs: the runtime-compiled fused kernels of the operator API.
// Sources are compiled for gfx950 with the same arithmetic flags as the rest
// of the library (no FMA contraction, no packed FP32), and code objects are
// cached by a 64-bit FNV-1a hash of (source, kernel name, options): in memory
// per context and on disk ($CFDARCO_JIT_CACHE, default ~/.cache/cfdarco/jit).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

#include "common.hpp"

struct cfd_program_s {
    cfd_ctx_t ctx = nullptr;
    hipModule_t module = nullptr;
    hipFunction_t fn = nullptr;
    std::string name;
    uint64_t key = 0;
};

namespace cfd {

static const char* kJitOptions[] = {
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
};

static uint64_t fnv1a(uint64_t h, const std::string& s) {
    for (unsigned char c : s) {
        h ^= c;
        h *= 1099511628211ull;
    }
    return h;
}

static uint64_t program_key(const std::string& src, const std::string& name) {
    uint64_t h = 1469598103934665603ull;
    h = fnv1a(h, src);
    h = fnv1a(h, "\x1f" + name);
    for (const char* o : kJitOptions) h = fnv1a(h, std::string("\x1f") + o);
    return h;
}

static std::string cache_dir() {
    const char* env = std::getenv("CFDARCO_JIT_CACHE");
    if (env && *env) return env;
    const char* home = std::getenv("HOME");
    return std::string(home ? home : "/tmp") + "/.cache/cfdarco/jit";
}

static void mkdirs(const std::string& path) {
    std::string cur;
    std::stringstream ss(path);
    std::string part;
    if (!path.empty() && path[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        ::mkdir(cur.c_str(), 0755);
    }
}

// Compiles `src` to a gfx950 code object; throws CFD_E_COMPILE with the log.
static std::vector<char> compile_to_code(const std::string& src, const std::string& name) {
    hiprtcProgram prog;
    hiprtcResult r = hiprtcCreateProgram(&prog, src.c_str(), (name + ".hip").c_str(), 0, nullptr, nullptr);
    if (r != HIPRTC_SUCCESS) throw Error(CFD_E_COMPILE, std::string("hiprtcCreateProgram: ") + hiprtcGetErrorString(r));
    const int nopt = int(sizeof(kJitOptions) / sizeof(kJitOptions[0]));
    r = hiprtcCompileProgram(prog, nopt, kJitOptions);
    if (r != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        throw Error(CFD_E_COMPILE, "hipRTC compilation of " + name + " failed:\n" + log);
    }
    size_t sz = 0;
    hiprtcGetCodeSize(prog, &sz);
    std::vector<char> code(sz);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return code;
}

// code object for (src, name), from the disk cache or freshly compiled
static std::vector<char> code_for(const std::string& src, const std::string& name, uint64_t key) {
    char hex[17];
    std::snprintf(hex, sizeof(hex), "%016llx", (unsigned long long)key);
    const std::string dir = cache_dir();
    const std::string path = dir + "/" + hex + ".co";
    {
        std::ifstream in(path, std::ios::binary);
        if (in) {
            std::vector<char> code((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
            if (!code.empty()) return code;
        }
    }
    std::vector<char> code = compile_to_code(src, name);
    mkdirs(dir);
    const std::string tmp = path + ".tmp" + std::to_string(::getpid());
    {
        std::ofstream out(tmp, std::ios::binary);
        out.write(code.data(), std::streamsize(code.size()));
    }
    std::rename(tmp.c_str(), path.c_str());
    return code;
}

struct ModuleCache {
    std::mutex mu;
    std::map<std::pair<cfd_ctx_t, uint64_t>, hipModule_t> mods;
};
static ModuleCache& module_cache() {
    static ModuleCache c;
    return c;
}

}  // namespace cfd

using namespace cfd;

extern "C" {

int cfd_program_compile(cfd_ctx_t ctx, const char* src, const char* kernel_name, cfd_program_t* out) {
    return guarded([&] {
        CFD_REQUIRE(ctx && src && kernel_name && out, "cfd_program_compile: null argument");
        const std::string s(src), name(kernel_name);
        const uint64_t key = program_key(s, name);
        CFD_HIP(hipSetDevice(ctx->device));
        auto& mc = module_cache();
        hipModule_t mod = nullptr;
        {
            std::lock_guard<std::mutex> lk(mc.mu);
            auto it = mc.mods.find({ctx, key});
            if (it != mc.mods.end()) mod = it->second;
        }
        if (!mod) {
            std::vector<char> code = code_for(s, name, key);
            CFD_HIP(hipModuleLoadData(&mod, code.data()));
            std::lock_guard<std::mutex> lk(mc.mu);
            mc.mods[{ctx, key}] = mod;
        }
        auto p = std::make_unique<cfd_program_s>();
        p->ctx = ctx;
        p->module = mod;
        p->name = name;
        p->key = key;
        CFD_HIP(hipModuleGetFunction(&p->fn, mod, name.c_str()));
        *out = p.release();
    });
}

int cfd_program_compile_only(const char* src, const char* kernel_name, int64_t* code_bytes) {
    return guarded([&] {
        CFD_REQUIRE(src && kernel_name, "cfd_program_compile_only: null argument");
        std::vector<char> code = compile_to_code(src, kernel_name);
        if (code_bytes) *code_bytes = int64_t(code.size());
    });
}

int cfd_program_launch(cfd_program_t prog, const uint32_t grid[3], const uint32_t block[3], void** args,
                       void* stream) {
    return guarded([&] {
        CFD_REQUIRE(prog && grid && block, "cfd_program_launch: null argument");
        CFD_REQUIRE(grid[0] && grid[1] && grid[2] && block[0] && block[1] && block[2],
                    "cfd_program_launch: empty launch");
        CFD_REQUIRE(uint64_t(block[0]) * block[1] * block[2] <= 1024, "cfd_program_launch: block too large");
        CFD_HIP(hipModuleLaunchKernel(prog->fn, grid[0], grid[1], grid[2], block[0], block[1], block[2], 0,
                                      as_stream(stream), args, nullptr));
    });
}

int cfd_program_destroy(cfd_program_t prog) {
    return guarded([&] { delete prog; });  // modules stay cached for the process lifetime
}

}  // extern "C"
