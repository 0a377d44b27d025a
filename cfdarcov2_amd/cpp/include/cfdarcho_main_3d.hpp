// Process-wide state of the operator API (include/cfdarcho_main_3d.hpp:33-79
// of the reference): command-line flags, the output directory and the device
// context.  One process drives one GPU; world size and rank are 1 and 0.
#pragma once

#include <filesystem>
#include <string>
#include <vector>

#include "../../../include/cfdarco.h"

namespace fs = std::filesystem;

enum class DistributionStrategy { Linear, Cluster };

class CFDArcoGlobalInit {
public:
    static void initialize(int argc, char** argv, bool skip_history_, const fs::path& store_path = "./dumps");
    static void finalize();
    static inline int get_rank() { return world_rank; }
    static inline int get_size() { return world_size; }
    // Selects the MI355X device path.  `use_cuda` (-c) and `use_hip` (-hip)
    // both run on the HIP library; the flag chosen is reported through
    // cuda_enabled / hip_enabled as in the reference.
    static void enable_cuda(int cuda_ranks, bool use_cuda, bool use_hip);
    static inline void enable_cuda() { enable_cuda(0, true, false); }
    // library context of the device path (throws if the device path is off)
    static cfd_ctx_t context();
    static cfd_ctx_t context_or_null();

    static bool cuda_enabled;
    static bool hip_enabled;
    static bool skip_history;
    static bool store_stepping;
    static fs::path store_dir;

    CFDArcoGlobalInit(CFDArcoGlobalInit& other) = delete;
    void operator=(const CFDArcoGlobalInit&) = delete;

private:
    ~CFDArcoGlobalInit() = default;
    static int world_size;
    static int world_rank;
    static cfd_ctx_t ctx_;
};
