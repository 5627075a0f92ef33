// Library core: error reporting, device query, the seeded-field generator and the hipRTC
// compile/load/launch path used by the fused executor.
#include <hip/hiprtc.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "sp_common.h"

namespace sp {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

}  // namespace sp

using namespace sp;

extern "C" int sp_abi_version(void) { return SP_ABI_VERSION; }

extern "C" const char* sp_last_error(void) { return g_last_error.c_str(); }

extern "C" int sp_device_info(char* name, int name_len, int* n_cu, int* n_xcd) {
    int dev = 0;
    SP_HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    SP_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    if (name && name_len > 0) {
        std::strncpy(name, prop.gcnArchName, (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    if (n_cu) *n_cu = prop.multiProcessorCount;
    // gfx950 (MI355X) has 8 accelerator complex dies of 32 CUs each.
    if (n_xcd) *n_xcd = (std::strncmp(prop.gcnArchName, "gfx95", 5) == 0 ||
                         std::strncmp(prop.gcnArchName, "gfx94", 5) == 0) ? 8 : 1;
    return SP_OK;
}

// ---------------------------------------------------------------------------- generator
namespace {

struct FourierModes {
    int32_t k[8][3];
    double a[8];
    double phi[8];
};

__device__ __forceinline__ uint64_t splitmix_at(uint64_t seed, uint64_t lin) {
    uint64_t z = seed + (lin + 1ull) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256) k_gen_fourier(float* __restrict__ out, int64_t n0,
                                                     int64_t n1, int64_t n2, int64_t z0,
                                                     int64_t nz, FourierModes m,
                                                     uint64_t seed) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t y = blockIdx.y;
    const int64_t zl = blockIdx.z;
    if (x >= n2) return;
    const int64_t z = z0 + zl;
    const double two_pi = 6.283185307179586476925286766559;
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        double ph = two_pi * ((double)m.k[q][0] * (double)z / (double)n0 +
                              (double)m.k[q][1] * (double)y / (double)n1 +
                              (double)m.k[q][2] * (double)x / (double)n2) + m.phi[q];
        acc += m.a[q] * cos(ph);
    }
    const uint64_t lin = ((uint64_t)z * (uint64_t)n1 + (uint64_t)y) * (uint64_t)n2 + (uint64_t)x;
    const double u = (double)(splitmix_at(seed, lin) >> 11) * (1.0 / 9007199254740992.0);
    acc += (2.0 * u - 1.0) * 1e-3;
    out[(zl * n1 + y) * n2 + x] = (float)acc;
}

}  // namespace

extern "C" int sp_gen_fourier(sp_tensor* out, const int64_t n[3], int64_t z0, const int32_t* k,
                              const double* a, const double* phi, uint64_t noise_seed,
                              void* stream) {
    SP_REQUIRE(out && out->ptr && out->dtype == SP_F32 && out->ndim == 3, "gen: out must be f32 3-D");
    SP_REQUIRE(is_contiguous(out), "gen: out must be contiguous");
    SP_REQUIRE(out->shape[1] == n[1] && out->shape[2] == n[2], "gen: plane shape mismatch");
    SP_REQUIRE(z0 >= 0 && z0 + out->shape[0] <= n[0], "gen: plane range outside the field");
    SP_REQUIRE(n[1] <= 65535 && out->shape[0] <= 65535, "gen: grid too large");
    FourierModes m;
    for (int q = 0; q < 8; ++q) {
        for (int d = 0; d < 3; ++d) m.k[q][d] = k[3 * q + d];
        m.a[q] = a[q];
        m.phi[q] = phi[q];
    }
    if (out->shape[0] == 0) return SP_OK;
    dim3 grid(cdiv(n[2], 256), (unsigned)n[1], (unsigned)out->shape[0]);
    hipLaunchKernelGGL(k_gen_fourier, grid, dim3(256), 0, as_stream(stream), (float*)out->ptr,
                       n[0], n[1], n[2], z0, out->shape[0], m, noise_seed);
    return launch_status("sp_gen_fourier");
}

// ---------------------------------------------------------------------------- hipRTC
struct sp_module {
    hipModule_t module = nullptr;
    hipFunction_t fn = nullptr;
};

extern "C" int sp_rtc_compile(const char* source, const char* name, const char* options,
                              void** code, size_t* code_size, char* log, size_t log_len) {
    SP_REQUIRE(source && code && code_size, "rtc: null argument");
    *code = nullptr;
    *code_size = 0;
    std::vector<std::string> opts;
    if (options) {
        std::string s(options);
        size_t pos = 0;
        while (pos <= s.size()) {
            size_t nl = s.find('\n', pos);
            if (nl == std::string::npos) nl = s.size();
            if (nl > pos) opts.push_back(s.substr(pos, nl - pos));
            pos = nl + 1;
        }
    }
    std::vector<const char*> optp;
    for (auto& o : opts) optp.push_back(o.c_str());
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, source, name ? name : "fused.hip", 0, nullptr, nullptr) !=
        HIPRTC_SUCCESS)
        return fail(SP_ECOMPILE, "hiprtcCreateProgram failed");
    hiprtcResult r = hiprtcCompileProgram(prog, (int)optp.size(), optp.data());
    size_t lsz = 0;
    hiprtcGetProgramLogSize(prog, &lsz);
    std::string lg(lsz + 1, '\0');
    if (lsz) hiprtcGetProgramLog(prog, &lg[0]);
    if (log && log_len) {
        std::strncpy(log, lg.c_str(), log_len - 1);
        log[log_len - 1] = 0;
    }
    if (r != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return fail(SP_ECOMPILE, std::string("hipRTC compile failed: ") + lg.substr(0, 4000));
    }
    size_t csz = 0;
    hiprtcGetCodeSize(prog, &csz);
    void* buf = std::malloc(csz);
    if (!buf) {
        hiprtcDestroyProgram(&prog);
        return fail(SP_ECOMPILE, "rtc: out of host memory");
    }
    hiprtcGetCode(prog, (char*)buf);
    hiprtcDestroyProgram(&prog);
    *code = buf;
    *code_size = csz;
    return SP_OK;
}

extern "C" void sp_free_code(void* code) { std::free(code); }

extern "C" int sp_module_load(const void* code, size_t code_size, const char* kernel_name,
                              sp_module** mod) {
    SP_REQUIRE(code && code_size && kernel_name && mod, "module_load: null argument");
    sp_module* m = new sp_module();
    hipError_t e = hipModuleLoadData(&m->module, code);
    if (e != hipSuccess) {
        delete m;
        return fail(SP_EHIP, std::string("hipModuleLoadData: ") + hipGetErrorString(e));
    }
    e = hipModuleGetFunction(&m->fn, m->module, kernel_name);
    if (e != hipSuccess) {
        hipModuleUnload(m->module);
        delete m;
        return fail(SP_EHIP, std::string("hipModuleGetFunction: ") + hipGetErrorString(e));
    }
    *mod = m;
    return SP_OK;
}

extern "C" int sp_module_launch(sp_module* mod, const unsigned grid[3], const unsigned block[3],
                                unsigned shared_bytes, void* stream, const void* args,
                                size_t args_size) {
    SP_REQUIRE(mod && mod->fn, "module_launch: module not loaded");
    SP_REQUIRE(grid[0] && grid[1] && grid[2], "module_launch: empty grid");
    // The argument block is copied into an 8-byte aligned local buffer.
    std::vector<uint64_t> buf((args_size + 7) / 8 + 1, 0);
    if (args_size) std::memcpy(buf.data(), args, args_size);
    size_t sz = args_size;
    void* config[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, buf.data(), HIP_LAUNCH_PARAM_BUFFER_SIZE,
                      &sz, HIP_LAUNCH_PARAM_END};
    SP_HIP_CHECK(hipModuleLaunchKernel(mod->fn, grid[0], grid[1], grid[2], block[0], block[1],
                                       block[2], shared_bytes, as_stream(stream), nullptr,
                                       config));
    return SP_OK;
}

extern "C" int sp_module_attributes(sp_module* mod, int* vgprs, int* lds_bytes,
                                    int* spill_bytes, int* max_threads) {
    SP_REQUIRE(mod && mod->fn, "module_attributes: module not loaded");
    int v = 0;
    if (vgprs) {
        SP_HIP_CHECK(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_NUM_REGS, mod->fn));
        *vgprs = v;
    }
    if (lds_bytes) {
        SP_HIP_CHECK(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, mod->fn));
        *lds_bytes = v;
    }
    if (spill_bytes) {
        SP_HIP_CHECK(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, mod->fn));
        *spill_bytes = v;
    }
    if (max_threads) {
        SP_HIP_CHECK(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, mod->fn));
        *max_threads = v;
    }
    return SP_OK;
}

extern "C" int sp_module_occupancy(sp_module* mod, int block_threads, size_t dyn_shared,
                                   int* blocks_per_cu) {
    SP_REQUIRE(mod && mod->fn && blocks_per_cu && block_threads > 0, "module_occupancy: bad arguments");
    int n = 0;
    SP_HIP_CHECK(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&n, mod->fn, block_threads, dyn_shared));
    *blocks_per_cu = n;
    return SP_OK;
}

extern "C" void sp_module_unload(sp_module* mod) {
    if (!mod) return;
    if (mod->module) hipModuleUnload(mod->module);
    delete mod;
}
