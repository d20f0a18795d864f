// The C ABI of include/maxichem_amd.h over hiprtc and the HIP module API (host code only; the kernels are HIP
// source compiled at run time by mxc_compile).
#include "maxichem_amd.h"

#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

struct mxc_module {
  hipModule_t module;
  int device;
};

namespace {
thread_local std::string last_error;
}  // namespace

namespace mxc_detail {  // shared with solver.cpp
int fail(const std::string& what) {
  last_error = what;
  return 1;
}
void clear_error() { last_error.clear(); }
}  // namespace mxc_detail

namespace {

using mxc_detail::fail;

int check(hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  return fail(std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")");
}

int check_rtc(hiprtcResult e, const char* what) {
  if (e == HIPRTC_SUCCESS) return 0;
  return fail(std::string(what) + ": " + hiprtcGetErrorString(e));
}

}  // namespace

extern "C" {

int mxc_version(void) { return 3; }

int mxc_compiler_version(int* major, int* minor) {
  last_error.clear();
  return check_rtc(hiprtcVersion(major, minor), "hiprtcVersion");
}

const char* mxc_last_error(void) { return last_error.c_str(); }

void mxc_free(void* p) { std::free(p); }

int mxc_compile(const char* source, const char* name, const char* const* options, int n_options, void** code,
                size_t* code_size, char* log, size_t log_size) {
  last_error.clear();
  if (log && log_size) log[0] = '\0';
  hiprtcProgram prog;
  if (check_rtc(hiprtcCreateProgram(&prog, source, name, 0, nullptr, nullptr), "hiprtcCreateProgram")) return 1;
  const hiprtcResult built = hiprtcCompileProgram(prog, n_options, const_cast<const char**>(options));
  size_t n = 0;
  if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n > 1 && log && log_size) {
    std::vector<char> text(n + 1, '\0');
    if (hiprtcGetProgramLog(prog, text.data()) == HIPRTC_SUCCESS) {
      std::snprintf(log, log_size, "%s", text.data());
    }
  }
  if (built != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return check_rtc(built, "hiprtcCompileProgram (see the compiler log)");
  }
  size_t size = 0;
  if (check_rtc(hiprtcGetCodeSize(prog, &size), "hiprtcGetCodeSize")) {
    hiprtcDestroyProgram(&prog);
    return 1;
  }
  void* buffer = std::malloc(size);
  if (!buffer) {
    hiprtcDestroyProgram(&prog);
    return fail("mxc_compile: out of host memory");
  }
  const int bad = check_rtc(hiprtcGetCode(prog, static_cast<char*>(buffer)), "hiprtcGetCode");
  hiprtcDestroyProgram(&prog);
  if (bad) {
    std::free(buffer);
    return 1;
  }
  *code = buffer;
  *code_size = size;
  return 0;
}

int mxc_module_load(const void* code, size_t code_size, int device, mxc_module** module) {
  last_error.clear();
  (void)code_size;
  if (check(hipSetDevice(device), "hipSetDevice")) return 1;
  hipModule_t m;
  if (check(hipModuleLoadData(&m, code), "hipModuleLoadData")) return 1;
  *module = new mxc_module{m, device};
  return 0;
}

int mxc_module_function(mxc_module* module, const char* name, void** function) {
  last_error.clear();
  hipFunction_t f;
  if (check(hipModuleGetFunction(&f, module->module, name), (std::string("hipModuleGetFunction ") + name).c_str()))
    return 1;
  *function = reinterpret_cast<void*>(f);
  return 0;
}

int mxc_module_unload(mxc_module* module) {
  last_error.clear();
  if (!module) return 0;
  int bad = check(hipSetDevice(module->device), "hipSetDevice");
  if (!bad) bad = check(hipModuleUnload(module->module), "hipModuleUnload");
  delete module;
  return bad;
}

int mxc_launch(void* function, unsigned grid, unsigned block, unsigned lds_bytes, void* stream, void** args) {
  last_error.clear();
  return check(hipModuleLaunchKernel(reinterpret_cast<hipFunction_t>(function), grid, 1, 1, block, 1, 1, lds_bytes,
                                     reinterpret_cast<hipStream_t>(stream), args, nullptr),
               "hipModuleLaunchKernel");
}

int mxc_occupancy(void* function, int block, size_t lds_bytes, int* blocks) {
  last_error.clear();
  return check(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(blocks, reinterpret_cast<hipFunction_t>(function),
                                                                  block, lds_bytes),
               "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
}

int mxc_function_attribute(void* function, int attribute, int* value) {
  last_error.clear();
  static const hipFunction_attribute table[] = {HIP_FUNC_ATTRIBUTE_NUM_REGS, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES,
                                                HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES,
                                                HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK};
  if (attribute < 0 || attribute > 3) return fail("mxc_function_attribute: attribute 0-3 expected");
  return check(hipFuncGetAttribute(value, table[attribute], reinterpret_cast<hipFunction_t>(function)),
               "hipFuncGetAttribute");
}

int mxc_device_info(int device, int* compute_units, int* lds_per_block, int* wave_size) {
  last_error.clear();
  if (check(hipDeviceGetAttribute(compute_units, hipDeviceAttributeMultiprocessorCount, device), "compute units"))
    return 1;
  if (check(hipDeviceGetAttribute(lds_per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, device), "LDS per block"))
    return 1;
  return check(hipDeviceGetAttribute(wave_size, hipDeviceAttributeWarpSize, device), "wavefront size");
}

}  // extern "C"
