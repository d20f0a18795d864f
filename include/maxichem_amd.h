/* maxichem_amd: the C ABI between the Python host code and the HIP kernels of the batched Rodas3 chemistry step.
 *
 * Plain pointers, sizes and integer return codes (0 = success; the message of a failure is read with
 * mxc_last_error). No exceptions cross this boundary and no PyTorch types appear in it. Device pointers are the
 * data_ptr() of PyTorch tensors, a stream is torch.cuda.current_stream().cuda_stream (a hipStream_t).
 *
 * Each entry point replaces one function of the reference's driver binding, maxichem/_cuda.py, which binds the CUDA
 * driver and NVRTC through ctypes; the cited lines are that file's.
 */
#ifndef MAXICHEM_AMD_H
#define MAXICHEM_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mxc_module mxc_module; /* a loaded code object and its device */

/* The library's ABI version; the Python binding refuses another. */
int mxc_version(void);

/* The hiprtc version that mxc_compile uses (part of the key of a cache of compiled code). Part of the NVRTC binding
 * the reference loads (_cuda.py:22-41). */
int mxc_compiler_version(int* major, int* minor);

/* Compiles HIP source with hiprtc for the options given (e.g. --offload-arch=gfx950) into a code object, returned
 * in a buffer the caller frees with mxc_free. The compiler log, if any, goes to log (NUL-terminated, truncated to
 * log_size). Replaces compile_cubin (_cuda.py:52-79), which compiles with NVRTC. */
int mxc_compile(const char* source, const char* name, const char* const* options, int n_options, void** code,
                size_t* code_size, char* log, size_t log_size);

/* Frees a buffer that mxc_compile returned. */
void mxc_free(void* p);

/* Loads a code object on device and returns the module. Replaces load_module (_cuda.py:89-105). */
int mxc_module_load(const void* code, size_t code_size, int device, mxc_module** module);

/* The kernel of that name in the module (a hipFunction_t). Part of load_module (_cuda.py:89-105). */
int mxc_module_function(mxc_module* module, const char* name, void** function);

/* Unloads a module (its device is made current for the call). Replaces unload_modules (_cuda.py:108-113). */
int mxc_module_unload(mxc_module* module);

/* Launches a kernel of a loaded module on stream: grid x block (1-D), dynamic LDS bytes, args = one pointer to
 * each argument's value in the kernel's parameter order. Replaces launch_kernel (_cuda.py:82-86). */
int mxc_launch(void* function, unsigned grid, unsigned block, unsigned lds_bytes, void* stream, void** args);

/* Blocks of this size and dynamic LDS one compute unit holds at once. Replaces blocks_per_multiprocessor
 * (_cuda.py:128-136). */
int mxc_occupancy(void* function, int block, size_t lds_bytes, int* blocks);

/* A kernel attribute: 0 registers per lane, 1 private (scratch) bytes per lane, 2 static LDS bytes, 3 the largest
 * block. Replaces function_attribute (_cuda.py:160-164). */
int mxc_function_attribute(void* function, int attribute, int* value);

/* The device's compute units, LDS bytes a block may take, and its wavefront size. Replaces the device properties
 * the reference reads through torch (kincuda.py:139-140). */
int mxc_device_info(int device, int* compute_units, int* lds_per_block, int* wave_size);

/* The message of this thread's last failure (empty when none). Replaces check_driver and check_nvrtc's messages
 * (_cuda.py:44-49). */
const char* mxc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
