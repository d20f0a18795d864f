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

/* ---- A solver without Python, from an export folder ----
 *
 * CudaSolver.export(folder) (maxichem_amd/_export.py) writes the call's code object, tables and meta.txt; these
 * entry points load and run them with the Python call's kernel arguments, so a call gives the same bits. They
 * replace the reference's C++ caller, maxichem::Solver (cpp/maxichem.hpp:44-92, cpp/maxichem.cpp:200-398); the C++
 * class maxichem::Solver of include/maxichem.hpp wraps them for libtorch callers. Calls of one solver share its
 * counts and work buffers: they run one at a time (one stream, or synchronized between streams). */

typedef struct mxc_solver mxc_solver; /* an export loaded on one device */

/* mxc_solver_call's code when the call refused cells (bad input, or T outside the solver's range): the message is
 * mxc_last_error's, the outputs and counts are written. */
#define MXC_REFUSED 2

/* Loads the export in folder on device. Every check of meta.txt and the table files (format, argument lists of the
 * kernels, table sizes) comes before any device work; the device must be the code object's architecture. Replaces
 * Solver::Solver (cpp/maxichem.cpp:200-263). */
int mxc_solver_open(const char* folder, int device, mxc_solver** solver);

/* Waits for the device, frees the solver's memory and unloads its code. Replaces Solver::~Solver
 * (cpp/maxichem.cpp:265-269). */
int mxc_solver_close(mxc_solver* solver);

/* Species, photolysis branches, elements and count slots. Replaces Solver::ns, nj (cpp/maxichem.hpp:65-66). */
int mxc_solver_sizes(const mxc_solver* solver, int* ns, int* nj, int* nel, int* count_slots);

/* The call's default rtol, atol and max_steps (the Python call's defaults, maxichem/_constants.py:12-25). Replaces
 * kDefaultRtol, kDefaultMaxSteps and Solver::default_atol (cpp/maxichem.hpp:33-34,69). */
int mxc_solver_defaults(const mxc_solver* solver, double* rtol, double* atol, int* max_steps);

/* A name: list 0 species (x's column order), 1 elements (the atoms table's columns), 2 count slots (the queue, the
 * ten counts of the Python solver's COUNTS, the recovery's queue); NULL out of range. Replaces Solver::species and
 * elements (cpp/maxichem.hpp:67-68). */
const char* mxc_solver_name(const mxc_solver* solver, int list, int index);

/* One call on stream (a hipStream_t; NULL: the null stream), all arrays on the solver's device:
 * x [cells, ns], T, P, dt, h [cells] (float64); J [cells, nj] (float64; NULL without photolysis, required with it);
 * out: x_out [cells, ns], h_out [cells] (the next step sizes), done [cells] (int32, 1 when the call finished and the
 * result passed the guard), steps [cells] (accepted steps, float64), flag [cells] (int32 0/1/2 as the Python launch's
 * flag; NULL: not returned). Returns 0, MXC_REFUSED, or 1 for a form error (before any launch) or a runtime failure.
 * Synchronizes the stream once (the counts). Replaces Solver::operator() and Solver::refuse
 * (cpp/maxichem.cpp:271-398) and CudaSolver.launch (maxichem/kincuda.py:230-320). */
int mxc_solver_call(mxc_solver* solver, void* stream, int cells, const double* x, const double* T, const double* P,
                    const double* dt, const double* h, const double* J, double rtol, double atol, int max_steps,
                    double* x_out, double* h_out, int* done, double* steps, int* flag);

/* The most cells one mxc_solver_call_static call takes. */
#define MXC_STATIC_CELLS 262144

/* mxc_solver_call without any device allocation: the call's work buffers (each cell's status, guard flag, recovery
 * list entry and tries: 20 bytes per cell) are the library's own device array of MXC_STATIC_CELLS cells, one copy
 * per device, whose address mxc_solver_open looks up; mxc_solver_call allocates them per solver instead (hipMalloc
 * when the batch grows). Same arguments, results, counts, refusals and return codes as mxc_solver_call, bit for bit;
 * more than MXC_STATIC_CELLS cells are refused (return 1, mxc_last_error says so) before any device work. Every
 * call of this entry point on one device shares that array, so such calls must not overlap on a device (one stream,
 * or synchronized between streams), whichever solver they use. No counterpart in the reference: it is for callers
 * that cannot allocate device memory in their call path (tight serving loops, allocation-free harnesses). */
int mxc_solver_call_static(mxc_solver* solver, void* stream, int cells, const double* x, const double* T,
                           const double* P, const double* dt, const double* h, const double* J, double rtol,
                           double atol, int max_steps, double* x_out, double* h_out, int* done, double* steps,
                           int* flag);

/* Host copies of the last call's counts (int [count_slots]) and their sum since mxc_solver_open (long long
 * [count_slots]); either may be NULL. Replaces Solver::count and total (cpp/maxichem.hpp:72-73). */
int mxc_solver_counts(const mxc_solver* solver, int* count, long long* total);

/* A host copy of the atoms table, [ns][nel] float64 (network species order). Replaces Solver::atoms
 * (cpp/maxichem.hpp:71). */
int mxc_solver_atoms(const mxc_solver* solver, double* atoms);

/* The parameter names mxc_solver_call passes to kernel (mxc_call or mxc_recover), in order, separated by
 * spaces, into names (size bytes with the NUL). An export must list the same. Replaces the reference caller's
 * kKernelArguments (cpp/maxichem.cpp:27-49). */
int mxc_kernel_arguments(const char* kernel, char* names, size_t size);

#ifdef __cplusplus
}
#endif
#endif
