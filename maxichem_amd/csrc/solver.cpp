// A solver without Python: loads the folder a Python CudaSolver exported (code object, tables, meta.txt) and runs
// the call kernel (which guards each cell's result) and the recovery on a HIP stream with the arguments the Python
// call passes, so a call gives the same bits. The C ABI of include/maxichem_amd.h (mxc_solver_*); the reference's counterpart is the C++
// caller cpp/maxichem.cpp:200-398, whose checks of the export come before any device work, as here.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "maxichem_amd.h"

namespace mxc_detail {
int fail(const std::string& what);  // sets the thread's last error (maxichem_hip.cpp); returns 1
void clear_error();
}  // namespace mxc_detail

using mxc_detail::fail;

namespace {

constexpr int kExportFormat = 3;  // 2: the recovery workgroup's size (rec_threads); 3: the guard in the call kernel
constexpr int kWave = 64;  // lanes of a wavefront: one wavefront steps one cell
// The kernels' Net struct, one device pointer per table in this order (kernels/common.hip).
const char* const kNetFields[] = {"perm",  "rc",   "rf",  "nasa",  "rxe",   "svw",
                                  "rowp",  "diagw", "leadw", "chw", "atoms", "rord"};
constexpr int kNetCount = sizeof(kNetFields) / sizeof(kNetFields[0]);

// The position of a Net field by name (the guard's atoms table is read on the host as well).
int net_field(const char* name) {
  for (int i = 0; i < kNetCount; ++i)
    if (std::string(kNetFields[i]) == name) return i;
  return -1;
}
// Each kernel's parameters in order: what mxc_solver_call passes. An export listing another order is refused.
struct KernelArguments {
  const char* kernel;
  std::vector<std::string> names;
};
const KernelArguments kKernels[] = {
    {"mxc_call", {"net", "x0", "Ta", "Pa", "dta", "hio", "xout", "doneo", "noko", "status", "redo", "list", "count",
                  "tries", "order", "Jin", "B", "rtol", "atol", "max_steps", "xprop", "rscratch", "fp64"}},
    {"mxc_recover", {"net", "x0", "Ta", "Pa", "dta", "hin", "xout", "hio", "doneo", "noko", "redo", "list", "count",
                     "tries", "Jin", "rtol", "atol", "max_steps", "rscratch"}},
};

int check(hipError_t e, const std::string& what) {
  if (e == hipSuccess) return 0;
  return fail(what + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")");
}

bool read_file(const std::string& path, std::string& out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  out = ss.str();
  return true;
}

// %g as Python's format(v, "g") prints it ("nan" for every NaN)
std::string g(double v) {
  if (std::isnan(v)) return "nan";
  char b[64];
  std::snprintf(b, sizeof b, "%g", v);
  return b;
}

struct Table {
  std::string name, dtype;
  std::vector<long long> shape;
  size_t bytes = 0;
};

size_t dtype_size(const std::string& d) {
  if (d == "float64" || d == "int64") return 8;
  if (d == "int32") return 4;
  return 0;
}

// The work buffers of mxc_solver_call_static, per cell: tries (double), then status, redo and list (int), for at
// most MXC_STATIC_CELLS cells. One copy per device, held until the process ends; in host code the name stands for
// no device address: its address on a device comes from hipGetSymbolAddress with that device current.
constexpr size_t kStaticBytes = static_cast<size_t>(MXC_STATIC_CELLS) * (sizeof(double) + 3 * sizeof(int));
__device__ __align__(256) unsigned char scratch[kStaticBytes];

}  // namespace

struct mxc_solver {
  int device = 0;
  std::string arch;  // the code object's architecture
  hipModule_t module = nullptr;
  hipFunction_t call = nullptr, recover = nullptr;
  int ns = 0, nj = 0, nr = 0, nel = 0, cpw = 1, lds_call = 0, lds_rec = 0, rates_global = 0;
  int rec_threads = 0;  // the recovery workgroup: wavefront 0 steps the cell, all of them share its dense LU
  int rec_steps = 0, max_steps = 0, max_blocks = 0, rec_blocks = 0;
  int bad_slot = -1, cold_slot = -1, hot_slot = -1;
  double rtol = 0.0, atol = 0.0, t_min = 0.0, t_max = 0.0, x_sum_tol = 0.0;
  std::vector<std::string> species, elements, counts;
  std::vector<void*> tables;  // device copies, in the Net struct's order
  void* net = nullptr;        // the Net struct on the device
  double* rscratch = nullptr;
  int* count = nullptr;  // [slots] on the device
  std::vector<int> last;
  std::vector<long long> total;
  int capacity = 0;  // cells the work buffers hold
  int *status = nullptr, *redo = nullptr, *list = nullptr;
  double* tries = nullptr;
  unsigned char* static_work = nullptr;  // this device's copy of scratch (mxc_solver_call_static)
};

namespace {

void release(mxc_solver* s) {
  if (!s) return;  // best effort: the failure being reported, or the close's own, comes first
  (void)hipSetDevice(s->device);
  for (void* p : s->tables) (void)hipFree(p);
  for (void* p : {(void*)s->net, (void*)s->rscratch, (void*)s->count, (void*)s->status, (void*)s->redo,
                  (void*)s->list, (void*)s->tries})
    if (p) (void)hipFree(p);
  if (s->module) (void)hipModuleUnload(s->module);
  delete s;
}

// meta.txt and the table files: every check before any device work. Fills s and the tables' host bytes.
int read_export(const std::string& folder, mxc_solver* s, std::string& code, std::vector<std::string>& bytes) {
  std::string text;
  if (!read_file(folder + "/meta.txt", text)) return fail("mxc_solver_open: cannot read " + folder + "/meta.txt");
  std::map<std::string, std::vector<std::string>> meta, args;
  std::vector<Table> tabs;
  std::istringstream lines(text);
  for (std::string line; std::getline(lines, line);) {
    std::istringstream ws(line);
    std::vector<std::string> tok;
    for (std::string t; ws >> t;) tok.push_back(t);
    if (tok.empty()) continue;
    if (tok[0] == "table") {
      if (tok.size() < 3) return fail("mxc_solver_open: meta.txt: a table line without name and dtype");
      Table t{tok[1], tok[2], {}, dtype_size(tok[2])};
      if (!t.bytes) return fail("mxc_solver_open: meta.txt: table " + t.name + " has dtype " + t.dtype);
      for (size_t i = 3; i < tok.size(); ++i) {
        const long long n = std::atoll(tok[i].c_str());
        if (n < 0) return fail("mxc_solver_open: meta.txt: table " + t.name + " has a negative size");
        t.shape.push_back(n);
        t.bytes *= static_cast<size_t>(n);
      }
      tabs.push_back(t);
    } else if (tok[0] == "args") {
      if (tok.size() < 2) return fail("mxc_solver_open: meta.txt: an args line without a kernel");
      args[tok[1]].assign(tok.begin() + 2, tok.end());
    } else {
      meta[tok[0]].assign(tok.begin() + 1, tok.end());
    }
  }
  auto one = [&](const char* key, std::string& out) {
    auto it = meta.find(key);
    if (it == meta.end() || it->second.size() != 1) return fail(std::string("mxc_solver_open: meta.txt lacks ") + key);
    out = it->second[0];
    return 0;
  };
  auto integer = [&](const char* key, int& out) {
    std::string v;
    if (one(key, v)) return 1;
    out = std::atoi(v.c_str());
    return 0;
  };
  auto real = [&](const char* key, double& out) {
    std::string v;
    if (one(key, v)) return 1;
    out = std::strtod(v.c_str(), nullptr);
    return 0;
  };
  int format = 0;
  if (integer("export_format", format)) return 1;
  if (format != kExportFormat)
    return fail("mxc_solver_open: export format " + std::to_string(format) + ", this library reads " +
                std::to_string(kExportFormat) + "; export again with this version");
  int wave = 0;
  if (integer("ns", s->ns) || integer("nj", s->nj) || integer("nr", s->nr) || integer("nel", s->nel) ||
      integer("wave", wave) || integer("cpw", s->cpw) || integer("lds_call", s->lds_call) ||
      integer("lds_rec", s->lds_rec) || integer("rates_global", s->rates_global) ||
      integer("rec_threads", s->rec_threads) ||
      integer("rec_steps", s->rec_steps) ||
      integer("max_steps", s->max_steps) || real("atol", s->atol) || real("rtol", s->rtol) ||
      real("t_min", s->t_min) || real("t_max", s->t_max) || real("x_sum_tol", s->x_sum_tol))
    return 1;
  if (wave != kWave || s->ns < 1 || s->nj < 0 || s->nr < 1 || s->nel < 1 || s->cpw < 1 ||
      s->rec_threads < kWave || s->rec_threads % kWave ||
      s->rec_steps < 1 || s->lds_call < 0 || s->lds_rec < 0)
    return fail("mxc_solver_open: meta.txt: sizes out of range");
  s->species = meta["species"];
  s->elements = meta["elements"];
  s->counts = meta["counts"];
  if (static_cast<int>(s->species.size()) != s->ns || static_cast<int>(s->elements.size()) != s->nel)
    return fail("mxc_solver_open: meta.txt names " + std::to_string(s->species.size()) + " species and " +
                std::to_string(s->elements.size()) + " elements; ns is " + std::to_string(s->ns) + ", nel " +
                std::to_string(s->nel));
  auto slot = [&](const char* name, int& out) {
    auto it = std::find(s->counts.begin(), s->counts.end(), name);
    if (it == s->counts.end()) return fail(std::string("mxc_solver_open: meta.txt counts lack ") + name);
    out = static_cast<int>(it - s->counts.begin());
    return 0;
  };
  if (slot("bad_input", s->bad_slot) || slot("cold", s->cold_slot) || slot("hot", s->hot_slot)) return 1;
  // each kernel's argument list: the names this library passes, each once, in its order
  for (const auto& k : kKernels) {
    auto refuse = [&](const std::string& what) {
      return fail(std::string("mxc_solver_open: kernel ") + k.kernel + " " + what + "; export again with this version");
    };
    auto it = args.find(k.kernel);
    if (it == args.end()) return refuse("has no args line in meta.txt");
    const auto& got = it->second;
    for (const auto& name : got) {
      if (std::find(k.names.begin(), k.names.end(), name) == k.names.end())
        return refuse("takes argument " + name + ", which this library does not pass");
      if (std::count(got.begin(), got.end(), name) > 1) return refuse("lists argument " + name + " twice");
    }
    for (const auto& name : k.names)
      if (std::find(got.begin(), got.end(), name) == got.end()) return refuse("lacks argument " + name);
    for (size_t i = 0; i < k.names.size(); ++i)
      if (got[i] != k.names[i])
        return refuse("has argument " + got[i] + " at position " + std::to_string(i) + ", where this library passes " +
                      k.names[i]);
  }
  // the Net struct's fields and their tables
  const auto net = meta["net"];
  if (net.size() != kNetCount) return fail("mxc_solver_open: meta.txt: the net line lists another set of tables");
  for (int i = 0; i < kNetCount; ++i)
    if (net[i] != kNetFields[i])
      return fail("mxc_solver_open: meta.txt: net field " + std::to_string(i) + " is " + net[i] + ", the kernels read " +
                  kNetFields[i] + " there");
  bytes.assign(kNetCount, std::string());
  for (int i = 0; i < kNetCount; ++i) {
    auto t = std::find_if(tabs.begin(), tabs.end(), [&](const Table& q) { return q.name == kNetFields[i]; });
    if (t == tabs.end()) return fail(std::string("mxc_solver_open: meta.txt lists no table ") + kNetFields[i]);
    const std::string path = folder + "/tables/" + t->name + ".bin";
    if (!read_file(path, bytes[i])) return fail("mxc_solver_open: cannot read " + path);
    if (bytes[i].size() != t->bytes)
      return fail("mxc_solver_open: table " + t->name + " has " + std::to_string(bytes[i].size()) +
                  " bytes; its dtype and shape in meta.txt need " + std::to_string(t->bytes));
  }
  const size_t atoms_bytes = bytes[net_field("atoms")].size();
  if (atoms_bytes != sizeof(double) * s->ns * s->nel)
    return fail("mxc_solver_open: the atoms table is not species x elements float64");
  std::string code_name;
  if (one("code", code_name) || one("arch", s->arch)) return 1;
  if (!read_file(folder + "/" + code_name, code) || code.empty())
    return fail("mxc_solver_open: cannot read " + folder + "/" + code_name);
  return 0;
}

// the device's architecture (gcnArchName up to its first ':', e.g. gfx950)
int device_arch(int device, std::string& arch) {
  hipDeviceProp_t p;
  if (check(hipGetDeviceProperties(&p, device), "hipGetDeviceProperties")) return 1;
  arch = p.gcnArchName;
  arch = arch.substr(0, arch.find(':'));
  return 0;
}

int grow(mxc_solver* s, int cells) {
  if (cells <= s->capacity) return 0;
  for (void* p : {(void*)s->status, (void*)s->redo, (void*)s->list, (void*)s->tries})
    if (p && check(hipFree(p), "hipFree")) return 1;
  s->status = s->redo = s->list = nullptr;
  s->tries = nullptr;
  s->capacity = 0;
  const size_t n = static_cast<size_t>(cells);
  if (check(hipMalloc((void**)&s->status, n * sizeof(int)), "hipMalloc") ||
      check(hipMalloc((void**)&s->redo, n * sizeof(int)), "hipMalloc") ||
      check(hipMalloc((void**)&s->list, n * sizeof(int)), "hipMalloc") ||
      check(hipMalloc((void**)&s->tries, n * sizeof(double)), "hipMalloc"))
    return 1;
  s->capacity = cells;
  return 0;
}

// The refusal message of the call just made (the Python solver's, maxichem/_solver.py:183-224), or empty.
int refusal(mxc_solver* s, hipStream_t stream, int B, const int* status_d, const double* x, const double* T,
            const double* P, const double* dt, const double* J, std::string& message) {
  const int bad = s->last[s->bad_slot], cold = s->last[s->cold_slot], hot = s->last[s->hot_slot];
  message.clear();
  if (!bad && !cold && !hot) return 0;
  std::vector<int> status(B);
  std::vector<double> Th(B);
  if (check(hipMemcpyAsync(status.data(), status_d, sizeof(int) * B, hipMemcpyDeviceToHost, stream), "copy") ||
      check(hipMemcpyAsync(Th.data(), T, sizeof(double) * B, hipMemcpyDeviceToHost, stream), "copy") ||
      check(hipStreamSynchronize(stream), "hipStreamSynchronize"))
    return 1;
  auto first = [&](int kind) {  // 0 cold, 1 hot, 2 neither
    for (int i = 0; i < B; ++i) {
      if (!status[i]) continue;
      const bool c = Th[i] < s->t_min, h = Th[i] > s->t_max;
      if ((kind == 0 && c) || (kind == 1 && h) || (kind == 2 && !c && !h)) return i;
    }
    return -1;
  };
  std::vector<std::string> parts;
  const std::string range = g(s->t_min) + "-" + g(s->t_max) + " K";
  if (cold) {
    const int i = first(0);
    parts.push_back(std::to_string(cold) + " cells below " + g(s->t_min) + " K, outside the solver's range " + range +
                    " (first: cell " + std::to_string(i) + ", T = " + g(i >= 0 ? Th[i] : NAN) +
                    " K); leave them out of the call");
  }
  if (hot) {
    const int i = first(1);
    parts.push_back(std::to_string(hot) + " cells above " + g(s->t_max) + " K, outside the solver's range " + range +
                    " (first: cell " + std::to_string(i) + ", T = " + g(i >= 0 ? Th[i] : NAN) +
                    " K); leave them out of the call");
  }
  if (bad) {
    const int i = first(2);
    std::string where = "cell " + std::to_string(i);
    if (i >= 0) {
      std::vector<double> row(s->ns), jrow(s->nj);
      double p = 0.0, d = 0.0;
      if (check(hipMemcpyAsync(row.data(), x + static_cast<size_t>(i) * s->ns, sizeof(double) * s->ns,
                               hipMemcpyDeviceToHost, stream),
                "copy") ||
          check(hipMemcpyAsync(&p, P + i, sizeof(double), hipMemcpyDeviceToHost, stream), "copy") ||
          check(hipMemcpyAsync(&d, dt + i, sizeof(double), hipMemcpyDeviceToHost, stream), "copy") ||
          (s->nj && check(hipMemcpyAsync(jrow.data(), J + static_cast<size_t>(i) * s->nj, sizeof(double) * s->nj,
                                         hipMemcpyDeviceToHost, stream),
                          "copy")) ||
          check(hipStreamSynchronize(stream), "hipStreamSynchronize"))
        return 1;
      auto least = [](const std::vector<double>& v) {  // NaN if any entry is NaN
        double m = INFINITY;
        for (double q : v) m = std::isnan(q) || std::isnan(m) ? NAN : std::min(m, q);
        return m;
      };
      double sum = 0.0;  // in index order
      for (double q : row) sum += q;
      where += ": sum(x) - 1 = " + g(sum - 1.0) +
               " (added in another order than the kernel's, so it can differ in the last bits), min(x) = " +
               g(least(row)) + ", T = " + g(Th[i]) + " K, P = " + g(p) + " dyn cm^-2, dt = " + g(d) + " s";
      if (s->nj) where += ", min(J) = " + g(least(jrow)) + " s^-1";
    }
    parts.push_back(std::to_string(bad) + " cells of bad input: x negative or non-finite, |sum(x) - 1| above " +
                    g(s->x_sum_tol) +
                    " or no x above 0; T non-finite, P non-finite, not above 0 or so large that P / (k_B T) "
                    "overflows, dt negative or non-finite, or J negative or non-finite (first: " +
                    where + "); pass x normalized to sum 1 in float64, clamped at 0");
  }
  message = "the call refused cells: ";
  for (size_t k = 0; k < parts.size(); ++k) message += (k ? "; " : "") + parts[k];
  return 0;
}

}  // namespace

extern "C" {

int mxc_solver_open(const char* folder, int device, mxc_solver** solver) {
  mxc_detail::clear_error();
  if (!folder || !solver) return fail("mxc_solver_open: null argument");
  *solver = nullptr;
  auto* s = new mxc_solver();
  s->device = device;
  std::string code;
  std::vector<std::string> bytes;
  if (read_export(folder, s, code, bytes)) {
    delete s;
    return 1;
  }
  // device work from here on
  std::string arch;
  if (check(hipSetDevice(device), "hipSetDevice") || device_arch(device, arch)) {
    delete s;
    return 1;
  }
  if (arch != s->arch) {
    const std::string want = s->arch;
    delete s;
    return fail("mxc_solver_open: the export's code object is for " + want + ", device " + std::to_string(device) +
                " is " + arch + "; export for this device");
  }
  int cus = 0, occ_call = 0, occ_rec = 0;
  const bool bad =
      check(hipModuleLoadData(&s->module, code.data()), "hipModuleLoadData") ||
      check(hipModuleGetFunction(&s->call, s->module, "mxc_call"), "kernel mxc_call") ||
      check(hipModuleGetFunction(&s->recover, s->module, "mxc_recover"), "kernel mxc_recover") ||
      check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "compute units") ||
      check(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&occ_call, s->call, kWave * s->cpw, s->lds_call),
            "occupancy") ||
      check(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&occ_rec, s->recover, s->rec_threads, s->lds_rec),
            "occupancy");
  if (bad) {
    release(s);
    return 1;
  }
  s->max_blocks = cus * std::max(1, occ_call);
  s->rec_blocks = cus * std::max(1, occ_rec);
  std::vector<void*> ptrs;
  for (const auto& b : bytes) {
    void* d = nullptr;
    if (check(hipMalloc(&d, std::max<size_t>(b.size(), 8)), "hipMalloc")) {
      release(s);
      return 1;
    }
    s->tables.push_back(d);
    if (check(hipMemcpy(d, b.data(), b.size(), hipMemcpyHostToDevice), "hipMemcpy")) {
      release(s);
      return 1;
    }
  }
  const int slots = static_cast<int>(s->counts.size());
  if (check(hipMalloc(&s->net, sizeof(void*) * kNetCount), "hipMalloc") ||
      check(hipMemcpy(s->net, s->tables.data(), sizeof(void*) * kNetCount, hipMemcpyHostToDevice), "hipMemcpy") ||
      check(hipMalloc((void**)&s->count, sizeof(int) * slots), "hipMalloc") ||
      check(hipMemset(s->count, 0, sizeof(int) * slots), "hipMemset")) {
    release(s);
    return 1;
  }
  if (s->rates_global) {  // each call wavefront's and each recovery workgroup's slot of rate coefficients
    const size_t n = static_cast<size_t>(std::max(s->max_blocks * s->cpw, s->rec_blocks)) * 2 * s->nr;
    if (check(hipMalloc((void**)&s->rscratch, n * sizeof(double)), "hipMalloc")) {
      release(s);
      return 1;
    }
  }
  void* work = nullptr;  // this device's work buffers for mxc_solver_call_static (the device is current)
  if (check(hipGetSymbolAddress(&work, HIP_SYMBOL(scratch)), "hipGetSymbolAddress")) {
    release(s);
    return 1;
  }
  s->static_work = static_cast<unsigned char*>(work);
  s->last.assign(slots, 0);
  s->total.assign(slots, 0);
  *solver = s;
  return 0;
}

int mxc_solver_close(mxc_solver* solver) {
  mxc_detail::clear_error();
  if (!solver) return 0;
  if (check(hipSetDevice(solver->device), "hipSetDevice")) return 1;
  const int bad = check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  release(solver);
  return bad;
}

int mxc_solver_sizes(const mxc_solver* s, int* ns, int* nj, int* nel, int* count_slots) {
  mxc_detail::clear_error();
  if (!s) return fail("mxc_solver_sizes: null solver");
  if (ns) *ns = s->ns;
  if (nj) *nj = s->nj;
  if (nel) *nel = s->nel;
  if (count_slots) *count_slots = static_cast<int>(s->counts.size());
  return 0;
}

int mxc_solver_defaults(const mxc_solver* s, double* rtol, double* atol, int* max_steps) {
  mxc_detail::clear_error();
  if (!s) return fail("mxc_solver_defaults: null solver");
  if (rtol) *rtol = s->rtol;
  if (atol) *atol = s->atol;
  if (max_steps) *max_steps = s->max_steps;
  return 0;
}

const char* mxc_solver_name(const mxc_solver* s, int list, int index) {
  if (!s || index < 0) return nullptr;
  const std::vector<std::string>* v = list == 0 ? &s->species : list == 1 ? &s->elements : list == 2 ? &s->counts
                                                                                                     : nullptr;
  if (!v || index >= static_cast<int>(v->size())) return nullptr;
  return (*v)[index].c_str();
}

namespace {

// The argument checks of a call, before any device work.
int check_call(const mxc_solver* s, int cells, const double* x, const double* T, const double* P, const double* dt,
               const double* h, const double* J, double rtol, double atol, int max_steps, double* x_out,
               double* h_out, int* done, double* steps, const char* who) {
  if (!s) return fail(std::string(who) + ": null solver");
  if (cells < 0) return fail(std::string(who) + ": cells: at least 0 expected, got " + std::to_string(cells));
  if (cells > 0 && (!x || !T || !P || !dt || !h || !x_out || !h_out || !done || !steps))
    return fail(std::string(who) + ": null input or output pointer");
  if (s->nj && cells > 0 && !J)
    return fail("J [B, " + std::to_string(s->nj) + "] is required with photolysis: pass zeros for dark cells");
  if (!s->nj && J) return fail("J: none expected, since this solver has no photolysis");
  if (!(std::isfinite(rtol) && rtol > 0.0 && std::isfinite(atol) && atol > 0.0))
    return fail("rtol and atol finite and positive expected, got " + g(rtol) + ", " + g(atol));
  const int most = std::numeric_limits<int>::max() / s->rec_steps;
  if (max_steps < 1 || max_steps > most)
    return fail("max_steps from 1 to " + std::to_string(most) + " expected, got " + std::to_string(max_steps));
  return 0;
}

// A call with the given work buffers [cells] on the device: status, redo (the flags), list and tries.
int run_call(mxc_solver* s, hipStream_t stream, int cells, const double* x, const double* T, const double* P,
             const double* dt, const double* h, const double* J, double rtol, double atol, int max_steps,
             double* x_out, double* h_out, int* done, double* steps, int* status, int* redo, int* list,
             double* tries) {
  const int slots = static_cast<int>(s->counts.size());
  const size_t B = static_cast<size_t>(cells);
  if ((cells && check(hipMemcpyAsync(h_out, h, B * sizeof(double), hipMemcpyDeviceToDevice, stream), "copy h")) ||
      (cells && check(hipMemsetAsync(tries, 0, B * sizeof(double), stream), "clear tries")) ||
      check(hipMemsetAsync(s->count, 0, slots * sizeof(int), stream), "clear counts"))
    return 1;
  const int* order = nullptr;  // cells in index order (results do not depend on the order)
  const double* xprop = nullptr;
  int* fp64 = nullptr;  // the mixed form's cells that fell back to FP64 in an earlier call: none kept here
  int n = cells, rec_steps = s->rec_steps * max_steps;
  // argument values by name, in each kernel's order
  std::map<std::string, void*> v{
      {"net", &s->net},   {"x0", &x},       {"Ta", &T},        {"Pa", &P},          {"dta", &dt},
      {"hin", &h},        {"hio", &h_out},  {"xout", &x_out},  {"doneo", &done},    {"noko", &steps},
      {"status", &status}, {"redo", &redo}, {"list", &list},  {"count", &s->count}, {"tries", &tries},
      {"order", &order},  {"Jin", &J},      {"B", &n},         {"rtol", &rtol},     {"atol", &atol},
      {"max_steps", &max_steps}, {"xprop", &xprop}, {"rscratch", &s->rscratch}, {"fp64", &fp64}};
  auto launch = [&](int k, hipFunction_t f, unsigned grid, unsigned block, unsigned lds) {
    std::vector<void*> a;
    for (const auto& name : kKernels[k].names) a.push_back(v.at(name));
    return check(hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, lds, stream, a.data(), nullptr),
                 std::string("launch ") + kKernels[k].kernel);
  };
  const unsigned call_grid = static_cast<unsigned>(std::max(1, std::min((cells + s->cpw - 1) / s->cpw, s->max_blocks)));
  if (launch(0, s->call, call_grid, kWave * s->cpw, s->lds_call)) return 1;
  v["max_steps"] = &rec_steps;
  const unsigned rec_grid = static_cast<unsigned>(std::max(1, std::min(cells, s->rec_blocks)));
  if (launch(1, s->recover, rec_grid, s->rec_threads, s->lds_rec)) return 1;
  if (check(hipMemcpyAsync(s->last.data(), s->count, slots * sizeof(int), hipMemcpyDeviceToHost, stream),
            "copy counts") ||
      check(hipStreamSynchronize(stream), "hipStreamSynchronize"))  // the call's one host sync
    return 1;
  for (int i = 0; i < slots; ++i) s->total[i] += s->last[i];
  std::string message;
  if (refusal(s, stream, cells, status, x, T, P, dt, J, message)) return 1;
  if (!message.empty()) {
    fail(message);
    return MXC_REFUSED;
  }
  return 0;
}

}  // namespace

int mxc_solver_call(mxc_solver* s, void* stream_, int cells, const double* x, const double* T, const double* P,
                    const double* dt, const double* h, const double* J, double rtol, double atol, int max_steps,
                    double* x_out, double* h_out, int* done, double* steps, int* flag) {
  mxc_detail::clear_error();
  if (check_call(s, cells, x, T, P, dt, h, J, rtol, atol, max_steps, x_out, h_out, done, steps, "mxc_solver_call"))
    return 1;
  if (check(hipSetDevice(s->device), "hipSetDevice") || grow(s, std::max(cells, 1))) return 1;
  return run_call(s, static_cast<hipStream_t>(stream_), cells, x, T, P, dt, h, J, rtol, atol, max_steps, x_out,
                  h_out, done, steps, s->status, flag ? flag : s->redo, s->list, s->tries);
}

int mxc_solver_call_static(mxc_solver* s, void* stream_, int cells, const double* x, const double* T,
                           const double* P, const double* dt, const double* h, const double* J, double rtol,
                           double atol, int max_steps, double* x_out, double* h_out, int* done, double* steps,
                           int* flag) {
  mxc_detail::clear_error();
  if (check_call(s, cells, x, T, P, dt, h, J, rtol, atol, max_steps, x_out, h_out, done, steps,
                 "mxc_solver_call_static"))
    return 1;
  if (cells > MXC_STATIC_CELLS)
    return fail("mxc_solver_call_static: at most " + std::to_string(MXC_STATIC_CELLS) + " cells per call, got " +
                std::to_string(cells) + "; split the batch, or call mxc_solver_call");
  if (check(hipSetDevice(s->device), "hipSetDevice")) return 1;
  // this device's scratch, carved as its comment says: tries, then status, redo and list
  const size_t n = static_cast<size_t>(MXC_STATIC_CELLS);
  double* tries = reinterpret_cast<double*>(s->static_work);
  int* status = reinterpret_cast<int*>(s->static_work + n * sizeof(double));
  int* redo = status + n;
  int* list = redo + n;
  return run_call(s, static_cast<hipStream_t>(stream_), cells, x, T, P, dt, h, J, rtol, atol, max_steps, x_out,
                  h_out, done, steps, status, flag ? flag : redo, list, tries);
}

int mxc_solver_counts(const mxc_solver* s, int* count, long long* total) {
  mxc_detail::clear_error();
  if (!s) return fail("mxc_solver_counts: null solver");
  for (size_t i = 0; i < s->counts.size(); ++i) {
    if (count) count[i] = s->last[i];
    if (total) total[i] = s->total[i];
  }
  return 0;
}

int mxc_solver_atoms(const mxc_solver* s, double* atoms) {
  mxc_detail::clear_error();
  if (!s || !atoms) return fail("mxc_solver_atoms: null argument");
  if (check(hipSetDevice(s->device), "hipSetDevice")) return 1;
  return check(hipMemcpy(atoms, s->tables[net_field("atoms")], sizeof(double) * s->ns * s->nel,
                         hipMemcpyDeviceToHost),
               "hipMemcpy");
}

int mxc_kernel_arguments(const char* kernel, char* names, size_t size) {
  mxc_detail::clear_error();
  for (const auto& k : kKernels) {
    if (!kernel || std::strcmp(kernel, k.kernel)) continue;
    std::string all;
    for (const auto& n : k.names) all += (all.empty() ? "" : " ") + n;
    if (!names || size <= all.size()) return fail("mxc_kernel_arguments: a buffer of " + std::to_string(all.size() + 1) + " bytes needed");
    std::memcpy(names, all.c_str(), all.size() + 1);
    return 0;
  }
  return fail(std::string("mxc_kernel_arguments: no kernel ") + (kernel ? kernel : "(null)"));
}

}  // extern "C"
