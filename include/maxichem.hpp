#pragma once

// maxichem from C++ with libtorch tensors: the kernels a Python CudaSolver exported with CudaSolver.export(folder),
// run on torch's current HIP stream with the arguments the Python call passes, so a call gives the same bits. The
// class, its members and its errors are those of the reference's C++ caller (cpp/maxichem.hpp:1-94), over the C ABI
// of maxichem_amd.h (link libmaxichem_amd.so):
//
//   maxichem::Solver chem("export/NCHO_photo", torch::Device(torch::kCUDA, 0));
//   auto [x1, h1, done, steps] = chem(x, T, P, dt, h, J);
//
// x [B, ns] mole fractions in the network's species order, T [B] K, P [B] dyn cm^-2, dt [B] s, h [B] the starting
// step sizes (float64 on the solver's device); J [B, nj] s^-1 (float64 on the device), required for an export with
// photolysis (zeros for dark cells). Without photolysis (nj() == 0) the call takes no J (an undefined tensor) or an
// empty [B, 0] one and refuses a J with columns, the Python call's rule and message. Returns the new mole fractions,
// the next step sizes, done (bool) and the accepted steps per cell (float64). count() is the last call's counts
// (int32: queue, bad_input, cold, hot, nonfinite, unfinished, drift, recomputed, recovered, still_failed, nonfinite_h,
// recovery queue) and total() their sum since construction (int64), both on the device. The call reads the counts
// back (its one host sync) and throws std::invalid_argument, with the Python call's message after "maxichem: ", if it
// refused a cell (bad input, or T outside the solver's range); the counts stay readable. The constructor throws
// std::runtime_error, before any device work, for an export of another format or kernel argument lists, and for a
// device of another architecture than the code object's.
//
// Calls share the solver's counts and work buffers, so they run one at a time: on one stream, or synchronize before
// calling from another (as in Python).

#include <c10/core/DeviceGuard.h>
#include <c10/hip/HIPStream.h>
#include <torch/torch.h>

#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "maxichem_amd.h"

namespace maxichem {

// The Python call's defaults (maxichem_amd/_constants.py DEFAULT_RTOL, DEFAULT_MAX_STEPS; the export's meta.txt
// holds the same and default_atol()).
inline constexpr double kDefaultRtol = 1e-3;
inline constexpr int kDefaultMaxSteps = 4000;

// What a call returns, per cell, with the field names of the Python CallResult.
struct CallResult {
  torch::Tensor x;      // [B, ns] mole fractions at the end of the call (float64)
  torch::Tensor h;      // [B] the step size to start the cell's next call with, s (float64)
  torch::Tensor done;   // [B] the call finished and its result passed the guard (bool)
  torch::Tensor steps;  // [B] accepted steps (float64)
};

class Solver {
 public:
  Solver(std::string const& folder, torch::Device device) : device_(device) {
    TORCH_CHECK(device.is_cuda(), "maxichem::Solver: a GPU device");
    c10::DeviceGuard guard(device);
    c10::DeviceIndex const index = guard.current_device().index();  // "cuda" alone: the current device
    device_ = torch::Device(torch::kCUDA, index);
    torch::zeros({1}, torch::TensorOptions().device(device_));  // torch's runtime set up on the device first
    mxc_solver* s = nullptr;
    if (mxc_solver_open(folder.c_str(), index, &s)) throw std::runtime_error(std::string("maxichem: ") + mxc_last_error());
    solver_ = s;
    int slots = 0, nel = 0;
    mxc_solver_sizes(s, &ns_, &nj_, &nel, &slots);
    int steps = 0;
    double rtol = 0.0;
    mxc_solver_defaults(s, &rtol, &default_atol_, &steps);
    for (int i = 0; i < ns_; ++i) species_.emplace_back(mxc_solver_name(s, 0, i));
    for (int i = 0; i < nel; ++i) elements_.emplace_back(mxc_solver_name(s, 1, i));
    atoms_ = torch::empty({ns_, nel}, torch::kFloat64);
    check(mxc_solver_atoms(s, atoms_.data_ptr<double>()));
    atoms_ = atoms_.to(device_);
    count_ = torch::zeros({slots}, torch::TensorOptions().dtype(torch::kInt32).device(device_));
    total_ = torch::zeros({slots}, torch::TensorOptions().dtype(torch::kInt64).device(device_));
  }
  ~Solver() {
    if (solver_) mxc_solver_close(solver_);
  }
  Solver(Solver const&) = delete;
  Solver& operator=(Solver const&) = delete;

  CallResult operator()(torch::Tensor x, torch::Tensor T, torch::Tensor P, torch::Tensor dt, torch::Tensor h,
                        torch::Tensor J = torch::Tensor(), double rtol = kDefaultRtol, double atol = -1.0,
                        int max_steps = kDefaultMaxSteps) {
    std::pair<char const*, torch::Tensor*> given[] = {{"x", &x}, {"T", &T}, {"P", &P}, {"dt", &dt}, {"h", &h}};
    for (auto const& [name, a] : given)
      TORCH_CHECK(a->device() == device_ && a->scalar_type() == torch::kFloat64, "maxichem: ", name, ": float64 on ",
                  device_, " expected, got ", a->scalar_type(), " on ", a->device());
    TORCH_CHECK(x.dim() == 2 && x.size(1) == ns_, "maxichem: x: [B, ", ns_, "] expected, got ", x.sizes());
    TORCH_CHECK(x.size(0) <= std::numeric_limits<int>::max(), "maxichem: at most 2^31 - 1 cells per call, got ",
                x.size(0));
    int const B = static_cast<int>(x.size(0));
    for (auto const& [name, a] : given)
      TORCH_CHECK(a == &x || (a->dim() == 1 && a->size(0) == B), "maxichem: ", name, ": [", B, "] expected, got ",
                  a->sizes());
    if (atol < 0) atol = default_atol_;
    c10::DeviceGuard guard(device_);
    for (auto* a : {&x, &T, &P, &dt, &h}) *a = a->contiguous();
    // the Python call's photolysis rule and messages (maxichem_amd/_solver.py check_photolysis)
    double const* jp = nullptr;
    if (J.defined()) {
      TORCH_CHECK(nj_ || (J.dim() == 2 && J.size(1) == 0), "maxichem: J: none or an empty [", B,
                  ", 0] one expected, since this solver has no photolysis; got ", J.sizes());
      TORCH_CHECK(J.device() == device_ && J.scalar_type() == torch::kFloat64 && J.dim() == 2 && J.size(0) == B &&
                      J.size(1) == nj_,
                  "maxichem: J: float64 [", B, ", ", nj_, "] on ", device_, " expected, got ", J.scalar_type(), " ",
                  J.sizes(), " on ", J.device());
      if (nj_) {
        J = J.contiguous();
        jp = J.data_ptr<double>();
      }
    } else {
      TORCH_CHECK(nj_ == 0, "maxichem: J [B, ", nj_, "] is required with photolysis: pass zeros for dark cells");
    }
    auto xo = torch::empty_like(x), ho = torch::empty_like(h), steps = torch::empty_like(h);
    auto done = torch::empty({B}, x.options().dtype(torch::kInt32));
    void* stream = c10::hip::getCurrentHIPStream(device_.index()).stream();
    int const rc = mxc_solver_call(solver_, stream, B, x.data_ptr<double>(), T.data_ptr<double>(),
                                   P.data_ptr<double>(), dt.data_ptr<double>(), h.data_ptr<double>(), jp, rtol, atol,
                                   max_steps, xo.data_ptr<double>(), ho.data_ptr<double>(), done.data_ptr<int>(),
                                   steps.data_ptr<double>(), nullptr);
    std::string const message = std::string("maxichem: ") + mxc_last_error();  // before another entry point clears it
    if (rc == 1) throw std::runtime_error(message);
    read_counts();
    if (rc == MXC_REFUSED) throw std::invalid_argument(message);
    return {xo, ho, done.to(torch::kBool), steps};
  }

  // Read only: the accessors return copies, so a caller's change to them leaves the solver's own.
  int ns() const { return ns_; }                                          // species
  int nj() const { return nj_; }                                          // photolysis branches (0: thermal network)
  std::vector<std::string> const& species() const { return species_; }    // names, in x's column order
  std::vector<std::string> const& elements() const { return elements_; }  // names, in atoms()' column order
  double default_atol() const { return default_atol_; }                   // mole fraction
  torch::Tensor atoms() const { return atoms_.clone(); }  // species x elements (float64, on the device)
  torch::Tensor count() const { return count_.clone(); }  // the last call's counts (int32, on the device)
  torch::Tensor total() const { return total_.clone(); }  // their sum since construction (int64, on the device)

 private:
  static void check(int rc) {
    if (rc) throw std::runtime_error(std::string("maxichem: ") + mxc_last_error());
  }
  void read_counts() {
    auto c = torch::empty({count_.size(0)}, torch::kInt32);
    auto t = torch::empty({total_.size(0)}, torch::kInt64);
    static_assert(sizeof(long long) == sizeof(int64_t));
    check(mxc_solver_counts(solver_, c.data_ptr<int>(), reinterpret_cast<long long*>(t.data_ptr<int64_t>())));
    count_.copy_(c);
    total_.copy_(t);
  }

  mxc_solver* solver_ = nullptr;
  torch::Device device_;
  int ns_ = 0, nj_ = 0;
  double default_atol_ = 0.0;
  std::vector<std::string> species_, elements_;
  torch::Tensor atoms_, count_, total_;
};

}  // namespace maxichem
