// Runs maxichem::Solver (include/maxichem.hpp) on inputs a Python solver saved and compares the results bit for bit:
//   test_export <export folder> <case folder>
// The case folder holds raw little-endian files written by tests/test_export.py: the inputs x.bin [B, ns], T.bin,
// P.bin, dt.bin, h.bin [B] and, with photolysis, J.bin [B, nj] (float64); the Python call's results x_out.bin,
// h_out.bin, steps_out.bin (float64), done_out.bin (int8) and counts_out.bin (int32); B.txt and max_steps.txt. If the
// case holds refused.txt, the Python call refused it with that message: the C++ call must throw std::invalid_argument
// with the same message after "maxichem: " and leave the same counts. Otherwise the call runs twice and both results
// are compared in full, counts included; at the default max_steps the call uses the default arguments. A thermal
// export is also called with an empty [B, 0] J (the same bits) and a [B, 1] J (refused). Then the tensors the
// accessors return are changed in place, and a further call must still give the same bits, the solver's own totals
// and its own atoms table. Prints one line; exits 0 when every comparison holds.
#include <chrono>
#include <cstring>
#include <fstream>
#include <iostream>
#include <iterator>
#include <set>
#include <sstream>
#include <stdexcept>

#include "maxichem.hpp"

static torch::Tensor load(std::string const& path, std::vector<int64_t> shape, torch::Dtype dt) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot read " + path);
  std::ostringstream ss;
  ss << f.rdbuf();
  std::string const b = ss.str();
  auto t = torch::empty(shape, torch::TensorOptions().dtype(dt));
  if (b.size() != static_cast<size_t>(t.nbytes())) throw std::runtime_error("wrong size: " + path);
  std::memcpy(t.data_ptr(), b.data(), b.size());
  return t;
}

static int hip_runtimes() {  // libamdhip64 images mapped in this process
  std::ifstream maps("/proc/self/maps");
  std::set<std::string> paths;
  for (std::string line; std::getline(maps, line);)
    if (line.find("libamdhip64") != std::string::npos) paths.insert(line.substr(line.rfind(' ') + 1));
  return static_cast<int>(paths.size());
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::cerr << "usage: test_export <export folder> <case folder>\n";
    return 2;
  }
  std::string const ex = argv[1], cs = argv[2];
  torch::Device dev(torch::kCUDA, 0);
  maxichem::Solver chem(ex, dev);
  std::cout << "export loaded; " << std::flush;  // an export the constructor refuses never gets here
  int64_t B = 0;
  if (!(std::ifstream(cs + "/B.txt") >> B) || B < 0) throw std::runtime_error("cannot read " + cs + "/B.txt");
  auto const f64 = torch::kFloat64;
  auto x = load(cs + "/x.bin", {B, chem.ns()}, f64).to(dev);
  auto T = load(cs + "/T.bin", {B}, f64).to(dev);
  auto P = load(cs + "/P.bin", {B}, f64).to(dev);
  auto dt = load(cs + "/dt.bin", {B}, f64).to(dev);
  auto h = load(cs + "/h.bin", {B}, f64).to(dev);
  torch::Tensor J;
  if (chem.nj()) J = load(cs + "/J.bin", {B, chem.nj()}, f64).to(dev);
  int max_steps = 0;
  if (!(std::ifstream(cs + "/max_steps.txt") >> max_steps) || max_steps < 1)
    throw std::runtime_error("cannot read " + cs + "/max_steps.txt");
  bool const defaults = max_steps == maxichem::kDefaultMaxSteps;
  auto call = [&] {
    return defaults ? chem(x, T, P, dt, h, J) : chem(x, T, P, dt, h, J, maxichem::kDefaultRtol, -1.0, max_steps);
  };
  auto const co = load(cs + "/counts_out.bin", {chem.count().size(0)}, torch::kInt32);
  if (std::ifstream expected_file{cs + "/refused.txt"}) {  // the Python call refused this input: so must this one
    std::string const expected((std::istreambuf_iterator<char>(expected_file)), std::istreambuf_iterator<char>());
    std::string message;
    try {
      call();
    } catch (std::invalid_argument const& e) {
      message = e.what();
    }
    std::string const prefix = "maxichem: ";
    bool const same_message = message.rfind(prefix, 0) == 0 && message.substr(prefix.size()) == expected;
    bool const same_counts = torch::equal(chem.count().cpu(), co);
    std::cout << "C++ call " << (same_message && same_counts ? "refused as the Python call" : "DIFFERENT")
              << "; message: " << message << "; counts " << chem.count().cpu().view({1, -1}) << std::endl;
    return same_message && same_counts ? 0 : 1;
  }
  auto const xo = load(cs + "/x_out.bin", {B, chem.ns()}, f64);
  auto const ho = load(cs + "/h_out.bin", {B}, f64);
  auto const so = load(cs + "/steps_out.bin", {B}, f64);
  auto const dn = load(cs + "/done_out.bin", {B}, torch::kInt8).to(torch::kBool);
  auto same = [](torch::Tensor a, torch::Tensor b) {  // bit for bit
    return torch::equal(a.cpu().contiguous().view(torch::kInt64), b.contiguous().view(torch::kInt64));
  };
  auto all_same = [&](maxichem::CallResult const& q, torch::Tensor const& c) {
    auto const& [qx, qh, qdone, qsteps] = q;  // the structured binding the header promises, in field order
    bool const in_order = qx.is_same(q.x) && qh.is_same(q.h) && qdone.is_same(q.done) && qsteps.is_same(q.steps);
    return in_order && same(q.x, xo) && same(q.h, ho) && same(q.steps, so) && torch::equal(q.done.cpu(), dn) &&
           torch::equal(c, co);
  };
  auto r = call();
  auto counts = chem.count().cpu();
  torch::cuda::synchronize();
  auto t0 = std::chrono::steady_clock::now();
  auto r2 = call();
  auto counts2 = chem.count().cpu();
  torch::cuda::synchronize();
  double const ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  bool ok = all_same(r, counts) && all_same(r2, counts2);
  if (chem.nj() == 0) {  // without photolysis an empty [B, 0] J is no J; any other J is refused before a launch
    auto r3 = chem(x, T, P, dt, h, torch::zeros({B, 0}, x.options()), maxichem::kDefaultRtol, -1.0, max_steps);
    bool const same_empty = all_same(r3, chem.count().cpu());
    std::string refusal;
    try {
      chem(x, T, P, dt, h, torch::zeros({B, 1}, x.options()), maxichem::kDefaultRtol, -1.0, max_steps);
    } catch (c10::Error const& e) {
      refusal = e.what_without_backtrace();
    }
    ok = ok && same_empty && !refusal.empty();
    std::cout << "J [B, 0] " << (same_empty ? "gives the same bits" : "DIFFERENT") << ", J [B, 1] "
              << (refusal.empty() ? "NOT REFUSED" : "refused") << " (" << refusal << "); ";
  }
  // what the accessors return is the caller's to change: the solver's counts and tables stay as they were
  auto const total_before = chem.total().cpu(), atoms_before = chem.atoms().cpu();
  chem.count().resize_({1});
  chem.total().zero_();
  chem.atoms().zero_();
  auto r4 = call();
  bool const own = torch::equal(chem.total().cpu(), total_before + co.to(torch::kInt64)) &&
                   torch::equal(chem.atoms().cpu(), atoms_before) && all_same(r4, chem.count().cpu());
  ok = ok && own;
  int const runtimes = hip_runtimes();
  ok = ok && runtimes == 1;
  std::cout << (own ? "accessors return copies" : "ACCESSORS SHARE THE SOLVER'S TENSORS") << "; HIP runtimes mapped "
            << runtimes << "; cells " << B << ", species " << chem.ns() << ", branches " << chem.nj() << ": C++ call "
            << (ok ? "bitwise identical to the Python call" : "DIFFERENT") << "; second call " << ms << " ms"
            << "; max_steps " << max_steps << (defaults ? " (default arguments)" : "") << "; counts "
            << counts.view({1, -1}) << std::endl;
  return ok ? 0 : 1;
}
