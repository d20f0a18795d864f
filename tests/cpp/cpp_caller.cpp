// A plan-driven caller of maxichem::Solver (include/maxichem.hpp) for tests/test_export.py, which does every
// comparison itself:
//   cpp_caller <export folder> <case folder> <output folder>
// The case folder holds the inputs as raw little-endian float64 files (x.bin [B, ns], T.bin, P.bin, dt.bin, h.bin [B],
// J.bin [B, nj] when the export has photolysis), B.txt, and plan.txt: one step per line,
//   call <tag> <max_steps | default> <given | none | empty | columns>   one call; J as saved, absent, [B, 0] or [B, 1]
//   scramble                                                           change the accessors' tensors in place
//   state <tag>                                                        write the accessors' count, total and atoms
// A call writes <tag>.x.bin, <tag>.h.bin, <tag>.steps.bin (float64), <tag>.done.bin (int8) and <tag>.count.bin
// (int32), or, when it throws, <tag>.error.txt ("<kind>\n<message>", kind invalid_argument, runtime_error or c10).
// The program ends by writing runtimes.txt: the libamdhip64 images mapped into the process. Exit status 3: the
// constructor threw (its message on stderr); 0 otherwise.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "maxichem.hpp"

namespace {

std::string slurp(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("cannot read " + path);
  return std::string(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
}

// A raw file as a host tensor of the given shape and type; its size must match.
torch::Tensor raw(const std::string& path, torch::IntArrayRef shape, torch::Dtype type) {
  const std::string bytes = slurp(path);
  torch::Tensor t = torch::empty(shape, torch::TensorOptions().dtype(type));
  if (bytes.size() != static_cast<size_t>(t.nbytes())) throw std::runtime_error(path + ": unexpected size");
  std::copy(bytes.begin(), bytes.end(), static_cast<char*>(t.data_ptr()));
  return t;
}

void dump(const std::string& path, const torch::Tensor& t) {
  const torch::Tensor h = t.cpu().contiguous();
  std::ofstream out(path, std::ios::binary);
  out.write(static_cast<const char*>(h.data_ptr()), static_cast<std::streamsize>(h.nbytes()));
}

void note(const std::string& path, const std::string& text) { std::ofstream(path) << text; }

int mapped_hip_runtimes() {
  std::ifstream maps("/proc/self/maps");
  std::set<std::string> images;
  for (std::string line; std::getline(maps, line);) {
    const auto at = line.find("libamdhip64");
    if (at != std::string::npos) images.insert(line.substr(line.find('/')));
  }
  return static_cast<int>(images.size());
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 4) {
    std::cerr << "usage: cpp_caller <export folder> <case folder> <output folder>\n";
    return 2;
  }
  const std::string exported = argv[1], folder = std::string(argv[2]) + "/", out = std::string(argv[3]) + "/";
  const torch::Device gpu(torch::kCUDA, 0);
  std::unique_ptr<maxichem::Solver> chem;
  try {
    chem = std::make_unique<maxichem::Solver>(exported, gpu);
  } catch (const std::runtime_error& e) {
    std::cerr << e.what() << "\n";
    return 3;
  }
  std::cout << "opened: " << chem->ns() << " species, " << chem->nj() << " branches" << std::endl;
  const int64_t cells = std::stoll(slurp(folder + "B.txt"));
  const auto f64 = torch::kFloat64;
  std::vector<torch::Tensor> in;
  for (const char* name : {"x", "T", "P", "dt", "h"})
    in.push_back(raw(folder + name + ".bin", name[0] == 'x' ? torch::IntArrayRef{cells, chem->ns()}
                                                             : torch::IntArrayRef{cells},
                     f64)
                     .to(gpu));
  const torch::Tensor J = chem->nj() ? raw(folder + "J.bin", {cells, chem->nj()}, f64).to(gpu) : torch::Tensor();
  std::istringstream plan(slurp(folder + "plan.txt"));
  for (std::string line; std::getline(plan, line);) {
    std::istringstream words(line);
    std::string step, tag, steps, j;
    words >> step;
    if (step == "call") {
      words >> tag >> steps >> j;
      torch::Tensor jx = j == "given" ? J
                         : j == "empty" ? torch::zeros({cells, 0}, in[0].options())
                         : j == "columns" ? torch::zeros({cells, 1}, in[0].options())
                                          : torch::Tensor();
      std::string kind, message;
      try {
        const maxichem::CallResult r =
            steps == "default" ? (*chem)(in[0], in[1], in[2], in[3], in[4], jx)
                               : (*chem)(in[0], in[1], in[2], in[3], in[4], jx, maxichem::kDefaultRtol, -1.0,
                                         std::stoi(steps));
        const auto& [x, h, done, accepted] = r;  // field order as the Python CallResult
        dump(out + tag + ".x.bin", x);
        dump(out + tag + ".h.bin", h);
        dump(out + tag + ".steps.bin", accepted);
        dump(out + tag + ".done.bin", done.to(torch::kInt8));
        dump(out + tag + ".count.bin", chem->count());
      } catch (const std::invalid_argument& e) {
        kind = "invalid_argument", message = e.what();
      } catch (const c10::Error& e) {
        kind = "c10", message = e.what_without_backtrace();
      } catch (const std::runtime_error& e) {
        kind = "runtime_error", message = e.what();
      }
      if (!kind.empty()) {
        note(out + tag + ".error.txt", kind + "\n" + message);
        dump(out + tag + ".count.bin", chem->count());
      }
      std::cout << "call " << tag << ": " << (kind.empty() ? "returned" : "threw " + kind) << std::endl;
    } else if (step == "scramble") {
      chem->count().resize_({1});
      chem->total().zero_();
      chem->atoms().fill_(-1.0);
      std::cout << "scrambled the accessors' tensors" << std::endl;
    } else if (step == "state") {
      words >> tag;
      dump(out + tag + ".count.bin", chem->count());
      dump(out + tag + ".total.bin", chem->total());
      dump(out + tag + ".atoms.bin", chem->atoms());
      std::cout << "state " << tag << " written" << std::endl;
    } else if (!step.empty()) {
      std::cerr << "unknown plan step: " << line << "\n";
      return 2;
    }
  }
  note(out + "runtimes.txt", std::to_string(mapped_hip_runtimes()));
  return 0;
}
