// jit.h -- the network-specialised kernels (jit.cpp): the per-reaction
// description the source generator takes, the loaded kernel handles, and the
// compile / occupancy / launch entry points used by the C ABI.
#ifndef REBOP_JIT_H
#define REBOP_JIT_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "ssa_layout.h"

namespace rebop {

struct JitReaction {
  double k;
  std::vector<std::pair<uint32_t, uint32_t>> factors; // (species, c): factor = x[s] - c
  std::vector<int64_t> diff;                          // dense [S]
  // Rate::Expr: postfix tokens (rebop_expr.h) and their constants; empty for LMA
  std::vector<uint32_t> expr;
  std::vector<double> consts;
  // define_system reaction: reactant terms (species, multiplicity) in written order
  std::vector<std::pair<uint32_t, uint32_t>> terms;
};

struct JitKernel {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  hipFunction_t fn_sweep = nullptr;
  hipFunction_t fn_g2 = nullptr; // 2 lanes per trajectory (plain launches)
  hipFunction_t fn_g4 = nullptr; // 4 lanes per trajectory
};

// Kernel source for a network of S species and the given reactions; with
// `system`, the define_system arithmetic (integer mass-action products over
// `terms`, first-match choice); with `cap`, the loops check the draw cap.
std::string jit_source(uint32_t S, const std::vector<JitReaction> &rx, bool system = false,
                       bool cap = false);
// Compiles (or finds in the cache) and loads the kernel on the current device.
int jit_get(const std::string &src, JitKernel *out, std::string *err);
// Compiles for gfx950 without loading (no device needed).
int jit_compile_only(const std::string &src, size_t *code_size, std::string *err);
// Dynamic LDS bytes a launch of `run` takes (the LDS copy of the sample times).
size_t jit_dynamic_lds(const rebop_dev_run &run);
hipError_t jit_occupancy(const JitKernel &k, int threads, size_t lds, int *blocks_per_cu);
// group = lanes per trajectory (1, 2 or 4; 2 and 4 only for plain, non-sweep
// grid/target runs)
hipError_t jit_launch(const JitKernel &k, const rebop_dev_run &run, int blocks, int threads,
                      hipStream_t stream, int group = 1);

} // namespace rebop

#endif // REBOP_JIT_H
