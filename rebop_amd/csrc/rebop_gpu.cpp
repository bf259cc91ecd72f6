// rebop_gpu.cpp -- host side of the C ABI declared in include/rebop_gpu.h.
//
// Owns the network description (the Rust `gillespie::Gillespie` fields,
// src/gillespie.rs:157-163), packs it for the device (ssa_layout.h), chooses
// a kernel and launches it.  All device work of one call is enqueued on one
// HIP stream.
#include "../../include/rebop_gpu.h"

#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <utility>
#include <vector>

#include "jit.h"
#include "rebop_expr.h"
#include "scratch.h"
#include "ssa_layout.h"

namespace rebop {
int reg_shape_count();
int reg_shape_ms(int c);
int reg_shape_mr(int c);
hipError_t reg_occupancy(int c, int threads, size_t lds, int *blocks_per_cu);
hipError_t launch_reset_counters(unsigned long long *counters, hipStream_t stream);
hipError_t launch_zero_u64(unsigned long long *p, uint64_t n, hipStream_t stream);
hipError_t launch_grid(double *grid, const rebop_dev_run &run, hipStream_t stream);
hipError_t launch_raw_counts(int64_t *out, uint64_t n, int cus, hipStream_t stream);
hipError_t launch_reg(int c, const rebop_dev_net &net, const rebop_dev_run &run, int blocks,
                      int threads, size_t lds, hipStream_t stream);

int lds_checkpoints();
hipError_t lds_occupancy(int threads, size_t lds, int width, bool expr, bool grec,
                         int *blocks_per_cu);
hipError_t launch_lds(const rebop_dev_job *jobs, const rebop_dev_item *items, uint32_t n_items,
                      unsigned long long *counters, unsigned long long *job_events,
                      unsigned long long *job_queue, int blocks, int threads, size_t lds,
                      uint32_t rec_word_off, int width, bool expr, bool grec, bool rerun,
                      hipStream_t stream);

} // namespace rebop

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(REBOP_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int kThreads = 256;

struct DeviceInfo {
  bool probed = false;
  int status = REBOP_OK;
  std::string error;
  int device = -1;
  int cus = 0;
  size_t lds_max = 65536; // LDS bytes a workgroup may allocate
  std::string name;
};

DeviceInfo g_dev;

// The library runs only on gfx950; anything else is an error, never a
// silent fallback.
int ensure_device() {
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess) {
    return fail(REBOP_ERR_DEVICE, "no HIP device available");
  }
  if (g_dev.probed && g_dev.device == cur) {
    if (g_dev.status != REBOP_OK) return fail(g_dev.status, g_dev.error);
    return REBOP_OK;
  }
  g_dev = DeviceInfo();
  g_dev.probed = true;
  g_dev.device = cur;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, cur) != hipSuccess) {
    g_dev.status = REBOP_ERR_DEVICE;
    g_dev.error = "hipGetDeviceProperties failed";
    return fail(g_dev.status, g_dev.error);
  }
  g_dev.cus = prop.multiProcessorCount;
  if (prop.sharedMemPerBlock > 0) g_dev.lds_max = std::min<size_t>(prop.sharedMemPerBlock, 160 * 1024);
  g_dev.name = prop.gcnArchName;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    g_dev.status = REBOP_ERR_DEVICE;
    g_dev.error = std::string("device ") + prop.gcnArchName +
                  " is not gfx950; this library carries gfx950 code only";
    return fail(g_dev.status, g_dev.error);
  }
  return REBOP_OK;
}

struct Reaction {
  double k;
  // (species, order) in evaluation order; for a define_system network the
  // reactant terms (species, multiplicity) in written order
  std::vector<std::pair<uint32_t, uint32_t>> entries;
  std::vector<int64_t> diff;                          // dense [S]
  std::vector<uint32_t> expr;                         // Rate::Expr postfix tokens (else empty)
  std::vector<double> consts;                         // the expression's constants
};

// device buffer owned by RAII
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t ensure(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    release();
    hipError_t e = hipMalloc(&p, n ? n : 8);
    if (e == hipSuccess) bytes = n ? n : 8;
    return e;
  }
};

// The two timing events of one launch sequence, destroyed on every return path.
struct EventPair {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  EventPair() = default;
  EventPair(const EventPair &) = delete;
  EventPair &operator=(const EventPair &) = delete;
  ~EventPair() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
  hipError_t create() {
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    return e;
  }
};

// Device tables of rebop_run_jobs_device, one set per (device, stream).  The
// map is never destroyed: its buffers live until the process ends (freeing
// them from a static destructor could run after the HIP runtime is gone).
struct JobBuffers {
  DevBuf counters, djobs, ditems, dev_ev, dev_q;
};
std::mutex g_job_mu;
std::map<std::pair<int, hipStream_t>, JobBuffers> *g_job_bufs =
    new std::map<std::pair<int, hipStream_t>, JobBuffers>();

// finish_stats: a compact or narrow LDS launch saw a count beyond int16 /
// int32 (rerun it with the next wider counts)
constexpr int kRetryWide = -1;
// finish_stats: a register or specialised launch saw a count beyond the f64
// counts' exact range (rerun it on the LDS kernel's int64 counts)
constexpr int kRetryExact = -2;

// Draws per trajectory after which a run whose horizon is not finite (tmax
// = +inf or NaN) stops a trajectory that never reaches an absorbing state,
// when the handle sets no cap of its own.
constexpr uint64_t kInfiniteHorizonCap = 1ull << 20;

} // namespace

struct rebop_gillespie {
  uint32_t S = 0;
  bool sparse = false;
  uint64_t seed = 0;
  uint64_t draw_cap = 0; // watchdog: draws per trajectory and call (0 = none)
  std::vector<int64_t> x0;
  std::vector<Reaction> rx;
  uint32_t n_expr = 0; // reactions with a rate expression (specialised kernels only)
  bool system = false; // define_system semantics (specialised kernels only)
  // packed device copy of the network, rebuilt when reactions change
  bool dirty = true;
  int dev_id = -1;
  DevBuf dnet;
  DevBuf dx0;
  rebop_dev_net view{};
  uint32_t F = 0;
  DevBuf counters;
  // the autotuner's choice for AUTO launches (kernel < 0: not tuned)
  int tune_kernel = -1;
  int tune_threads = 256;
  double tune_occ = 1.0;
  // the network part of an LDS-kernel job (sparse jump, checkpoint blocks)
  rebop_dev_job job{};
  DevBuf djobs, ditems;
  // network-specialised kernels (run-time compiled), rebuilt when reactions
  // change: [0] plain, [1] with the draw-cap checks (compiled when first needed)
  bool jit_dirty[2] = {true, true};
  int jit_dev[2] = {-1, -1};
  rebop::JitKernel jitk[2]{};
  std::string jit_err;
  // written-out state of the trajectories a persistent launch hands to its tail
  DevBuf spill;
  // sample times of the current launch (register and specialised kernels)
  DevBuf grid;
};

struct rebop_batch {
  const rebop_gillespie *g = nullptr;
  uint64_t traj_begin = 0, n = 0;
  DevBuf x, t, draw;
};

static_assert(REBOP_EXPR_CONST == REBOP_OP_CONST && REBOP_EXPR_SPECIES == REBOP_OP_SPECIES &&
                  REBOP_EXPR_NEG == REBOP_OP_NEG && REBOP_EXPR_ADD == REBOP_OP_ADD &&
                  REBOP_EXPR_SUB == REBOP_OP_SUB && REBOP_EXPR_MUL == REBOP_OP_MUL &&
                  REBOP_EXPR_DIV == REBOP_OP_DIV && REBOP_EXPR_POW == REBOP_OP_POW &&
                  REBOP_EXPR_MAX == REBOP_OP_MAX && REBOP_EXPR_MIN == REBOP_OP_MIN &&
                  REBOP_EXPR_EXP == REBOP_OP_EXP,
              "public expression op codes match the kernels'");

namespace {

int check_x0(const int64_t *x, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    const int64_t v = x[i];
    if (v > (int64_t)REBOP_EXACT_LIMIT || v < -(int64_t)REBOP_EXACT_LIMIT)
      return fail(REBOP_ERR_VALUE, "initial count " + std::to_string(v) +
                                       " is beyond the exactly simulated range |n| <= 2^52");
  }
  return REBOP_OK;
}

// Packs the network into one device buffer (layout in ssa_layout.h).
int upload_net(rebop_gillespie *g) {
  int cur = -1;
  HIP_TRY(hipGetDevice(&cur));
  if (!g->dirty && g->dev_id == cur) return REBOP_OK;
  const uint32_t S = g->S, R = (uint32_t)g->rx.size();
  std::vector<double> k(R);
  std::vector<uint32_t> fb(R + 1, 0), fs;
  std::vector<double> fc, diff((size_t)R * S);
  for (uint32_t r = 0; r < R; ++r) {
    const Reaction &x = g->rx[r];
    k[r] = x.k;
    for (auto &e : x.entries) {
      // Rate::rate multiplies (n+1-e), ..., n: factor x[s] - c, c = e-1 .. 0
      for (uint32_t j = 0; j < e.second; ++j) {
        fs.push_back(e.first);
        fc.push_back((double)(e.second - 1u - j));
      }
      if (fs.size() > (1u << 24))
        return fail(REBOP_ERR_UNSUPPORTED, "more than 2^24 propensity factors in total");
    }
    fb[r + 1] = (uint32_t)fs.size();
    for (uint32_t s = 0; s < S; ++s) diff[(size_t)r * S + s] = (double)x.diff[s];
  }
  const uint32_t F = (uint32_t)fs.size();
  // sparse jump (Jump::sparse, gillespie.rs:126-140) and whether the
  // cumulative sum is guaranteed non-decreasing (all rate constants >= 0 and
  // every consumed species appears once in the rate with order >= its use)
  std::vector<uint32_t> jb(R + 1, 0), js;
  std::vector<int32_t> jd;
  bool linear = false;
  for (uint32_t r = 0; r < R; ++r) {
    const Reaction &x = g->rx[r];
    if (x.k < 0.0) linear = true;
    for (uint32_t sp = 0; sp < S; ++sp) {
      if (x.diff[sp] == 0) continue;
      js.push_back(sp);
      jd.push_back((int32_t)x.diff[sp]);
      if (x.diff[sp] < 0) {
        uint32_t seen = 0, order = 0;
        for (auto &e : x.entries)
          if (e.first == sp) {
            seen++;
            order = e.second;
          }
        if (seen != 1 || (int64_t)order < -x.diff[sp]) linear = true;
      }
    }
    jb[r + 1] = (uint32_t)js.size();
  }
  const uint32_t J = (uint32_t)js.size();
  // Rate::Expr programs for the LDS kernel's interpreter: one token array with
  // per-reaction offsets, constant tokens rebased onto one constant array
  std::vector<uint32_t> eb(R + 1, 0), ecode;
  std::vector<double> econst;
  for (uint32_t r = 0; r < R; ++r) {
    const Reaction &x = g->rx[r];
    const uint32_t cbase = (uint32_t)econst.size();
    for (uint32_t t : x.expr) {
      const uint32_t op = REBOP_EXPR_OPCODE(t), arg = REBOP_EXPR_ARG(t);
      ecode.push_back(op == REBOP_OP_CONST ? REBOP_EXPR_TOKEN(op, cbase + arg) : t);
    }
    econst.insert(econst.end(), x.consts.begin(), x.consts.end());
    if (econst.size() >= (1u << 24))
      return fail(REBOP_ERR_UNSUPPORTED, "more than 2^24 rate-expression constants in total");
    eb[r + 1] = (uint32_t)ecode.size();
    if (!x.expr.empty()) linear = true; // an expression may be negative or not monotone
  }
  const uint32_t EC = (uint32_t)ecode.size(), EK = (uint32_t)econst.size();
  // define_system reactant terms (species, multiplicity) in written order
  std::vector<uint32_t> tbeg(R + 1, 0), tsp, tmul;
  if (g->system) {
    for (uint32_t r = 0; r < R; ++r) {
      const Reaction &x = g->rx[r];
      for (auto &e : x.entries) {
        tsp.push_back(e.first);
        tmul.push_back(e.second);
      }
      tbeg[r + 1] = (uint32_t)tsp.size();
    }
    // a NaN or negative rate, or a term product that wraps, can make the sum
    // non-monotone: the choice scans linearly (it is exact either way)
    linear = true;
  }
  const uint32_t TS = (uint32_t)tsp.size();
  // LDS-kernel reaction records (layout in ssa_layout.h)
  std::vector<uint32_t> rec((size_t)REBOP_REC_WORDS * R, 0u);
  std::vector<uint32_t> jrec((size_t)REBOP_JREC_WORDS * R, 0u);
  bool any_slow = false;
  for (uint32_t r = 0; r < R; ++r) {
    uint32_t *w = rec.data() + (size_t)REBOP_REC_WORDS * r;
    uint64_t kb;
    std::memcpy(&kb, &k[r], sizeof kb);
    w[0] = (uint32_t)kb;
    w[1] = (uint32_t)(kb >> 32);
    const uint32_t nf = fb[r + 1] - fb[r];
    // an expression (the interpreter) or a define_system product: the slow path
    bool slow_f = nf > 2 || !g->rx[r].expr.empty() || g->system;
    for (uint32_t q = 0; q < nf && q < 2; ++q) {
      const uint32_t f = fb[r] + q;
      const uint32_t c = (uint32_t)fc[f];
      if (c >= 4096u || fs[f] >= 65535u) slow_f = true;
      // the factor's count row: species s at row s + 1 (row 0: ones)
      w[2 + q] = ((fs[f] + 1u) << 16) | (c & 0xfffu) | REBOP_REC_PRESENT;
    }
    if (slow_f) {
      w[3] |= REBOP_REC_SLOW_FACTORS;
      any_slow = true;
    }
    const uint32_t nj = jb[r + 1] - jb[r];
    bool slow_j = nj > 4;
    for (uint32_t e = jb[r]; e < jb[r + 1]; ++e) // species in 8 bits, difference in 24
      if (js[e] >= 256u || jd[e] < -(1 << 23) || jd[e] >= (1 << 23) || jd[e] == 0) slow_j = true;
    if (slow_j) {
      w[3] |= REBOP_REC_SLOW_JUMP;
    } else {
      uint32_t *jw = jrec.data() + (size_t)REBOP_JREC_WORDS * r;
      for (uint32_t q = 0; q < nj; ++q) {
        const uint32_t e = jb[r] + q;
        jw[q] = (js[e] & 0xffu) | ((uint32_t)jd[e] << 8);
      }
    }
  }
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t ok = 0;
  const size_t ofb = ok + al(sizeof(double) * R);
  const size_t ofs = ofb + al(sizeof(uint32_t) * (R + 1));
  const size_t ofc = ofs + al(sizeof(uint32_t) * F);
  const size_t odf = ofc + al(sizeof(double) * F);
  const size_t ojb = odf + al(sizeof(double) * R * S);
  const size_t ojs = ojb + al(sizeof(uint32_t) * (R + 1));
  const size_t ojd = ojs + al(sizeof(uint32_t) * J);
  const size_t orc = ojd + al(sizeof(int32_t) * J);
  // four records of padding after the last: the LDS kernel's pass reads the
  // next group of four ahead (never using what lies past R - 1)
  const size_t ojr = orc + al(sizeof(uint32_t) * (rec.size() + 4u * REBOP_REC_WORDS));
  const size_t oeb = ojr + al(sizeof(uint32_t) * jrec.size());
  const size_t oec = oeb + al(sizeof(uint32_t) * (R + 1));
  const size_t oek = oec + al(sizeof(uint32_t) * EC);
  const size_t otb = oek + al(sizeof(double) * EK);
  const size_t ots = otb + al(sizeof(uint32_t) * (R + 1));
  const size_t otm = ots + al(sizeof(uint32_t) * TS);
  const size_t total = otm + al(sizeof(uint32_t) * TS);
  std::vector<unsigned char> host(total, 0);
  std::memcpy(host.data() + ojb, jb.data(), sizeof(uint32_t) * (R + 1));
  if (J) std::memcpy(host.data() + ojs, js.data(), sizeof(uint32_t) * J);
  if (J) std::memcpy(host.data() + ojd, jd.data(), sizeof(int32_t) * J);
  if (R) std::memcpy(host.data() + orc, rec.data(), sizeof(uint32_t) * rec.size());
  if (R) std::memcpy(host.data() + ojr, jrec.data(), sizeof(uint32_t) * jrec.size());
  if (R) std::memcpy(host.data() + ok, k.data(), sizeof(double) * R);
  std::memcpy(host.data() + ofb, fb.data(), sizeof(uint32_t) * (R + 1));
  if (F) std::memcpy(host.data() + ofs, fs.data(), sizeof(uint32_t) * F);
  if (F) std::memcpy(host.data() + ofc, fc.data(), sizeof(double) * F);
  if (R * S) std::memcpy(host.data() + odf, diff.data(), sizeof(double) * R * S);
  if (EC) {
    std::memcpy(host.data() + oeb, eb.data(), sizeof(uint32_t) * (R + 1));
    std::memcpy(host.data() + oec, ecode.data(), sizeof(uint32_t) * EC);
  }
  if (EK) std::memcpy(host.data() + oek, econst.data(), sizeof(double) * EK);
  if (g->system) {
    std::memcpy(host.data() + otb, tbeg.data(), sizeof(uint32_t) * (R + 1));
    if (TS) {
      std::memcpy(host.data() + ots, tsp.data(), sizeof(uint32_t) * TS);
      std::memcpy(host.data() + otm, tmul.data(), sizeof(uint32_t) * TS);
    }
  }
  HIP_TRY(g->dnet.ensure(total));
  HIP_TRY(hipMemcpy(g->dnet.p, host.data(), total, hipMemcpyHostToDevice));
  HIP_TRY(g->dx0.ensure(sizeof(int64_t) * (S ? S : 1)));
  if (S) HIP_TRY(hipMemcpy(g->dx0.p, g->x0.data(), sizeof(int64_t) * S, hipMemcpyHostToDevice));
  HIP_TRY(g->counters.ensure(sizeof(unsigned long long) * REBOP_CTR_COUNT));
  unsigned char *base = (unsigned char *)g->dnet.p;
  g->view.S = S;
  g->view.R = R;
  g->view.F = F;
  g->view.k = (const double *)(base + ok);
  g->view.fac_begin = (const uint32_t *)(base + ofb);
  g->view.fac_s = (const uint32_t *)(base + ofs);
  g->view.fac_c = (const double *)(base + ofc);
  g->view.diff = (const double *)(base + odf);
  g->F = F;
  const uint32_t MB = (uint32_t)rebop::lds_checkpoints();
  rebop_dev_job jt{};
  jt.S = S;
  jt.R = R;
  // checkpoint blocks of a multiple of 4 reactions (the LDS kernel's pass
  // handles reactions in groups of 4); any block size gives the same sums
  jt.block = R ? ((R + MB - 1) / MB + 3u) & ~3u : 4;
  jt.nblocks = R ? (R + jt.block - 1) / jt.block : 0;
  jt.flags = (linear ? REBOP_JOB_LINEAR : 0u) | (any_slow ? REBOP_JOB_SLOW : 0u) |
             (g->system ? REBOP_JOB_SYSTEM : 0u);
  jt.k = g->view.k;
  jt.fac_begin = g->view.fac_begin;
  jt.fac_s = g->view.fac_s;
  jt.fac_c = g->view.fac_c;
  jt.jump_begin = (const uint32_t *)(base + ojb);
  jt.jump_s = (const uint32_t *)(base + ojs);
  jt.jump_d = (const int32_t *)(base + ojd);
  jt.rec = (const uint32_t *)(base + orc);
  jt.jrec = (const uint32_t *)(base + ojr);
  if (EC) {
    jt.expr_begin = (const uint32_t *)(base + oeb);
    jt.expr_code = (const uint32_t *)(base + oec);
    jt.expr_const = (const double *)(base + oek);
  }
  if (g->system) {
    jt.sys_begin = (const uint32_t *)(base + otb);
    jt.sys_species = (const uint32_t *)(base + ots);
    jt.sys_mult = (const uint32_t *)(base + otm);
  }
  g->job = jt;
  g->dirty = false;
  g->dev_id = cur;
  return REBOP_OK;
}

// The draw cap of a run of this handle up to tmax: the handle's own, else
// one for horizons that are not finite (where the reference's loop would only
// end in an absorbing state), else none.
uint64_t effective_cap(const rebop_gillespie *g, double tmax) {
  if (g->draw_cap) return g->draw_cap;
  return std::isfinite(tmax) ? 0ull : kInfiniteHorizonCap;
}

int pick_reg_shape(uint32_t S, uint32_t R) {
  for (int c = 0; c < rebop::reg_shape_count(); ++c)
    if ((int)S <= rebop::reg_shape_ms(c) && (int)R <= rebop::reg_shape_mr(c)) return c;
  return -1;
}

// Networks small enough to compile into a kernel (registers hold x[S] and cum[R]).
bool jit_eligible(const rebop_gillespie *g) {
  if (g->S > 32 || g->rx.size() > 64) return false;
  size_t F = 0;
  for (auto &r : g->rx) {
    for (auto &e : r.entries) F += e.second + (g->system ? 1u : 0u);
    F += r.expr.size();
  }
  return F <= 4096;
}

// Networks that a run puts on a specialised kernel whatever the kernel kind
// (other than LDS): define_system arithmetic and rate expressions are
// compiled into it when the network fits one (larger ones run on the LDS
// kernel); the data-driven register kernels take neither.
bool specialised_only(const rebop_gillespie *g) {
  return (g->system || g->n_expr) && jit_eligible(g);
}

// Lanes at or below which a wave of a persistent specialised launch hands
// its trajectories to the grouped tail launch once the queue is empty
// (REBOP_TAIL_LANES overrides 16; 0 turns the hand-off off), and the lanes
// per trajectory of that launch (REBOP_TAIL_GROUP: 2 or 4).
uint32_t tail_lanes() {
  const char *v = std::getenv("REBOP_TAIL_LANES");
  if (!v || !*v) return 16u;
  const long n = std::strtol(v, nullptr, 10);
  // below 64: a wave is thin only once some lane left, i.e. the queue is empty
  return n <= 0 ? 0u : n >= 63 ? 63u : (uint32_t)n;
}
int tail_group() {
  const char *v = std::getenv("REBOP_TAIL_GROUP");
  return v && std::strcmp(v, "2") == 0 ? 2 : 4;
}

// The autotuner's on-disk cache: one line per tuned shape,
//   <key> <kernel> <threads> <occupancy> <ms>
// keyed by the network shape (species, reactions, total reactant order,
// specialised-only arithmetic), nb_steps, the run size (power of two) and the
// device.  A cached choice only changes speed: every variant gives the same
// results.  REBOP_TUNE_CACHE names the file (default
// $XDG_CACHE_HOME/rebop_amd/tune_cache.txt or ~/.cache/...); an empty value
// turns the cache off; REBOP_TUNE_FORCE=1 re-times and rewrites the entry.
std::string tune_cache_path() {
  if (const char *v = std::getenv("REBOP_TUNE_CACHE")) return v;
  std::string base;
  if (const char *x = std::getenv("XDG_CACHE_HOME"); x && *x)
    base = x;
  else if (const char *h = std::getenv("HOME"); h && *h)
    base = std::string(h) + "/.cache";
  else
    return "";
  return base + "/rebop_amd/tune_cache.txt";
}

std::string tune_key(const rebop_gillespie *g, uint64_t nb_steps, uint64_t n_pilot,
                     const std::string &device) {
  uint64_t order = 0;
  for (auto &r : g->rx)
    for (auto &e : r.entries) order += e.second;
  int lg = 0;
  while ((1ull << (lg + 1)) <= n_pilot && lg < 62) ++lg;
  char buf[256];
  std::snprintf(buf, sizeof buf, "S%u-R%zu-O%llu-E%u-Y%d-nb%llu-n2^%d-%s", g->S, g->rx.size(),
                (unsigned long long)order, g->n_expr, g->system ? 1 : 0,
                (unsigned long long)nb_steps, lg, device.c_str());
  return buf;
}

bool tune_cache_get(const std::string &path, const std::string &key, int *kernel, int *threads,
                    double *occ, double *ms) {
  if (path.empty()) return false;
  FILE *f = std::fopen(path.c_str(), "r");
  if (!f) return false;
  char k[256];
  int kk = 0, th = 0;
  double oc = 0, t = 0;
  bool found = false;
  while (std::fscanf(f, "%255s %d %d %lf %lf", k, &kk, &th, &oc, &t) == 5)
    if (key == k) { // the last entry for the key wins
      *kernel = kk;
      *threads = th;
      *occ = oc;
      *ms = t;
      found = true;
    }
  std::fclose(f);
  return found;
}

void tune_cache_put(const std::string &path, const std::string &key, int kernel, int threads,
                    double occ, double ms) {
  if (path.empty()) return;
  const size_t slash = path.rfind('/');
  if (slash != std::string::npos) {
    // create the directory chain (failures are ignored: the cache is optional)
    for (size_t i = 1; i <= slash; ++i)
      if (path[i] == '/' || i == slash) (void)mkdir(path.substr(0, i == slash ? slash : i).c_str(), 0755);
  }
  if (FILE *f = std::fopen(path.c_str(), "a")) {
    std::fprintf(f, "%s %d %d %.6f %.6f\n", key.c_str(), kernel, threads, occ, ms);
    std::fclose(f);
  }
}

bool jit_disabled() {
  const char *v = std::getenv("REBOP_NO_SPECIALIZE");
  return v && *v && std::strcmp(v, "0") != 0;
}

std::string jit_source_of(const rebop_gillespie *g, bool cap = false) {
  std::vector<rebop::JitReaction> jr;
  jr.reserve(g->rx.size());
  for (auto &r : g->rx) {
    rebop::JitReaction j;
    j.k = r.k;
    if (g->system)
      j.terms = r.entries;
    else
      for (auto &e : r.entries)
        for (uint32_t q = 0; q < e.second; ++q) j.factors.emplace_back(e.first, e.second - 1u - q);
    j.diff = r.diff;
    j.expr = r.expr;
    j.consts = r.consts;
    jr.push_back(std::move(j));
  }
  std::string src = rebop::jit_source(g->S, jr, g->system, cap);
  // REBOP_DUMP_SPECIALIZED=<file>: write the generated kernel source there
  if (const char *path = std::getenv("REBOP_DUMP_SPECIALIZED")) {
    if (FILE *f = std::fopen(path, "w")) {
      std::fputs(src.c_str(), f);
      std::fclose(f);
    }
  }
  return src;
}

// Is the specialised kernel (cap: its variant with the draw-cap checks) loaded
// on device `dev` for the handle's current network?
bool jit_ready(const rebop_gillespie *g, bool cap, int dev) {
  const int v = cap ? 1 : 0;
  return !g->jit_dirty[v] && g->jit_dev[v] == dev && g->jitk[v].fn;
}

// Compiles (cached) and loads the network-specialised kernel on the current device.
int prepare_jit(rebop_gillespie *g, bool cap = false) {
  int cur = -1;
  HIP_TRY(hipGetDevice(&cur));
  if (jit_ready(g, cap, cur)) return REBOP_OK;
  const int v = cap ? 1 : 0;
  std::string err;
  if (rebop::jit_get(jit_source_of(g, cap), &g->jitk[v], &err) != 0) {
    g->jit_err = err;
    return fail(REBOP_ERR_DEVICE, "specialised kernel: " + err);
  }
  g->jit_dirty[v] = false;
  g->jit_dev[v] = cur;
  return REBOP_OK;
}

void mark_jit_dirty(rebop_gillespie *g) {
  g->jit_dirty[0] = true;
  g->jit_dirty[1] = true;
}

constexpr int kLdsThreads = 64;

// Items of 64 trajectories per job, largest expected cost first.
std::vector<rebop_dev_item> make_items(const std::vector<rebop_dev_job> &jobs,
                                       const std::vector<double> &cost) {
  std::vector<uint32_t> order(jobs.size());
  for (uint32_t j = 0; j < order.size(); ++j) order[j] = j;
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
  std::vector<rebop_dev_item> items;
  for (uint32_t j : order)
    for (uint64_t f = 0; f < jobs[j].n_traj; f += 64) {
      rebop_dev_item it;
      it.job = j;
      it.count = (uint32_t)std::min<uint64_t>(64, jobs[j].n_traj - f);
      it.first = f;
      items.push_back(it);
    }
  return items;
}

// Where enqueue_lds puts the job and item tables: growable buffers, or a
// fixed region of caller scratch (which must be large enough).
struct TableDest {
  DevBuf *jobs = nullptr, *items = nullptr;
  void *fixed_jobs = nullptr, *fixed_items = nullptr;
  size_t jobs_cap = 0, items_cap = 0; // bytes of the fixed regions
};

// Uploads the job and item tables and enqueues the LDS kernel.
// job_queue: one zeroed 64-bit trajectory counter per job.
// REBOP_LDS_COMPACT: 0 never starts the LDS kernel with int16 counts, 1
// does whenever the initial counts fit, unset (-1) where that runs more
// waves per CU (compact_pays).
int lds_compact_mode() {
  const char *e = std::getenv("REBOP_LDS_COMPACT");
  if (!e || !*e) return -1;
  return std::strcmp(e, "0") == 0 ? 0 : 1;
}

// The LDS kernel's count width (bytes) a launch starts with: int64 when
// asked (REBOP_KERNEL_WIDE), for a resumable batch (whose state a rerun
// could not restore) or when the shared initial counts `x0` are beyond
// int32; int16 when they fit it and a launch that meets a larger count can
// be rerun (`rerun`: its counters are read) -- unless REBOP_LDS_COMPACT=0;
// int32 otherwise (device-side initial counts, x0 NULL, included).  The
// callers keep int16 only where it runs more waves (compact_pays).
int lds_start_width(int kernel, uint32_t mode, bool rerun, const std::vector<int64_t> *x0) {
  if ((kernel & REBOP_KERNEL_WIDE) || mode == REBOP_MODE_TARGET) return 8;
  if (!x0) return 4;
  int64_t m = 0;
  for (int64_t v : *x0) {
    if (v > REBOP_LDS_LIMIT || v < -REBOP_LDS_LIMIT) return 8;
    m = std::max(m, v < 0 ? -v : v);
  }
  return rerun && lds_compact_mode() != 0 && m <= REBOP_LDS16_LIMIT ? 2 : 4;
}

// rebop_stats.kernel of an LDS launch of a count width
int lds_kind(int width) {
  return REBOP_KERNEL_LDS | (width == 8 ? REBOP_KERNEL_WIDE : width == 2 ? REBOP_KERNEL_COMPACT : 0);
}

// The LDS of one workgroup of an LDS launch: counts [1 + S_max][64] of
// `width` bytes (row 0 holds ones, for absent factors; 16-byte aligned),
// then the reaction records and four of padding, unless they are read in
// place (grec).
// `xbytes`: the counts' part.
size_t lds_bytes(uint32_t smax, uint32_t rmax, int width, bool grec, size_t *xbytes = nullptr) {
  const size_t xb = ((size_t)(smax + 1u) * kLdsThreads * (size_t)width + 15) & ~(size_t)15;
  if (xbytes) *xbytes = xb;
  return xb + (grec ? 0u : (size_t)(rmax + 4u) * REBOP_REC_WORDS * sizeof(uint32_t));
}

// Workgroups (waves) of the LDS kernel per CU for a shape, or 0 when its LDS
// does not fit.
int lds_per_cu(uint32_t smax, uint32_t rmax, int width, bool expr, bool grec) {
  if (expr && grec) return 0;
  const size_t lds = lds_bytes(smax, rmax, width, grec);
  int n = 0;
  if (lds > g_dev.lds_max ||
      rebop::lds_occupancy(kLdsThreads, lds, width, expr, grec, &n) != hipSuccess)
    return 0;
  return n;
}

// Whether the records are read in place: when the LDS they would take
// leaves fewer waves per CU (the registers bound most networks first, and
// there the LDS copy is faster).
bool lds_global_records(uint32_t smax, uint32_t rmax, int width, bool expr) {
  return lds_per_cu(smax, rmax, width, expr, true) > lds_per_cu(smax, rmax, width, expr, false);
}

// The largest species and reaction counts of a set of jobs, and whether one
// has rate expressions (the interpreting kernel).
void job_shape(const std::vector<rebop_dev_job> &jobs, uint32_t *smax, uint32_t *rmax, bool *expr) {
  *smax = 1;
  *rmax = 1;
  *expr = false;
  for (auto &j : jobs) {
    *smax = std::max(*smax, j.S);
    *rmax = std::max(*rmax, j.R);
    *expr = *expr || j.expr_begin != nullptr;
  }
}

// Whether int16 counts let more waves share a CU than int32 counts.  Only
// then are they worth the risk of a rerun: for most networks the registers,
// not the LDS, bound the waves per CU, and the two widths run alike.
bool compact_pays(const std::vector<rebop_dev_job> &jobs) {
  uint32_t smax, rmax;
  bool expr;
  job_shape(jobs, &smax, &rmax, &expr);
  auto per_cu = [&](int width) {
    return std::max(lds_per_cu(smax, rmax, width, expr, false),
                    lds_per_cu(smax, rmax, width, expr, true));
  };
  return lds_compact_mode() == 1 || per_cu(2) > per_cu(4);
}

int enqueue_lds(const std::vector<rebop_dev_job> &jobs, const std::vector<double> &cost,
                const TableDest &dst, unsigned long long *counters,
                unsigned long long *job_events, unsigned long long *job_queue, int width,
                bool rerun, hipStream_t stream) {
  uint32_t smax, rmax;
  bool expr;
  job_shape(jobs, &smax, &rmax, &expr);
  size_t xbytes = 0;
  const bool grec = lds_global_records(smax, rmax, width, expr);
  const size_t lds = lds_bytes(smax, rmax, width, grec, &xbytes);
  if (lds > g_dev.lds_max)
    return fail(REBOP_ERR_UNSUPPORTED,
                "LDS kernel: " + std::to_string(smax) + " species and " + std::to_string(rmax) +
                    " reactions need " + std::to_string(lds) + " bytes of LDS per workgroup (" +
                    std::to_string(g_dev.lds_max) + " available)");
  const std::vector<rebop_dev_item> items = make_items(jobs, cost);
  if (items.empty()) return REBOP_OK;
  if (items.size() > 0xffffffffull) return fail(REBOP_ERR_VALUE, "too many work items");
  const size_t jbytes = sizeof(rebop_dev_job) * jobs.size();
  const size_t ibytes = sizeof(rebop_dev_item) * items.size();
  void *pj = dst.fixed_jobs, *pi = dst.fixed_items;
  if (dst.jobs) {
    HIP_TRY(dst.jobs->ensure(jbytes));
    HIP_TRY(dst.items->ensure(ibytes));
    pj = dst.jobs->p;
    pi = dst.items->p;
  } else if (jbytes > dst.jobs_cap || ibytes > dst.items_cap) {
    return fail(REBOP_ERR_VALUE, "the LDS kernel's work items (" + std::to_string(items.size()) +
                                     ") exceed the scratch limit of " +
                                     std::to_string(dst.items_cap / sizeof(rebop_dev_item)));
  }
  // stream-ordered: a previous launch on this stream may still read the tables
  HIP_TRY(hipMemcpyAsync(pj, jobs.data(), jbytes, hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMemcpyAsync(pi, items.data(), ibytes, hipMemcpyHostToDevice, stream));
  int per_cu = 0;
  HIP_TRY(rebop::lds_occupancy(kLdsThreads, lds, width, expr, grec, &per_cu));
  if (per_cu < 1) per_cu = 1;
  const uint64_t waves_per_block = kLdsThreads / 64;
  uint64_t blocks = (uint64_t)per_cu * (uint64_t)g_dev.cus;
  const uint64_t need = (items.size() + waves_per_block - 1) / waves_per_block;
  if (blocks > need) blocks = need;
  HIP_TRY(rebop::launch_lds((const rebop_dev_job *)pj, (const rebop_dev_item *)pi,
                            (uint32_t)items.size(), counters, job_events, job_queue, (int)blocks,
                            kLdsThreads, lds, (uint32_t)(xbytes / sizeof(uint32_t)), width, expr,
                            grec, rerun && width < 8, stream));
  return REBOP_OK;
}

// Reads the counters of a finished launch into `stats` (waits for the stream).
int finish_stats(rebop_stats *stats, unsigned long long *counters, const EventPair &ev, int kind,
                 hipStream_t stream, uint64_t draw_cap) {
  unsigned long long ctr[REBOP_CTR_COUNT] = {0};
  HIP_TRY(hipMemcpyAsync(ctr, counters, sizeof ctr, hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  float ms = 0.f;
  if (ev.e0) HIP_TRY(hipEventElapsedTime(&ms, ev.e0, ev.e1));
  stats->events = ctr[REBOP_CTR_EVENTS];
  stats->draws = ctr[REBOP_CTR_DRAWS];
  stats->trajectories = ctr[REBOP_CTR_TRAJ];
  stats->kernel = (uint64_t)kind;
  stats->kernel_ms = ms;
  stats->tail = ctr[REBOP_CTR_SPILL];
  stats->capped = ctr[REBOP_CTR_CAPPED];
  if (ctr[REBOP_CTR_FLAGS] & REBOP_FLAG_INTERNAL)
    return fail(REBOP_ERR_DEVICE, "internal error: a kernel launch invariant did not hold");
  if (ctr[REBOP_CTR_FLAGS] & REBOP_FLAG_NARROW) return kRetryWide;
  if (ctr[REBOP_CTR_FLAGS] & REBOP_FLAG_EXACT) return kRetryExact;
  if (ctr[REBOP_CTR_FLAGS] & REBOP_FLAG_RANGE)
    return fail(REBOP_ERR_RANGE, "a count overflowed int64 (or, in a resumable batch, left "
                                 "|x| <= 2^52), or a total propensity was infinite");
  if (ctr[REBOP_CTR_FLAGS] & REBOP_FLAG_CAP)
    return fail(REBOP_ERR_CAP, std::to_string(ctr[REBOP_CTR_CAPPED]) +
                                   " trajectories reached the cap of " + std::to_string(draw_cap) +
                                   " draws (their results stop there)");
  return REBOP_OK;
}

// Enqueues one SSA launch (plus the counter reset) on `stream`.  When
// `stats` is set, synchronises and fills it.  With `sc`, the launch counters,
// sample-time table and LDS work tables live in that caller scratch instead
// of the handle's buffers, nothing is allocated (the handle must be prepared:
// network uploaded, kernel compiled), the tail hand-off is off (its results
// are the same) and the kernel is not timed.
int launch_run(rebop_gillespie *g, rebop_dev_run run, int kernel, hipStream_t stream,
               rebop_stats *stats, int threads = 0, double occ = 1.0,
               const rebop::Scratch *sc = nullptr) {
  const rebop_dev_run run_in = run; // for a rerun on the LDS kernel's int64 counts
  int rc = ensure_device();
  if (rc) return rc;
  if (kernel == REBOP_KERNEL_AUTO && g->tune_kernel >= 0 && threads == 0) {
    kernel = g->tune_kernel;
    threads = g->tune_threads;
    occ = g->tune_occ;
    // a grouped choice is kept only while the run still leaves at most one
    // wave per SIMD (the pilot may have been smaller than this run)
    const int grp = (kernel & REBOP_KERNEL_GROUP4) ? 4 : (kernel & REBOP_KERNEL_GROUP2) ? 2 : 1;
    if (grp > 1 && (run.n_traj * (uint64_t)grp > (uint64_t)g_dev.cus * 256u || run.rates ||
                    run.mode == REBOP_MODE_EVENTS))
      kernel = REBOP_KERNEL_PERSISTENT | REBOP_KERNEL_SPECIALIZED;
  }
  if (threads == 0) threads = kThreads;
  if (threads != 64 && threads != 128 && threads != 256)
    return fail(REBOP_ERR_VALUE, "workgroup size must be 64, 128 or 256");
  rc = upload_net(g);
  if (rc) return rc;
  run.draw_cap = effective_cap(g, run.tmax);
  const uint32_t S = g->S, R = (uint32_t)g->rx.size();
  // kernel choice: LDS when asked; the specialised kernel when asked or (AUTO)
  // when the network is small enough; else the data-driven register kernel
  // when it fits, else the LDS kernel
  int variant = kernel & 0xF;
  if ((g->system || g->n_expr) && variant != REBOP_KERNEL_LDS &&
      !(kernel & REBOP_KERNEL_SPECIALIZED)) {
    // define_system arithmetic and rate expressions: compiled into the
    // specialised kernel when the network fits one, else run by the LDS
    // kernel (never the data-driven register kernels)
    if (jit_eligible(g)) {
      kernel |= REBOP_KERNEL_SPECIALIZED;
    } else {
      kernel = (kernel & ~0xF) | REBOP_KERNEL_LDS;
      variant = REBOP_KERNEL_LDS;
    }
  }
  // shared initial counts beyond the f64 counts' exact range (|x| <= 2^52):
  // only the LDS kernel's int64 counts hold them
  if (!run.x0_div && !run.st_x)
    for (int64_t v : g->x0)
      if (v > (int64_t)REBOP_EXACT_LIMIT || v < -(int64_t)REBOP_EXACT_LIMIT) {
        kernel = REBOP_KERNEL_LDS | REBOP_KERNEL_WIDE;
        variant = REBOP_KERNEL_LDS;
        break;
      }
  const bool fits_regs = !g->n_expr && !g->system && pick_reg_shape(S, R) >= 0;
  bool spec = false, use_lds = variant == REBOP_KERNEL_LDS;
  if (!use_lds && (kernel & REBOP_KERNEL_SPECIALIZED)) {
    if (!jit_eligible(g))
      return fail(REBOP_ERR_UNSUPPORTED,
                  "network too large for a specialised kernel (S <= 32, R <= 64)");
    rc = prepare_jit(g, run.draw_cap != 0);
    if (rc) return rc;
    spec = true;
  } else if (!use_lds && variant == REBOP_KERNEL_AUTO) {
    spec = jit_eligible(g) && !jit_disabled() && prepare_jit(g, run.draw_cap != 0) == REBOP_OK;
    use_lds = !spec && !fits_regs;
  } else if (!use_lds && !fits_regs) {
    return fail(REBOP_ERR_UNSUPPORTED, "network with " + std::to_string(S) + " species and " +
                                           std::to_string(R) +
                                           " reactions exceeds the register kernels");
  }
  if (use_lds && run.rates)
    return fail(REBOP_ERR_UNSUPPORTED,
                "parameter sweeps run on the register and specialised kernels (S <= 32, R <= 64)");
  unsigned long long *const ctrs =
      sc ? (unsigned long long *)(sc->base + rebop::kScratchCounters)
         : (unsigned long long *)g->counters.p;
  if (use_lds) {
    // LDS kernel: one job, items of 64 trajectories
    rebop_dev_job job = g->job;
    job.seed = run.seed;
    job.traj_begin = run.traj_begin;
    job.n_traj = run.n_traj;
    job.tmax = run.tmax;
    job.nb_steps = run.nb_steps;
    job.mode = run.mode;
    job.draw_cap = run.draw_cap;
    job.x0_div = run.x0_div;
    job.x0 = run.x0;
    job.out = run.out;
    job.final_x = run.final_x;
    job.st_x = run.st_x;
    job.st_t = run.st_t;
    job.st_draw = run.st_draw;
    job.n_records = run.n_records;
    job.rec_offsets = run.rec_offsets;
    job.rec_times = run.rec_times;
    job.rec_states = run.rec_states;
    TableDest td;
    if (sc) {
      td.fixed_jobs = sc->base + rebop::kScratchJob;
      td.jobs_cap = rebop::kScratchJobBytes;
      td.fixed_items = sc->base + rebop::kScratchItems;
      td.items_cap = rebop::kScratchItemsBytes;
    } else {
      td.jobs = &g->djobs;
      td.items = &g->ditems;
    }
    int width = lds_start_width(kernel, run.mode, stats != nullptr,
                                run.x0_div ? nullptr : &g->x0);
    if (width == 2 && !compact_pays(std::vector<rebop_dev_job>{job})) width = 4;
    for (;;) {
      EventPair ev;
      if (stats && !sc) HIP_TRY(ev.create());
      HIP_TRY(rebop::launch_reset_counters(ctrs, stream));
      if (ev.e0) HIP_TRY(hipEventRecord(ev.e0, stream));
      rc = enqueue_lds(std::vector<rebop_dev_job>{job}, std::vector<double>{1.0}, td, ctrs,
                       nullptr, ctrs + REBOP_CTR_JOBQ, width, stats != nullptr, stream);
      if (rc) return rc;
      if (!stats) return REBOP_OK;
      if (ev.e1) HIP_TRY(hipEventRecord(ev.e1, stream));
      rc = finish_stats(stats, ctrs, ev, lds_kind(width), stream, run.draw_cap);
      if (rc == kRetryExact) return fail(REBOP_ERR_DEVICE, "internal error: the LDS kernel flagged f64");
      if (rc != kRetryWide) return rc;
      if (width == 8) return fail(REBOP_ERR_DEVICE, "internal error: the wide LDS kernel flagged int32");
      width *= 2;
    }
  }
  // lanes per trajectory: asked for, or (AUTO) when even 4 (2) lanes per
  // trajectory leave at most one wave per SIMD
  int group = (kernel & REBOP_KERNEL_GROUP4) ? 4 : (kernel & REBOP_KERNEL_GROUP2) ? 2 : 1;
  if (group > 1) {
    if (!spec)
      return fail(REBOP_ERR_UNSUPPORTED,
                  "lanes per trajectory > 1 needs the specialised kernel (REBOP_KERNEL_SPECIALIZED)");
    if (variant == REBOP_KERNEL_PERSISTENT || run.rates)
      return fail(REBOP_ERR_UNSUPPORTED,
                  "lanes per trajectory > 1 applies to plain runs without rate tables");
  } else if (variant == REBOP_KERNEL_AUTO && spec && !run.rates) {
    const uint64_t lanes = (uint64_t)g_dev.cus * 4u * 64u;
    if (run.n_traj * 4u <= lanes)
      group = 4;
    else if (run.n_traj * 2u <= lanes)
      group = 2;
  }
  const bool persistent = group == 1 && variant != REBOP_KERNEL_PLAIN;
  const int shape = spec ? -1 : pick_reg_shape(S, R);
  if (!spec && shape < 0)
    return fail(REBOP_ERR_UNSUPPORTED, "network with " + std::to_string(S) + " species and " +
                                           std::to_string(R) +
                                           " reactions exceeds the register kernels");
  run.persistent = persistent ? 1 : 0;
  run.counters = ctrs;
  run.queue = ctrs + REBOP_CTR_QUEUE;
  const size_t lds = spec ? 0 : sizeof(double) * (size_t)R * S;
  const uint64_t need_blocks = (run.n_traj * (uint64_t)group + threads - 1) / threads;
  uint64_t blocks = need_blocks;
  if (persistent) {
    int per_cu = 0;
    if (spec)
      HIP_TRY(rebop::jit_occupancy(g->jitk[run.draw_cap ? 1 : 0], threads,
                                   rebop::jit_dynamic_lds(run), &per_cu));
    else
      HIP_TRY(rebop::reg_occupancy(shape, threads, lds, &per_cu));
    if (per_cu < 1) per_cu = 1;
    uint64_t resident = (uint64_t)((double)per_cu * (double)g_dev.cus * occ);
    if (resident < (uint64_t)g_dev.cus) resident = (uint64_t)g_dev.cus;
    if (blocks > resident) blocks = resident;
  }
  if (blocks == 0) blocks = 1;
  if (blocks > 0x7fffffffull) return fail(REBOP_ERR_VALUE, "too many trajectories for one launch");
  const int kind = (persistent ? REBOP_KERNEL_PERSISTENT : REBOP_KERNEL_PLAIN) |
                   (spec ? REBOP_KERNEL_SPECIALIZED : 0) |
                   (group == 4 ? REBOP_KERNEL_GROUP4 : group == 2 ? REBOP_KERNEL_GROUP2 : 0);
  // tail hand-off (see rebop_dev_run): a wave spills at most spill_below
  // trajectories, once, so the arrays need (waves x spill_below) entries
  const uint32_t below = tail_lanes();
  const bool tail = persistent && spec && !run.rates && run.mode != REBOP_MODE_EVENTS &&
                    !(kernel & REBOP_KERNEL_WHOLE) && below > 0 && run.n_traj > 0 && !sc;
  const int tail_k = tail_group();
  uint64_t tail_blocks = 0;
  if (tail) {
    const uint64_t cap = blocks * (uint64_t)threads / 64u * below;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_draw = al(sizeof(uint64_t) * cap), o_mark = o_draw + al(sizeof(uint64_t) * cap);
    const size_t o_t = o_mark + al(sizeof(uint64_t) * cap), o_i = o_t + al(sizeof(double) * cap);
    const size_t o_x = o_i + al(sizeof(uint32_t) * cap);
    const size_t bytes = o_x + sizeof(double) * cap * (S ? S : 1);
    HIP_TRY(g->spill.ensure(bytes));
    char *b = (char *)g->spill.p;
    run.spill_below = below;
    run.spill_cap = cap;
    run.spill_j = (uint64_t *)b;
    run.spill_draw = (uint64_t *)(b + o_draw);
    run.spill_mark = (uint64_t *)(b + o_mark);
    run.spill_t = (double *)(b + o_t);
    run.spill_i = (uint32_t *)(b + o_i);
    run.spill_x = (double *)(b + o_x);
    tail_blocks = (cap * (uint64_t)tail_k + threads - 1) / threads;
  }

  EventPair ev;
  if (stats && !sc) HIP_TRY(ev.create());
  HIP_TRY(rebop::launch_reset_counters(run.counters, stream));
  // the sample times, read by the kernels at every sample
  const size_t n_grid = run.mode == REBOP_MODE_GRID ? (size_t)run.nb_steps + 1u : 1u;
  double *grid = nullptr;
  if (sc) {
    if (n_grid > rebop::kScratchMaxSamples)
      return fail(REBOP_ERR_VALUE, "nb_steps + 1 = " + std::to_string(n_grid) +
                                       " sample times exceed the scratch limit of " +
                                       std::to_string(rebop::kScratchMaxSamples));
    grid = (double *)(sc->base + rebop::kScratchGrid);
  } else {
    HIP_TRY(g->grid.ensure(sizeof(double) * n_grid));
    grid = (double *)g->grid.p;
  }
  run.grid = grid;
  HIP_TRY(rebop::launch_grid(grid, run, stream));
  if (ev.e0) HIP_TRY(hipEventRecord(ev.e0, stream));
  if (run.n_traj > 0) {
    if (spec)
      HIP_TRY(rebop::jit_launch(g->jitk[run.draw_cap ? 1 : 0], run, (int)blocks, threads, stream,
                                group));
    else
      HIP_TRY(rebop::launch_reg(shape, g->view, run, (int)blocks, threads, lds, stream));
  }
  if (tail) {
    // the handed-off trajectories, tail_k lanes each (entry e on group e)
    rebop_dev_run r2 = run;
    r2.persistent = 0;
    r2.spill_below = 0;
    r2.resume = 1;
    HIP_TRY(rebop::jit_launch(g->jitk[run.draw_cap ? 1 : 0], r2, (int)tail_blocks, threads, stream,
                              tail_k));
  }
  // the samples were written as f64 bits: convert them to int64 in place
  if (run.out && run.mode == REBOP_MODE_GRID && run.n_traj > 0)
    HIP_TRY(rebop::launch_raw_counts(run.out, (run.nb_steps + 1u) * (uint64_t)S * run.n_traj,
                                     g_dev.cus, stream));
  if (ev.e1) HIP_TRY(hipEventRecord(ev.e1, stream));

  if (stats) {
    rc = finish_stats(stats, run.counters, ev, kind, stream, run.draw_cap);
    if (rc != kRetryExact) return rc;
    // a count left |x| <= 2^52: the f64 counts of these kernels no longer
    // hold it exactly, the LDS kernel's int64 counts do (the rerun rewrites
    // every output).  Rate tables have no LDS kernel, and a resumable batch
    // has already advanced its state in place.
    if (run.rates || run.st_x)
      return fail(REBOP_ERR_RANGE, std::string("a count left the exactly simulated range |x| <= "
                                               "2^52 of ") +
                                       (run.rates ? "a run with rate tables" : "a resumable batch"));
    return launch_run(g, run_in, REBOP_KERNEL_LDS | REBOP_KERNEL_WIDE, stream, stats, 0, 1.0, sc);
  }
  return REBOP_OK;
}

int check_handle(const void *p, const char *what) {
  if (!p) return fail(REBOP_ERR_VALUE, std::string(what) + " is NULL");
  return REBOP_OK;
}

} // namespace

namespace rebop {

int report_error(int code, const char *msg) { return fail(code, msg); }

int run_device_scratch(const rebop_gillespie *gc, double tmax, uint64_t nb_steps,
                       uint64_t traj_begin, uint64_t n_traj, const int64_t *d_x0, int64_t *d_out,
                       int64_t *d_final, int kernel, void *stream, rebop_stats *stats,
                       const Scratch *sc) {
  if (int rc = check_handle(gc, "gillespie")) return rc;
  rebop_gillespie *g = const_cast<rebop_gillespie *>(gc);
  if (nb_steps == 0)
    return fail(REBOP_ERR_VALUE, "nb_steps must be > 0 for a grid run");
  // sample positions are 31-bit in the kernels
  if (nb_steps >= (1ull << 31)) return fail(REBOP_ERR_VALUE, "nb_steps must be below 2^31");
  if (stats) std::memset(stats, 0, sizeof *stats);
  int rc = ensure_device();
  if (rc) return rc;
  if (sc) {
    // nothing may be allocated: the network must be on this device already,
    // and the specialised kernel compiled where the launch would use it
    int cur = -1;
    HIP_TRY(hipGetDevice(&cur));
    if (g->dirty || g->dev_id != cur)
      return fail(REBOP_ERR_VALUE, "the network is not prepared on this device: call "
                                   "rebop_gillespie_prepare first");
    const int variant = kernel & 0xF;
    const bool wants_jit = specialised_only(g) || (kernel & REBOP_KERNEL_SPECIALIZED) ||
                           (variant == REBOP_KERNEL_AUTO && jit_eligible(g) && !jit_disabled());
    if (wants_jit && !jit_ready(g, effective_cap(g, tmax) != 0, cur))
      return fail(REBOP_ERR_VALUE, "the specialised kernel is not compiled on this device: call "
                                   "rebop_gillespie_prepare with this kernel kind (after setting "
                                   "any draw cap) first");
  } else {
    rc = upload_net(g);
    if (rc) return rc;
  }
  rebop_dev_run run{};
  run.seed = g->seed;
  run.traj_begin = traj_begin;
  run.n_traj = n_traj;
  run.tmax = tmax;
  run.nb_steps = nb_steps;
  run.mode = REBOP_MODE_GRID;
  run.x0 = d_x0 ? d_x0 : (const int64_t *)g->dx0.p;
  run.x0_div = d_x0 ? 1 : 0;
  run.out = d_out;
  run.final_x = d_final;
  return launch_run(g, run, kernel, (hipStream_t)stream, stats, 0, 1.0, sc);
}

} // namespace rebop

extern "C" {

const char *rebop_last_error(void) { return g_last_error.c_str(); }

int rebop_device_info(char *name, size_t name_len, int *cu_count) {
  int rc = ensure_device();
  if (rc) return rc;
  if (name && name_len) {
    std::snprintf(name, name_len, "%s", g_dev.name.c_str());
  }
  if (cu_count) *cu_count = g_dev.cus;
  return REBOP_OK;
}

int rebop_gillespie_new_with_seed(const int64_t *species, uint32_t nb_species, int sparse,
                                  uint64_t seed, rebop_gillespie **out) {
  if (!out) return fail(REBOP_ERR_VALUE, "out is NULL");
  *out = nullptr;
  if (nb_species && !species) return fail(REBOP_ERR_VALUE, "species is NULL");
  rebop_gillespie *g = new (std::nothrow) rebop_gillespie();
  if (!g) return fail(REBOP_ERR_ALLOC, "out of host memory");
  g->S = nb_species;
  g->sparse = sparse != 0;
  g->seed = seed;
  g->x0.assign(species, species + nb_species);
  *out = g;
  return REBOP_OK;
}

int rebop_gillespie_new(const int64_t *species, uint32_t nb_species, int sparse,
                        rebop_gillespie **out) {
  std::random_device rd;
  const uint64_t seed = ((uint64_t)rd() << 32) ^ (uint64_t)rd();
  return rebop_gillespie_new_with_seed(species, nb_species, sparse, seed, out);
}

int rebop_system_new(const int64_t *species, uint32_t nb_species, uint64_t seed,
                     rebop_gillespie **out) {
  const int rc = rebop_gillespie_new_with_seed(species, nb_species, 1, seed, out);
  if (rc == REBOP_OK) (*out)->system = true;
  return rc;
}

int rebop_system_add_reaction(rebop_gillespie *g, double rate, const uint32_t *reactants,
                              const uint32_t *reactant_counts, uint32_t n_reactants,
                              const uint32_t *products, const uint32_t *product_counts,
                              uint32_t n_products) {
  if (int rc = check_handle(g, "gillespie")) return rc;
  if (!g->system)
    return fail(REBOP_ERR_UNSUPPORTED, "rebop_system_add_reaction needs a rebop_system_new handle");
  if (n_reactants && (!reactants || !reactant_counts))
    return fail(REBOP_ERR_VALUE, "reactants is NULL");
  if (n_products && (!products || !product_counts))
    return fail(REBOP_ERR_VALUE, "products is NULL");
  if (std::isinf(rate)) return fail(REBOP_ERR_VALUE, "rate constant must be finite (or NaN)");
  std::vector<int64_t> d(g->S, 0);
  for (int side = 0; side < 2; ++side) {
    const uint32_t n = side ? n_products : n_reactants;
    const uint32_t *sp = side ? products : reactants;
    const uint32_t *m = side ? product_counts : reactant_counts;
    for (uint32_t i = 0; i < n; ++i) {
      if (sp[i] >= g->S)
        return fail(REBOP_ERR_INDEX, "species index " + std::to_string(sp[i]) +
                                         " out of range for " + std::to_string(g->S) + " species");
      if (m[i] > REBOP_SYSTEM_MAX_COUNT)
        return fail(REBOP_ERR_VALUE, "stoichiometric coefficient " + std::to_string(m[i]) +
                                         " is beyond " + std::to_string(REBOP_SYSTEM_MAX_COUNT));
      d[sp[i]] += side ? (int64_t)m[i] : -(int64_t)m[i];
    }
  }
  for (uint32_t s = 0; s < g->S; ++s)
    if (d[s] > REBOP_DIFF_LIMIT || d[s] < -REBOP_DIFF_LIMIT)
      return fail(REBOP_ERR_VALUE, "stoichiometric difference " + std::to_string(d[s]) +
                                       " is beyond |d| <= 2^20");
  Reaction r;
  r.k = rate;
  for (uint32_t i = 0; i < n_reactants; ++i) r.entries.emplace_back(reactants[i], reactant_counts[i]);
  r.diff = std::move(d);
  g->rx.push_back(std::move(r));
  g->dirty = true;
  mark_jit_dirty(g);
  return REBOP_OK;
}

void rebop_gillespie_free(rebop_gillespie *g) { delete g; }

int rebop_gillespie_set_draw_cap(rebop_gillespie *g, uint64_t max_draws) {
  if (int rc = check_handle(g, "gillespie")) return rc;
  g->draw_cap = max_draws;
  return REBOP_OK;
}

int rebop_gillespie_seed(rebop_gillespie *g, uint64_t seed) {
  if (int rc = check_handle(g, "gillespie")) return rc;
  g->seed = seed;
  return REBOP_OK;
}

uint32_t rebop_gillespie_nb_species(const rebop_gillespie *g) { return g ? g->S : 0; }
uint32_t rebop_gillespie_nb_reactions(const rebop_gillespie *g) {
  return g ? (uint32_t)g->rx.size() : 0;
}

static int add_reaction_common(rebop_gillespie *g, double rate, bool rate_sparse,
                               const uint32_t *idx, const uint32_t *orders, uint32_t n,
                               const int64_t *diffs, uint32_t nd) {
  if (int rc = check_handle(g, "gillespie")) return rc;
  if (g->system)
    return fail(REBOP_ERR_UNSUPPORTED, "a define_system network takes rebop_system_add_reaction");
  if (n && !orders) return fail(REBOP_ERR_VALUE, "orders is NULL");
  if (rate_sparse && n && !idx) return fail(REBOP_ERR_VALUE, "species_idx is NULL");
  if (nd && !diffs) return fail(REBOP_ERR_VALUE, "differences is NULL");
  // src/gillespie.rs:227-235: the rate must operate on existing species
  if (!rate_sparse) {
    if (n != g->S)
      return fail(REBOP_ERR_SHAPE, "assertion `left == right` failed: rate has " +
                                       std::to_string(n) + " orders for " +
                                       std::to_string(g->S) + " species");
  } else {
    for (uint32_t i = 0; i < n; ++i)
      if (idx[i] >= g->S)
        return fail(REBOP_ERR_INDEX, "assertion failed: species index " + std::to_string(idx[i]) +
                                         " out of range for " + std::to_string(g->S) +
                                         " species");
  }
  // src/gillespie.rs:237: one difference per species
  if (nd != g->S)
    return fail(REBOP_ERR_SHAPE, "assertion `left == right` failed: " + std::to_string(nd) +
                                     " differences for " + std::to_string(g->S) + " species");
  // an infinite constant makes E/total == 0 forever (time never advances);
  // NaN is accepted: it makes the state absorbing, as in the reference
  if (std::isinf(rate))
    return fail(REBOP_ERR_VALUE, "rate constant must be finite (or NaN)");
  for (uint32_t s = 0; s < nd; ++s)
    if (diffs[s] > REBOP_DIFF_LIMIT || diffs[s] < -REBOP_DIFF_LIMIT)
      return fail(REBOP_ERR_VALUE, "stoichiometric difference " + std::to_string(diffs[s]) +
                                       " is beyond |d| <= 2^20");
  Reaction r;
  r.k = rate;
  if (!rate_sparse) {
    // dense LMA: ascending species; order 0 contributes no factor
    for (uint32_t s = 0; s < n; ++s)
      if (orders[s] > 0) r.entries.emplace_back(s, orders[s]);
  } else {
    // LMASparse: the given pair order, duplicates kept (Rate::rate :79-86);
    // a sparse net converts dense rates to this same ascending list (Rate::sparse)
    for (uint32_t i = 0; i < n; ++i)
      if (orders[i] > 0) r.entries.emplace_back(idx[i], orders[i]);
  }
  r.diff.assign(diffs, diffs + nd);
  g->rx.push_back(std::move(r));
  g->dirty = true;
  mark_jit_dirty(g);
  return REBOP_OK;
}

int rebop_gillespie_add_reaction_lma(rebop_gillespie *g, double rate, const uint32_t *orders,
                                     uint32_t n_orders, const int64_t *differences,
                                     uint32_t n_differences) {
  return add_reaction_common(g, rate, false, nullptr, orders, n_orders, differences,
                             n_differences);
}

int rebop_gillespie_add_reaction_expr(rebop_gillespie *g, const uint32_t *code, uint32_t n_code,
                                      const double *consts, uint32_t n_consts,
                                      const int64_t *differences, uint32_t n_differences) {
  if (int rc = check_handle(g, "gillespie")) return rc;
  if (g->system)
    return fail(REBOP_ERR_UNSUPPORTED, "a define_system network takes rebop_system_add_reaction");
  if (n_code && !code) return fail(REBOP_ERR_VALUE, "code is NULL");
  if (n_consts && !consts) return fail(REBOP_ERR_VALUE, "consts is NULL");
  if (n_differences && !differences) return fail(REBOP_ERR_VALUE, "differences is NULL");
  if (n_code == 0) return fail(REBOP_ERR_VALUE, "empty rate expression");
  if (n_code > REBOP_EXPR_MAX_CODE)
    return fail(REBOP_ERR_VALUE, "rate expression longer than 4096 tokens");
  // the postfix program must leave exactly one value, never underflow the
  // stack nor exceed its depth, and refer to existing constants and species
  uint32_t depth = 0;
  for (uint32_t i = 0; i < n_code; ++i) {
    const uint32_t op = REBOP_EXPR_OPCODE(code[i]), arg = REBOP_EXPR_ARG(code[i]);
    const std::string at = " at token " + std::to_string(i);
    switch (op) {
    case REBOP_OP_CONST:
      if (arg >= n_consts) return fail(REBOP_ERR_INDEX, "constant index out of range" + at);
      depth += 1;
      break;
    case REBOP_OP_SPECIES:
      if (arg >= g->S)
        return fail(REBOP_ERR_INDEX, "assertion failed: species index " + std::to_string(arg) +
                                         " out of range for " + std::to_string(g->S) +
                                         " species" + at);
      depth += 1;
      break;
    case REBOP_OP_NEG:
    case REBOP_OP_EXP:
      if (depth < 1) return fail(REBOP_ERR_VALUE, "rate expression stack underflow" + at);
      break;
    case REBOP_OP_ADD:
    case REBOP_OP_SUB:
    case REBOP_OP_MUL:
    case REBOP_OP_DIV:
    case REBOP_OP_POW:
    case REBOP_OP_MAX:
    case REBOP_OP_MIN:
      if (depth < 2) return fail(REBOP_ERR_VALUE, "rate expression stack underflow" + at);
      depth -= 1;
      break;
    default:
      return fail(REBOP_ERR_VALUE, "unknown rate expression operation " + std::to_string(op) + at);
    }
    if (depth > REBOP_EXPR_MAX_DEPTH)
      return fail(REBOP_ERR_VALUE, "rate expression deeper than 32 operands" + at);
  }
  if (depth != 1)
    return fail(REBOP_ERR_VALUE, "rate expression leaves " + std::to_string(depth) + " values");
  if (n_differences != g->S)
    return fail(REBOP_ERR_SHAPE, "assertion `left == right` failed: " +
                                     std::to_string(n_differences) + " differences for " +
                                     std::to_string(g->S) + " species");
  for (uint32_t s = 0; s < n_differences; ++s)
    if (differences[s] > REBOP_DIFF_LIMIT || differences[s] < -REBOP_DIFF_LIMIT)
      return fail(REBOP_ERR_VALUE, "stoichiometric difference " + std::to_string(differences[s]) +
                                       " is beyond |d| <= 2^20");
  Reaction r;
  r.k = 0.0;
  r.expr.assign(code, code + n_code);
  r.consts.assign(consts, consts + n_consts);
  r.diff.assign(differences, differences + n_differences);
  g->rx.push_back(std::move(r));
  g->n_expr += 1;
  g->dirty = true;
  mark_jit_dirty(g);
  return REBOP_OK;
}

int rebop_gillespie_add_reaction_lma_sparse(rebop_gillespie *g, double rate,
                                            const uint32_t *species_idx, const uint32_t *orders,
                                            uint32_t n_pairs, const int64_t *differences,
                                            uint32_t n_differences) {
  return add_reaction_common(g, rate, true, species_idx, orders, n_pairs, differences,
                             n_differences);
}

int rebop_gillespie_set_species(rebop_gillespie *g, const int64_t *species, uint32_t n) {
  if (int rc = check_handle(g, "gillespie")) return rc;
  if (n != g->S)
    return fail(REBOP_ERR_SHAPE, "assertion `left == right` failed: " + std::to_string(n) +
                                     " counts for " + std::to_string(g->S) + " species");
  if (n && !species) return fail(REBOP_ERR_VALUE, "species is NULL");
  g->x0.assign(species, species + n);
  g->dirty = true;
  return REBOP_OK;
}

int rebop_gillespie_get_species(const rebop_gillespie *g, int64_t *species, uint32_t n) {
  if (int rc = check_handle(g, "gillespie")) return rc;
  if (n != g->S) return fail(REBOP_ERR_SHAPE, "length mismatch");
  if (n) std::memcpy(species, g->x0.data(), sizeof(int64_t) * n);
  return REBOP_OK;
}

int rebop_gillespie_prepare(const rebop_gillespie *gc, int kernel) {
  if (int rc = check_handle(gc, "gillespie")) return rc;
  rebop_gillespie *g = const_cast<rebop_gillespie *>(gc);
  int cur = -1;
  const bool has_dev = hipGetDevice(&cur) == hipSuccess;
  // AUTO compiles the specialised kernel where an AUTO run would use it
  const bool auto_jit = (kernel & 0xF) == REBOP_KERNEL_AUTO && !(kernel & REBOP_KERNEL_SPECIALIZED) &&
                        jit_eligible(g) && !jit_disabled();
  if (!(kernel & REBOP_KERNEL_SPECIALIZED) && !auto_jit && !specialised_only(g)) {
    if (!has_dev) return REBOP_OK;
    if (int rc = ensure_device()) return rc;
    return upload_net(g); // the network on the device (the data-driven and LDS kernels read it)
  }
  if (!jit_eligible(g))
    return fail(REBOP_ERR_UNSUPPORTED, "network too large for a specialised kernel");
  if (!has_dev) {
    // no device: compile only, to check the generated kernel
    std::string err;
    size_t cs = 0;
    if (rebop::jit_compile_only(jit_source_of(g, g->draw_cap != 0), &cs, &err) != 0)
      return fail(REBOP_ERR_DEVICE, "specialised kernel: " + err);
    return REBOP_OK;
  }
  if (int rc = ensure_device()) return rc;
  if (int rc = upload_net(g)) return rc;
  // with a draw cap set, the variant that checks it too
  if (g->draw_cap)
    if (int rc = prepare_jit(g, true)) return rc;
  return prepare_jit(g);
}

int rebop_run_device(const rebop_gillespie *gc, double tmax, uint64_t nb_steps,
                     uint64_t traj_begin, uint64_t n_traj, const int64_t *d_x0, int64_t *d_out,
                     int64_t *d_final, int kernel, void *stream, rebop_stats *stats) {
  return rebop::run_device_scratch(gc, tmax, nb_steps, traj_begin, n_traj, d_x0, d_out, d_final,
                                   kernel, stream, stats, nullptr);
}

int rebop_autotune(rebop_gillespie *g, double tmax, uint64_t nb_steps, uint64_t n_pilot,
                   rebop_tune_result *out) {
  if (int rc = check_handle(g, "gillespie")) return rc;
  if (nb_steps == 0) return fail(REBOP_ERR_VALUE, "nb_steps must be > 0");
  if (nb_steps >= (1ull << 31)) return fail(REBOP_ERR_VALUE, "nb_steps must be below 2^31");
  if (n_pilot == 0) return fail(REBOP_ERR_VALUE, "n_pilot must be > 0");
  if (int rc = check_x0(g->x0.data(), g->x0.size())) return rc;
  int rc = ensure_device();
  if (rc) return rc;
  rc = upload_net(g);
  if (rc) return rc;
  const uint32_t S = g->S, R = (uint32_t)g->rx.size();
  const std::string cache = tune_cache_path(), key = tune_key(g, nb_steps, n_pilot, g_dev.name);
  const char *force = std::getenv("REBOP_TUNE_FORCE");
  {
    int k = 0, th = 0;
    double oc = 0, ms = 0;
    if (!(force && *force && std::strcmp(force, "0") != 0) &&
        tune_cache_get(cache, key, &k, &th, &oc, &ms)) {
      g->tune_kernel = k;
      g->tune_threads = th;
      g->tune_occ = oc;
      if (out) {
        out->kernel = k;
        out->threads = th;
        out->occupancy = oc;
        out->ms = ms;
        out->n_tried = 0; // from the cache
      }
      return REBOP_OK;
    }
  }
  struct Cand {
    int kernel, threads;
    double occ;
  };
  std::vector<Cand> cands;
  const bool jit_ok = jit_eligible(g) && !jit_disabled() && prepare_jit(g) == REBOP_OK;
  const bool reg_ok = pick_reg_shape(S, R) >= 0;
  for (int threads : {64, 128, 256}) {
    for (double occ : {1.0, 0.6, 0.4, 0.25}) {
      if (jit_ok)
        cands.push_back({REBOP_KERNEL_PERSISTENT | REBOP_KERNEL_SPECIALIZED, threads, occ});
      if (reg_ok) cands.push_back({REBOP_KERNEL_PERSISTENT, threads, occ});
    }
    if (jit_ok) {
      cands.push_back({REBOP_KERNEL_PLAIN | REBOP_KERNEL_SPECIALIZED, threads, 1.0});
      cands.push_back({REBOP_KERNEL_PLAIN | REBOP_KERNEL_SPECIALIZED | REBOP_KERNEL_GROUP2, threads, 1.0});
      cands.push_back({REBOP_KERNEL_PLAIN | REBOP_KERNEL_SPECIALIZED | REBOP_KERNEL_GROUP4, threads, 1.0});
    }
    if (reg_ok) cands.push_back({REBOP_KERNEL_PLAIN, threads, 1.0});
  }
  cands.push_back({REBOP_KERNEL_LDS, 64, 1.0});
  rebop_dev_run run{};
  run.seed = g->seed;
  run.traj_begin = 0;
  run.n_traj = n_pilot;
  run.tmax = tmax;
  run.nb_steps = nb_steps;
  run.mode = REBOP_MODE_GRID;
  run.x0 = (const int64_t *)g->dx0.p;
  const int saved = g->tune_kernel;
  g->tune_kernel = -1;
  double best = 1e300;
  Cand pick = cands.back();
  uint32_t tried = 0;
  for (const Cand &c : cands) {
    rebop_stats st{};
    // one warm launch, then the timed one
    rc = launch_run(g, run, c.kernel, nullptr, &st, c.threads, c.occ);
    if (rc == REBOP_OK) rc = launch_run(g, run, c.kernel, nullptr, &st, c.threads, c.occ);
    if (rc != REBOP_OK) continue; // a variant that does not apply
    tried++;
    if (st.kernel_ms < best) {
      best = st.kernel_ms;
      pick = c;
    }
  }
  if (tried == 0) {
    g->tune_kernel = saved;
    return fail(REBOP_ERR_UNSUPPORTED, "no kernel variant ran for this network");
  }
  g->tune_kernel = pick.kernel;
  g->tune_threads = pick.threads;
  g->tune_occ = pick.occ;
  tune_cache_put(cache, key, pick.kernel, pick.threads, pick.occ, best);
  if (out) {
    out->kernel = pick.kernel;
    out->threads = pick.threads;
    out->occupancy = pick.occ;
    out->ms = best;
    out->n_tried = tried;
  }
  return REBOP_OK;
}

int rebop_run_sweep_device(const rebop_gillespie *gc, uint64_t n_points, uint64_t traj_per_point,
                           const double *d_rates, const int64_t *d_x0, double tmax,
                           uint64_t nb_steps, uint64_t traj_begin, int64_t *d_out,
                           int64_t *d_final, int kernel, void *stream, rebop_stats *stats) {
  if (int rc = check_handle(gc, "gillespie")) return rc;
  rebop_gillespie *g = const_cast<rebop_gillespie *>(gc);
  if (nb_steps == 0) return fail(REBOP_ERR_VALUE, "nb_steps must be > 0 for a grid run");
  if (nb_steps >= (1ull << 31)) return fail(REBOP_ERR_VALUE, "nb_steps must be below 2^31");
  if (traj_per_point == 0) return fail(REBOP_ERR_VALUE, "traj_per_point must be > 0");
  if (d_rates && g->n_expr)
    return fail(REBOP_ERR_UNSUPPORTED,
                "rate tables apply to mass-action rates; this network has rate expressions");
  if (n_points && traj_per_point > UINT64_MAX / n_points)
    return fail(REBOP_ERR_VALUE, "n_points * traj_per_point overflows");
  if (!d_x0) {
    if (int rc = check_x0(g->x0.data(), g->x0.size())) return rc;
  }
  if (stats) std::memset(stats, 0, sizeof *stats);
  int rc = ensure_device();
  if (rc) return rc;
  rc = upload_net(g);
  if (rc) return rc;
  rebop_dev_run run{};
  run.seed = g->seed;
  run.traj_begin = traj_begin;
  run.n_traj = n_points * traj_per_point;
  run.tmax = tmax;
  run.nb_steps = nb_steps;
  run.mode = REBOP_MODE_GRID;
  run.x0 = d_x0 ? d_x0 : (const int64_t *)g->dx0.p;
  run.x0_div = d_x0 ? traj_per_point : 0;
  run.rates = d_rates;
  run.rates_div = traj_per_point;
  run.out = d_out;
  run.final_x = d_final;
  return launch_run(g, run, kernel, (hipStream_t)stream, stats);
}

int rebop_run(const rebop_gillespie *gc, double tmax, uint64_t nb_steps, uint64_t traj_begin,
              uint64_t n_traj, const int64_t *x0, int64_t *out, double *times,
              rebop_stats *stats) {
  if (int rc = check_handle(gc, "gillespie")) return rc;
  if (nb_steps == 0) return fail(REBOP_ERR_VALUE, "nb_steps must be > 0 for a grid run");
  if (nb_steps >= (1ull << 31)) return fail(REBOP_ERR_VALUE, "nb_steps must be below 2^31");
  const uint32_t S = gc->S;
  if (times)
    for (uint64_t i = 0; i <= nb_steps; ++i) times[i] = tmax * (double)i / (double)nb_steps;
  int rc = ensure_device();
  if (rc) return rc;
  DevBuf dx0, dout;
  const size_t out_bytes = sizeof(int64_t) * (size_t)(nb_steps + 1) * S * n_traj;
  if (x0 && n_traj && S) {
    HIP_TRY(dx0.ensure(sizeof(int64_t) * (size_t)n_traj * S));
    HIP_TRY(hipMemcpy(dx0.p, x0, sizeof(int64_t) * (size_t)n_traj * S, hipMemcpyHostToDevice));
  }
  if (out && out_bytes) HIP_TRY(dout.ensure(out_bytes));
  rebop_stats local{};
  rc = rebop_run_device(gc, tmax, nb_steps, traj_begin, n_traj,
                        (x0 && n_traj && S) ? (const int64_t *)dx0.p : nullptr,
                        (out && out_bytes) ? (int64_t *)dout.p : nullptr, nullptr,
                        REBOP_KERNEL_AUTO, nullptr, &local);
  if (rc) return rc;
  if (out && out_bytes) HIP_TRY(hipMemcpy(out, dout.p, out_bytes, hipMemcpyDeviceToHost));
  if (stats) *stats = local;
  return REBOP_OK;
}

int rebop_run_jobs_device(const rebop_job *jobs, uint32_t n_jobs, void *stream_,
                          rebop_stats *stats, uint64_t *events_per_job) {
  if (stats) std::memset(stats, 0, sizeof *stats);
  if (n_jobs && !jobs) return fail(REBOP_ERR_VALUE, "jobs is NULL");
  int rc = ensure_device();
  if (rc) return rc;
  hipStream_t stream = (hipStream_t)stream_;
  std::vector<rebop_dev_job> dj;
  std::vector<double> cost;
  dj.reserve(n_jobs);
  uint64_t cap_max = 0;
  for (uint32_t j = 0; j < n_jobs; ++j) {
    const rebop_job &J = jobs[j];
    if (int e = check_handle(J.net, "job net")) return e;
    if (J.nb_steps == 0 || J.nb_steps >= (1ull << 31))
      return fail(REBOP_ERR_VALUE, "job " + std::to_string(j) + ": nb_steps must be in 1 .. 2^31 - 1");
    rebop_gillespie *g = const_cast<rebop_gillespie *>(J.net);
    rc = upload_net(g);
    if (rc) return rc;
    rebop_dev_job d = g->job;
    d.seed = g->seed;
    d.traj_begin = J.traj_begin;
    d.n_traj = J.n_traj;
    d.tmax = J.tmax;
    d.nb_steps = J.nb_steps;
    d.mode = REBOP_MODE_GRID;
    d.draw_cap = effective_cap(g, J.tmax);
    cap_max = std::max(cap_max, d.draw_cap);
    d.x0_div = J.d_x0 ? 1 : 0;
    d.x0 = J.d_x0 ? J.d_x0 : (const int64_t *)g->dx0.p;
    d.out = J.d_out;
    d.final_x = J.d_final;
    dj.push_back(d);
    cost.push_back(J.cost_hint);
  }
  // device tables and counters, kept per (device, stream) and reused: calls
  // on one stream are ordered, so the next call rewrites them only after the
  // previous launch has read them
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  JobBuffers *jb = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_job_mu);
    jb = &(*g_job_bufs)[std::make_pair(dev, stream)];
  }
  HIP_TRY(jb->counters.ensure(sizeof(unsigned long long) * REBOP_CTR_COUNT));
  unsigned long long *ctrs = (unsigned long long *)jb->counters.p;
  if (events_per_job && n_jobs) HIP_TRY(jb->dev_ev.ensure(sizeof(unsigned long long) * n_jobs));
  unsigned long long *jev = events_per_job && n_jobs ? (unsigned long long *)jb->dev_ev.p : nullptr;
  // one trajectory queue per job
  HIP_TRY(jb->dev_q.ensure(sizeof(unsigned long long) * (n_jobs ? n_jobs : 1)));
  unsigned long long *jq = (unsigned long long *)jb->dev_q.p;
  TableDest td;
  td.jobs = &jb->djobs;
  td.items = &jb->ditems;
  // the counts' width: the narrowest that every job's shared initial counts
  // fit (int16 only when the counters are read, so that a launch that meets
  // a count beyond it can be rerun; device-side initial counts start at
  // int32); a launch that meets a count beyond its width is rerun with the
  // next wider counts when its counters are read (stats or events_per_job)
  const bool read = stats || jev;
  int width = 2;
  for (uint32_t j = 0; j < n_jobs; ++j)
    width = std::max(width, lds_start_width(REBOP_KERNEL_AUTO, REBOP_MODE_GRID, read,
                                            jobs[j].d_x0 ? nullptr : &jobs[j].net->x0));
  if (width == 2 && !compact_pays(dj)) width = 4;
  rebop_stats local{};
  for (;;) {
    if (jev) HIP_TRY(rebop::launch_zero_u64(jev, n_jobs, stream));
    HIP_TRY(rebop::launch_zero_u64(jq, n_jobs, stream));
    EventPair ev;
    if (stats) HIP_TRY(ev.create());
    HIP_TRY(rebop::launch_reset_counters(ctrs, stream));
    if (ev.e0) HIP_TRY(hipEventRecord(ev.e0, stream));
    rc = enqueue_lds(dj, cost, td, ctrs, jev, jq, width, read, stream);
    if (rc) return rc;
    if (ev.e1) HIP_TRY(hipEventRecord(ev.e1, stream));
    if (!read) return REBOP_OK;
    if (jev)
      HIP_TRY(hipMemcpyAsync(events_per_job, jev, sizeof(unsigned long long) * n_jobs,
                             hipMemcpyDeviceToHost, stream));
    rc = finish_stats(&local, ctrs, ev, lds_kind(width), stream, cap_max);
    if (rc == kRetryWide && width < 8) {
      width *= 2;
      continue;
    }
    if (rc == kRetryWide) return fail(REBOP_ERR_DEVICE, "internal error: the wide LDS kernel flagged int32");
    if (stats) *stats = local;
    return rc;
  }
}

int rebop_run_all_events(const rebop_gillespie *gc, double tmax, uint64_t traj_begin,
                         uint64_t n_traj, uint64_t *n_records, double *times, int64_t *states,
                         rebop_stats *stats) {
  if (int rc = check_handle(gc, "gillespie")) return rc;
  if (!n_records && n_traj) return fail(REBOP_ERR_VALUE, "n_records is NULL");
  rebop_gillespie *g = const_cast<rebop_gillespie *>(gc);
  if (stats) std::memset(stats, 0, sizeof *stats);
  int rc = ensure_device();
  if (rc) return rc;
  rc = upload_net(g);
  if (rc) return rc;
  const uint32_t S = g->S;
  DevBuf dcount, doff, dt, ds;
  HIP_TRY(dcount.ensure(sizeof(uint64_t) * (n_traj ? n_traj : 1)));
  rebop_dev_run run{};
  run.seed = g->seed;
  run.traj_begin = traj_begin;
  run.n_traj = n_traj;
  run.tmax = tmax;
  run.nb_steps = 0;
  run.mode = REBOP_MODE_EVENTS;
  run.x0 = (const int64_t *)g->dx0.p;
  run.x0_div = 0;
  run.n_records = (uint64_t *)dcount.p;
  const bool want = times != nullptr || states != nullptr;
  rebop_stats local{};
  if (!want) {
    // pass 1: count the records of every trajectory (also when the draw cap
    // stopped some: their counts end at the cap)
    rc = launch_run(g, run, REBOP_KERNEL_AUTO, nullptr, &local);
    if (rc && rc != REBOP_ERR_CAP) return rc;
    if (n_traj)
      HIP_TRY(hipMemcpy(n_records, dcount.p, sizeof(uint64_t) * n_traj, hipMemcpyDeviceToHost));
    if (stats) *stats = local;
    return rc;
  }
  // pass 2: the caller's counts give the offsets; the kernel checks them
  std::vector<uint64_t> off(n_traj ? n_traj : 1, 0);
  uint64_t total = 0;
  for (uint64_t j = 0; j < n_traj; ++j) {
    off[j] = total;
    total += n_records[j];
  }
  if (n_traj) {
    HIP_TRY(doff.ensure(sizeof(uint64_t) * n_traj));
    HIP_TRY(hipMemcpy(doff.p, off.data(), sizeof(uint64_t) * n_traj, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dcount.p, n_records, sizeof(uint64_t) * n_traj, hipMemcpyHostToDevice));
  }
  HIP_TRY(dt.ensure(sizeof(double) * (total ? total : 1)));
  HIP_TRY(ds.ensure(sizeof(int64_t) * (total * S ? total * S : 1)));
  run.rec_offsets = (const uint64_t *)doff.p;
  run.rec_times = (double *)dt.p;
  run.rec_states = (int64_t *)ds.p;
  rc = launch_run(g, run, REBOP_KERNEL_AUTO, nullptr, &local);
  if (rc == REBOP_ERR_RANGE)
    return fail(REBOP_ERR_VALUE, "n_records does not match this network and seed (or a count "
                                 "left the exactly simulated range)");
  if (rc && rc != REBOP_ERR_CAP) return rc;
  if (times && total) HIP_TRY(hipMemcpy(times, dt.p, sizeof(double) * total, hipMemcpyDeviceToHost));
  if (states && total * S)
    HIP_TRY(hipMemcpy(states, ds.p, sizeof(int64_t) * total * S, hipMemcpyDeviceToHost));
  if (stats) *stats = local;
  return rc;
}

int rebop_run_all_events_device(const rebop_gillespie *gc, double tmax, uint64_t traj_begin,
                                uint64_t n_traj, uint64_t *d_n_records, const uint64_t *d_offsets,
                                double *d_times, int64_t *d_states, void *stream,
                                rebop_stats *stats) {
  if (int rc = check_handle(gc, "gillespie")) return rc;
  if (!d_n_records && n_traj) return fail(REBOP_ERR_VALUE, "d_n_records is NULL");
  const bool want = d_offsets != nullptr || d_times != nullptr || d_states != nullptr;
  if (want && n_traj && (!d_offsets || !d_times || !d_states))
    return fail(REBOP_ERR_VALUE, "pass 2 needs d_offsets, d_times and d_states");
  rebop_gillespie *g = const_cast<rebop_gillespie *>(gc);
  if (stats) std::memset(stats, 0, sizeof *stats);
  int rc = ensure_device();
  if (rc) return rc;
  rc = upload_net(g);
  if (rc) return rc;
  rebop_dev_run run{};
  run.seed = g->seed;
  run.traj_begin = traj_begin;
  run.n_traj = n_traj;
  run.tmax = tmax;
  run.nb_steps = 0;
  run.mode = REBOP_MODE_EVENTS;
  run.x0 = (const int64_t *)g->dx0.p;
  run.x0_div = 0;
  run.n_records = d_n_records;
  if (want) {
    run.rec_offsets = d_offsets;
    run.rec_times = d_times;
    run.rec_states = d_states;
  }
  rebop_stats local{};
  rc = launch_run(g, run, REBOP_KERNEL_AUTO, (hipStream_t)stream, &local);
  if (want && rc == REBOP_ERR_RANGE)
    return fail(REBOP_ERR_VALUE, "d_n_records does not match this network and seed (or a count "
                                 "left the exactly simulated range)");
  if (rc) return rc;
  if (stats) *stats = local;
  return REBOP_OK;
}

int rebop_batch_new(const rebop_gillespie *g, uint64_t traj_begin, uint64_t n_traj,
                    const int64_t *x0, rebop_batch **out) {
  if (!out) return fail(REBOP_ERR_VALUE, "out is NULL");
  *out = nullptr;
  if (int rc = check_handle(g, "gillespie")) return rc;
  if (int rc = ensure_device()) return rc;
  const uint32_t S = g->S;
  const int64_t *src = x0 ? x0 : g->x0.data();
  if (int rc = check_x0(src, x0 ? (size_t)n_traj * S : S)) return rc;
  rebop_batch *b = new (std::nothrow) rebop_batch();
  if (!b) return fail(REBOP_ERR_ALLOC, "out of host memory");
  b->g = g;
  b->traj_begin = traj_begin;
  b->n = n_traj;
  std::vector<double> hx((size_t)S * n_traj);
  for (uint64_t j = 0; j < n_traj; ++j)
    for (uint32_t s = 0; s < S; ++s)
      hx[(size_t)s * n_traj + j] = (double)(x0 ? x0[j * S + s] : g->x0[s]);
  std::vector<double> ht(n_traj, 0.0);
  std::vector<uint64_t> hd(n_traj, 0);
  hipError_t e = b->x.ensure(sizeof(double) * hx.size());
  if (e == hipSuccess) e = b->t.ensure(sizeof(double) * n_traj);
  if (e == hipSuccess) e = b->draw.ensure(sizeof(uint64_t) * n_traj);
  if (e == hipSuccess && !hx.empty())
    e = hipMemcpy(b->x.p, hx.data(), sizeof(double) * hx.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && n_traj)
    e = hipMemcpy(b->t.p, ht.data(), sizeof(double) * n_traj, hipMemcpyHostToDevice);
  if (e == hipSuccess && n_traj)
    e = hipMemcpy(b->draw.p, hd.data(), sizeof(uint64_t) * n_traj, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    delete b;
    return fail(REBOP_ERR_DEVICE, std::string("batch allocation: ") + hipGetErrorString(e));
  }
  *out = b;
  return REBOP_OK;
}

void rebop_batch_free(rebop_batch *b) { delete b; }

int rebop_batch_advance_until(rebop_batch *b, double tmax, rebop_stats *stats) {
  if (int rc = check_handle(b, "batch")) return rc;
  rebop_gillespie *g = const_cast<rebop_gillespie *>(b->g);
  rebop_stats local{};
  rebop_dev_run run{};
  run.seed = g->seed;
  run.traj_begin = b->traj_begin;
  run.n_traj = b->n;
  run.tmax = tmax;
  run.nb_steps = 0;
  run.mode = REBOP_MODE_TARGET;
  run.st_x = (double *)b->x.p;
  run.st_t = (double *)b->t.p;
  run.st_draw = (uint64_t *)b->draw.p;
  int rc = launch_run(g, run, REBOP_KERNEL_AUTO, nullptr, &local);
  if (stats) *stats = local;
  return rc;
}

int rebop_batch_get_species(const rebop_batch *b, int64_t *species) {
  if (int rc = check_handle(b, "batch")) return rc;
  const uint32_t S = b->g->S;
  std::vector<double> hx((size_t)S * b->n);
  if (!hx.empty())
    HIP_TRY(hipMemcpy(hx.data(), b->x.p, sizeof(double) * hx.size(), hipMemcpyDeviceToHost));
  for (uint64_t j = 0; j < b->n; ++j)
    for (uint32_t s = 0; s < S; ++s) species[j * S + s] = (int64_t)hx[(size_t)s * b->n + j];
  return REBOP_OK;
}

int rebop_batch_set_species(rebop_batch *b, const int64_t *species) {
  if (int rc = check_handle(b, "batch")) return rc;
  const uint32_t S = b->g->S;
  if (int rc = check_x0(species, (size_t)S * b->n)) return rc;
  std::vector<double> hx((size_t)S * b->n);
  for (uint64_t j = 0; j < b->n; ++j)
    for (uint32_t s = 0; s < S; ++s) hx[(size_t)s * b->n + j] = (double)species[j * S + s];
  if (!hx.empty())
    HIP_TRY(hipMemcpy(b->x.p, hx.data(), sizeof(double) * hx.size(), hipMemcpyHostToDevice));
  return REBOP_OK;
}

int rebop_batch_get_time(const rebop_batch *b, double *t) {
  if (int rc = check_handle(b, "batch")) return rc;
  if (b->n) HIP_TRY(hipMemcpy(t, b->t.p, sizeof(double) * b->n, hipMemcpyDeviceToHost));
  return REBOP_OK;
}

int rebop_batch_set_time(rebop_batch *b, const double *t) {
  if (int rc = check_handle(b, "batch")) return rc;
  if (b->n) HIP_TRY(hipMemcpy(b->t.p, t, sizeof(double) * b->n, hipMemcpyHostToDevice));
  return REBOP_OK;
}

int rebop_batch_get_draws(const rebop_batch *b, uint64_t *draws) {
  if (int rc = check_handle(b, "batch")) return rc;
  if (b->n) HIP_TRY(hipMemcpy(draws, b->draw.p, sizeof(uint64_t) * b->n, hipMemcpyDeviceToHost));
  return REBOP_OK;
}

int rebop_batch_set_draws(rebop_batch *b, const uint64_t *draws) {
  if (int rc = check_handle(b, "batch")) return rc;
  if (b->n) HIP_TRY(hipMemcpy(b->draw.p, draws, sizeof(uint64_t) * b->n, hipMemcpyHostToDevice));
  return REBOP_OK;
}

} // extern "C"
