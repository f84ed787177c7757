// engine.hip -- host side of the MI355X macrospin engine: kernel registry,
// plan building (grouping, cost ordering, driver records), the drop-in
// single-junction run and the C ABI of include/cmtj_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "kernels.h"

using namespace cmtj;

// ---------------------------------------------------------------------------
// errors

static thread_local std::string g_err;
static thread_local std::string g_last_kernels;  // kernels of the calling thread's last launch

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(CMTJ_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------------------
// constants store (reference core/constants.hpp:39-53)

static const cmtj_constants kDefaults = {12.566e-7,           220880.0,      795774.715459,
                                         6.62607015e-34 / (2. * M_PI), 1.60217662e-19,
                                         1.380649e-23};
static cmtj_constants g_consts = kDefaults;
static std::mutex g_consts_mu;

// ---------------------------------------------------------------------------
// kernel registry

#include "registry.h"


// kinds: bit 0 = a record of another kind than step/sine/trapezoid, bit 1 =
// a step record, bit 2 = a sine record, bit 3 = a trapezoid record.
constexpr uint32_t KIND_OTHER = 1, KIND_STEP = 2, KIND_SINE = 4, KIND_TRAP = 8;

// Log layout the LOGF = 1 kernels compile in: log_mask | mr_mode << 8 | f32 << 16.
constexpr uint32_t kLogPimm = (CMTJ_LOG_MX | CMTJ_LOG_MR) | (CMTJ_MR_CLASSIC << 8) | (1u << 16);

static bool covers(const KEntry& k, uint32_t flags, uint32_t kinds, int ndyn, uint32_t logcfg = 0xffffffffu) {
  if (k.logf == 1 && logcfg != kLogPimm) return false;
  const uint32_t feat = k.fm & F_ALL;
  if ((feat & flags) != flags) return false;
  if ((k.fm & F_EXACT) && (feat != flags || ndyn > DREG)) return false;  // exact kernels keep every record in registers
  if ((k.fm & F_KSTEP) && (kinds & ~KIND_STEP)) return false;
  if ((k.fm & F_KSINE) && (kinds & ~KIND_SINE)) return false;
  if ((k.fm & F_KTRAP) && (kinds & ~KIND_TRAP)) return false;
  if (k.nd1 && k.nd1 != ndyn + 1) return false;
  return true;
}

// Lower is better: fewer layers of capacity, then exact feature sets, then
// driver-kind specialisations, then a compiled-in record count, then fewer
// compiled-in features.
static int rank(const KEntry& k) {
  return k.lm * 1000 + ((k.fm & F_EXACT) ? 0 : 200) + ((k.fm & (F_KSTEP | F_KSINE | F_KTRAP)) ? 0 : 100) +
         (k.nd1 ? 0 : 50) + (k.logf ? 0 : 25) + __builtin_popcount(k.fm & F_ALL);
}

// SIMDs of the target (MI355X: 256 CUs x 4): a batch whose waves number at
// most this many runs one wave per SIMD.
constexpr uint64_t kSimds = 1024;

static const KEntry* pick_kernel(int prec, int nl, int solver, uint32_t flags, uint32_t kinds, int ndyn, bool loge,
                                 int lanes, uint64_t n_inst, uint32_t logcfg) {
  static const KTable kTables[] = {kernels_f64(),    kernels_f64e(),    kernels_f32(),
                                   kernels_spec(),   kernels_ll_f64(),  kernels_ll_f32(),
                                   kernels_ll_spec(), kernels_exact(),  kernels_cl(),
                                   kernels_cl64(),    kernels_dp()};
  // [0] one lane per instance, [1] one lane per layer, [2] one lane per
  // component; cl64: the binary64 one-lane-per-component kernel of an FP32
  // group that no binary32 one covers
  const KEntry* best[3] = {nullptr, nullptr, nullptr};
  const KEntry* cl64 = nullptr;
  for (const KTable& t : kTables)
    for (size_t q = 0; q < t.n; ++q) {
      const KEntry& k = t.e[q];
      if (k.solver != solver || k.loge != loge || k.lm < nl) continue;
      if (!covers(k, flags, kinds, ndyn, logcfg)) continue;
      const int cat = (k.lpi == 1) ? 0 : (k.lpi == k.lm ? 1 : 2);
      if (k.prec != prec) {
        if (prec == CMTJ_FP32 && k.prec == CMTJ_FP64 && cat == 2 && solver == CMTJ_RK4 &&
            (!cl64 || rank(k) < rank(*cl64)))
          cl64 = &k;
        continue;
      }
      const KEntry*& b = best[cat];
      if (!b || rank(k) < rank(*b)) b = &k;
    }
  // FP32 mode, RK4: where only the binary64 component kernel covers a group
  // (spin torques, scalar drives), it stands in for the missing binary32
  // one: more accurate, and on batches that leave SIMDs idle faster than
  // binary32 one lane per layer (a 512-point STO sweep: 5.9e8 vs 3.2e8
  // junction-steps/s).  Not for Euler-Heun: its binary64 normals cost more
  // than the binary32 ones of one lane per layer.
  if (!best[2] && cl64) best[2] = cl64;
  if (lanes == CMTJ_LANES_ONE && best[0]) return best[0];
  if (lanes == CMTJ_LANES_PER_COMPONENT && best[2]) return best[2];
  // automatic choice: on a batch too small to fill the chip, the time of a
  // sweep is its longest instance's step count times the latency of one
  // wave step.  One lane per component issues about half the instructions
  // per wave step of one lane per layer, at four times the lanes: it wins
  // while its waves still fit one per SIMD, and costs about twice the
  // instructions per junction-step beyond that.
  if (lanes == CMTJ_LANES_AUTO && best[2] && n_inst * (uint64_t)best[2]->lpi <= 64 * kSimds) return best[2];
  // one lane per layer for multilayer instances: on full-chip batches it
  // measured no slower than one lane per instance (1M-instance CIMS ensemble
  // 6.6e10 vs 6.3e10 junction-steps/s, 65,536-point PIMM sweep 2.56e10 vs
  // 2.44e10 on MI355X): its register footprint is smaller, so more waves fit
  // per SIMD.
  if (best[1] && (nl >= 2 || !best[0])) return best[1];
  if (best[0]) return best[0];
  return best[1] ? best[1] : best[2];
}

// ---------------------------------------------------------------------------
// descriptor analysis

static int n_channels(int nl, int mr, int mask) {
  int n = 0;
  if (mask & CMTJ_LOG_MX) n += nl;
  if (mask & CMTJ_LOG_MYZ) n += 2 * nl;
  if (mask & CMTJ_LOG_MR) n += (mr == CMTJ_MR_CLASSIC) ? 1 : (mr == CMTJ_MR_STRIP ? 2 : 0);
  if (mask & CMTJ_LOG_ENERGY) n += CMTJ_ENERGY_CHANNELS * nl;
  if (mask & CMTJ_LOG_TIME) n += 1;
  return n;
}

static bool live(const cmtj_driver& d) { return d.type != CMTJ_CONSTANT || d.constant != 0.0; }

static uint32_t junction_flags(const cmtj_junction& j, const cmtj_layer* L) {
  uint32_t f = 0;
  const int nl = j.n_layers;
  for (int l = 0; l < nl; ++l) {
    const cmtj_layer& y = L[l];
    const double* N = y.demag;
    if (N[1] != 0 || N[2] != 0 || N[3] != 0 || N[5] != 0 || N[6] != 0 || N[7] != 0) f |= F_DEMAG_OFF;
    if (nl > 1)
      for (int q = 0; q < 9; ++q)
        if (y.dipole_top[q] != 0 || y.dipole_bottom[q] != 0) f |= F_DIPOLE;
    if (live(y.drv[CMTJ_SLOT_K])) f |= F_ANIS;
    if (live(y.drv[CMTJ_SLOT_K2])) f |= F_ANIS2;
    if (nl > 1 && (live(y.drv[CMTJ_SLOT_J_TOP]) || live(y.drv[CMTJ_SLOT_J_BOTTOM]))) f |= F_IEC;
    if (nl > 1 && (live(y.drv[CMTJ_SLOT_J2_TOP]) || live(y.drv[CMTJ_SLOT_J2_BOTTOM]))) f |= F_IEC2;
    if (nl > 1)
      for (int q = 0; q < 3; ++q)
        if (live(y.drv[CMTJ_SLOT_IDMI_TOP + q]) || live(y.drv[CMTJ_SLOT_IDMI_BOTTOM + q]))
          f |= F_IDMI;
    if (y.include_stt) f |= F_STT;
    if (!y.include_stt && y.include_sot) {
      f |= F_SOT;
      if (y.dynamic_sot && (live(y.drv[CMTJ_SLOT_FL2]) || live(y.drv[CMTJ_SLOT_DL2]))) f |= F_SOT2;
    }
    if (y.temperature_set && j.solver != CMTJ_RK4) f |= F_THERMAL;
    for (int s = 0; s < CMTJ_NSLOT; ++s) {
      if (y.drv[s].type != CMTJ_CONSTANT) f |= (s < CMTJ_SLOT_K) ? F_DYNH : F_DYN;
      if (y.drv[s].type == CMTJ_TABLE) f |= F_TABLE;
    }
  }
  return f;
}

static int validate(const cmtj_junction& j, size_t idx, const cmtj_layer* layers, size_t rows,
                    int prec, size_t n_table) {
  const std::string at = " (junction " + std::to_string(idx) + ")";
  if (j.n_layers < 1 || j.n_layers > CMTJ_MAX_LAYERS)
    return fail(CMTJ_ERR_INVALID, "n_layers must be in [1, 8]" + at);
  if ((size_t)j.layer_offset + (size_t)j.n_layers > rows)
    return fail(CMTJ_ERR_INVALID, "layer rows out of range" + at);
  const bool dp = j.solver == CMTJ_DORMAND_PRINCE;
  if (j.solver != CMTJ_RK4 && j.solver != CMTJ_HEUN && j.solver != CMTJ_EULER_HEUN && !dp)
    return fail(CMTJ_ERR_INVALID, "unknown solver" + at);
  if (!(j.dt > 0.0) || !std::isfinite(j.dt)) return fail(CMTJ_ERR_INVALID, "time step must be > 0" + at);
  if (dp) {
    // adaptive steps: 1-4 layers, binary64, closed-form drivers only
    // (callback tables are sampled at fixed stage times), no energy
    // channels, the time logged per entry
    if (j.n_layers > 4) return fail(CMTJ_ERR_INVALID, "Dormand-Prince runs take 1 to 4 layers" + at);
    if (prec != CMTJ_FP64) return fail(CMTJ_ERR_INVALID, "Dormand-Prince runs are FP64" + at);
    if (j.log_mask & CMTJ_LOG_ENERGY)
      return fail(CMTJ_ERR_INVALID, "no energy channels with Dormand-Prince" + at);
    if (j.log_mask && !(j.log_mask & CMTJ_LOG_TIME))
      return fail(CMTJ_ERR_INVALID, "a Dormand-Prince log needs the time channel" + at);
    if (!(j.total_time >= 0.0) || !std::isfinite(j.total_time) || !(j.write_freq > 0.0))
      return fail(CMTJ_ERR_INVALID, "total_time >= 0 and write_freq > 0 needed" + at);
    for (int l = 0; l < j.n_layers; ++l)
      for (int s = 0; s < CMTJ_NSLOT; ++s)
        if (layers[j.layer_offset + l].drv[s].type == CMTJ_TABLE)
          return fail(CMTJ_ERR_INVALID, "callback drivers need fixed stage times (not Dormand-Prince)" + at);
  } else {
    if (j.log_mask & CMTJ_LOG_TIME) return fail(CMTJ_ERR_INVALID, "the time channel is for Dormand-Prince" + at);
    if (j.write_every < 1) return fail(CMTJ_ERR_INVALID, "write_every must be >= 1" + at);
    const uint64_t ne = j.n_steps == 0 ? 0 : ((uint64_t)j.n_steps + j.write_every - 1) / j.write_every;
    if (ne != j.n_entries) return fail(CMTJ_ERR_INVALID, "n_entries != ceil(n_steps / write_every)" + at);
  }
  if ((j.log_mask & CMTJ_LOG_ENERGY) && prec != CMTJ_FP64)
    return fail(CMTJ_ERR_INVALID, "energy channels need an FP64 run" + at);
  if (j.mr_mode == CMTJ_MR_CLASSIC && j.n_layers > 2)
    return fail(CMTJ_ERR_INVALID, "CLASSIC magnetoresistance supports 1 or 2 layers" + at);
  const int sps = j.solver == CMTJ_RK4 ? 3 : (j.solver == CMTJ_HEUN ? 2 : 1);
  for (int l = 0; l < j.n_layers; ++l) {
    const cmtj_layer& y = layers[j.layer_offset + l];
    if (y.Ms * y.thickness == 0.0) return fail(CMTJ_ERR_INVALID, "Failed to divide vector by zero!");
    if (y.temperature_set && j.solver != CMTJ_RK4 && y.cell_surface * y.thickness == 0.0)
      return fail(CMTJ_ERR_INVALID, "Cell surface cannot be 0 during temp. calculations!");
    for (int s = 0; s < CMTJ_NSLOT; ++s) {
      const cmtj_driver& d = y.drv[s];
      if (d.type == CMTJ_TABLE) {
        if (d.table_index < 0 || (uint64_t)d.table_index + (uint64_t)j.n_steps * sps > n_table)
          return fail(CMTJ_ERR_INVALID, "driver table out of range" + at);
      } else if (d.type < 0 || d.type > CMTJ_GAUSSSTEP) {
        return fail(CMTJ_ERR_INVALID, "unknown driver type" + at);
      }
    }
  }
  return 0;
}

// Time-dependent records of one junction; identical records on several
// layers merge into one with a layer mask.
struct Rec {
  int32_t meta;
  DynRec num;
};

static void junction_records(const cmtj_junction& j, const cmtj_layer* L, std::vector<Rec>& out) {
  out.clear();
  for (int l = 0; l < j.n_layers; ++l) {
    for (int s = 0; s < CMTJ_NSLOT; ++s) {
      const cmtj_driver& d = L[l].drv[s];
      if (d.type == CMTJ_CONSTANT) continue;
      Rec r;
      std::memset(&r, 0, sizeof(r));
      double A, p1, p2, p3;
      wave_params(d, A, p1, p2, p3);
      r.meta = (d.type & 0xff) | (s << 8) | ((1 << l) << 16);
      r.num.tidx = d.type == CMTJ_TABLE ? d.table_index : 0;
      r.num.A = A;
      r.num.p1 = p1;
      r.num.p2 = p2;
      r.num.p3 = p3;
      bool merged = false;
      for (Rec& o : out) {
        if ((o.meta & 0xffff) == (r.meta & 0xffff) && o.num.tidx == r.num.tidx && o.num.A == r.num.A &&
            o.num.p1 == r.num.p1 && o.num.p2 == r.num.p2 && o.num.p3 == r.num.p3) {
          o.meta |= r.meta & 0xff0000;
          merged = true;
          break;
        }
      }
      if (!merged) out.push_back(r);
    }
  }
}

static int32_t layer_bits(const cmtj_layer& y) {
  return (y.ref_type & 3) | ((y.include_stt ? 1 : (y.include_sot ? 2 : 0)) << 2) |
         ((y.dynamic_sot ? 1 : 0) << 4) | ((y.alternative_stt ? 1 : 0) << 5) |
         ((y.temperature_set ? 1 : 0) << 6);
}

// ---------------------------------------------------------------------------
// plans

struct Group {
  const KEntry* k;
  GroupArgs a;
  uint32_t blocks;
  uint32_t threads = 64;  // per workgroup
  uint32_t* queue = nullptr;  // work counter of a queue launch (k_dp_q), else null
};

// Lanes of a work-queue launch: four waves per SIMD of the chip, or fewer
// when CMTJ_QUEUE_LANES says so (a smaller grid makes lanes refill more).
static uint32_t queue_lanes() {
  static const uint32_t v = [] {
    const char* e = std::getenv("CMTJ_QUEUE_LANES");
    const long x = e ? std::strtol(e, nullptr, 10) : 0;
    return x > 0 ? (uint32_t)((x + 63) / 64 * 64) : (uint32_t)(64 * 1024 * 4);
  }();
  return v;
}

struct cmtj_plan {
  void* arena = nullptr;
  bool owns = true;
  std::vector<Group> groups;
  cmtj_plan_info info;
  uint64_t log_extent = 0;
  // a captured execute (cmtj_plan_graph_create): one kernel node per group
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<GroupArgs> graph_args;
};

static void drop_graph(cmtj_plan* p) {
  if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->graph_exec = nullptr;
  p->graph = nullptr;
  p->graph_args.clear();
}

typedef void* (*ArenaFn)(size_t bytes, void* ctx);

static size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Structure signature of a junction: instances with equal signatures share a
// kernel launch and every structural branch inside it.
static void signature(const cmtj_junction& j, const cmtj_layer* L, const std::vector<Rec>& recs,
                      std::vector<int32_t>& sig) {
  sig.assign({j.n_layers, j.solver, j.log_mask, j.mr_mode, (int32_t)junction_flags(j, L)});
  for (int l = 0; l < j.n_layers; ++l) sig.push_back(layer_bits(L[l]));
  for (const Rec& r : recs) sig.push_back(r.meta);
}

// Builds every group of a plan into one device arena obtained from `get`.
// A group holds the junctions of one structure signature (layer count,
// solver, log layout, per-layer modes, driver kinds and slots), so every
// branch on structure inside the kernel is wave-uniform.
static int build_plan(cmtj_plan* P, const cmtj_junction* js, size_t n, const cmtj_layer* layers,
                      size_t rows, const cmtj_constants* c, const double* tables, size_t n_table,
                      int prec, int log_dtype, int lanes, ArenaFn get, void* gctx, hipStream_t stream,
                      size_t* size_only = nullptr, char* host_view = nullptr) {
  if (n == 0) return fail(CMTJ_ERR_INVALID, "empty batch");
  if (n > 0xFFFFFFFFull) return fail(CMTJ_ERR_INVALID, "too many instances");
  if (prec != CMTJ_FP64 && prec != CMTJ_FP32) return fail(CMTJ_ERR_INVALID, "unknown precision");
  if (log_dtype != CMTJ_LOG_F64 && log_dtype != CMTJ_LOG_F32) return fail(CMTJ_ERR_INVALID, "unknown log dtype");
  std::memset(&P->info, 0, sizeof(P->info));
  uint64_t total_steps = 0, max_steps = 0, extent = 0;
  int max_layers = 0;
  for (size_t i = 0; i < n; ++i) {
    const int r = validate(js[i], i, layers, rows, prec, n_table);
    if (r) return r;
    total_steps += js[i].n_steps;
    max_steps = std::max<uint64_t>(max_steps, js[i].n_steps);
    max_layers = std::max(max_layers, js[i].n_layers);
    const uint64_t e = js[i].log_offset +
                       (uint64_t)n_channels(js[i].n_layers, js[i].mr_mode, js[i].log_mask) * js[i].n_entries;
    extent = std::max(extent, e);
  }
  // signatures
  std::vector<std::vector<int32_t>> sigs;
  std::vector<std::vector<uint32_t>> members;
  std::vector<std::vector<Rec>> recs_of(n);
  std::vector<int32_t> sig;
  for (size_t i = 0; i < n; ++i) {
    const cmtj_junction& j = js[i];
    const cmtj_layer* L = layers + j.layer_offset;
    junction_records(j, L, recs_of[i]);
    if (recs_of[i].size() > (size_t)MAX_DYN)
      return fail(CMTJ_ERR_INVALID, "more than 16 time-dependent drivers in one junction");
    signature(j, L, recs_of[i], sig);
    size_t g = 0;
    for (; g < sigs.size(); ++g)
      if (sigs[g] == sig) break;
    if (g == sigs.size()) {
      sigs.push_back(sig);
      members.emplace_back();
    }
    members[g].push_back((uint32_t)i);
  }
  const size_t G = sigs.size();
  std::vector<std::vector<uint32_t>> orders(G);
  std::vector<std::vector<DynRec>> nums(G);
  std::vector<uint32_t> flags(G);
  for (size_t g = 0; g < G; ++g) {
    std::vector<uint32_t>& ord = orders[g];
    ord = members[g];
    // longest first: the waves of a group finish together and the hardware
    // dispatcher starts the long ones first
    std::stable_sort(ord.begin(), ord.end(),
                     [&](uint32_t a, uint32_t b) { return js[a].n_steps > js[b].n_steps; });
    uint32_t f = 0;
    const size_t nd = recs_of[ord[0]].size();
    nums[g].resize(std::max<size_t>(nd, 1) * ord.size());
    for (size_t p = 0; p < ord.size(); ++p) {
      const cmtj_junction& j = js[ord[p]];
      f |= junction_flags(j, layers + j.layer_offset);
      for (size_t k = 0; k < nd; ++k) nums[g][k * ord.size() + p] = recs_of[ord[p]][k].num;
    }
    flags[g] = f;
  }
  // arena layout
  size_t off = 0;
  const size_t o_js = off;
  off = align_up(off + n * sizeof(cmtj_junction));
  const size_t o_layers = off;
  off = align_up(off + rows * sizeof(cmtj_layer));
  const size_t o_tables = off;
  off = align_up(off + std::max<size_t>(n_table, 1) * sizeof(double));
  const size_t o_queue = off;  // one work counter per group (queue launches)
  off = align_up(off + G * sizeof(uint32_t));
  std::vector<size_t> o_order(G), o_dyn(G);
  for (size_t g = 0; g < G; ++g) {
    o_order[g] = off;
    off = align_up(off + orders[g].size() * sizeof(uint32_t));
    o_dyn[g] = off;
    off = align_up(off + nums[g].size() * sizeof(DynRec));
  }
  if (size_only) {
    *size_only = off;
    return 0;
  }
  char* arena = (char*)get(off, gctx);
  if (!arena) return CMTJ_ERR_HIP;
  P->arena = arena;
  // host-mapped arena: the kernels read the descriptors through the mapping,
  // the host writes them in place (no copy engine involved)
  auto upload = [&](size_t off, const void* src, size_t bytes) -> int {
    if (!bytes) return 0;
    if (host_view) {
      std::memcpy(host_view + off, src, bytes);
      return 0;
    }
    HIP_TRY(hipMemcpyAsync(arena + off, src, bytes, hipMemcpyHostToDevice, stream));
    return 0;
  };
  if (upload(o_js, js, n * sizeof(cmtj_junction))) return CMTJ_ERR_HIP;
  if (upload(o_layers, layers, rows * sizeof(cmtj_layer))) return CMTJ_ERR_HIP;
  if (n_table && upload(o_tables, tables, n_table * sizeof(double))) return CMTJ_ERR_HIP;
  P->groups.clear();
  std::string names;
  int max_lpi = 1;
  for (size_t g = 0; g < G; ++g) {
    if (upload(o_order[g], orders[g].data(), orders[g].size() * sizeof(uint32_t))) return CMTJ_ERR_HIP;
    if (upload(o_dyn[g], nums[g].data(), nums[g].size() * sizeof(DynRec))) return CMTJ_ERR_HIP;
    const cmtj_junction& j0 = js[orders[g][0]];
    const cmtj_layer* L0 = layers + j0.layer_offset;
    const bool loge = (j0.log_mask & CMTJ_LOG_ENERGY) != 0;
    uint32_t kinds = 0;
    for (const Rec& r : recs_of[orders[g][0]]) {
      const int ty = r.meta & 0xff;
      kinds |= (ty == CMTJ_STEP) ? KIND_STEP : (ty == CMTJ_SINE) ? KIND_SINE : (ty == CMTJ_TRAPEZOID) ? KIND_TRAP
                                                                                               : KIND_OTHER;
    }
    const KEntry* k = pick_kernel(prec, j0.n_layers, j0.solver, flags[g], kinds,
                                  (int)recs_of[orders[g][0]].size(), loge,
                                  lanes == CMTJ_LANES_QUEUE ? (int)CMTJ_LANES_AUTO : lanes, orders[g].size(),
                                  (uint32_t)j0.log_mask | ((uint32_t)j0.mr_mode << 8) |
                                      (log_dtype == CMTJ_LOG_F32 ? (1u << 16) : 0u));
    if (!k) return fail(CMTJ_ERR_INVALID, "no kernel for this junction configuration");
    Group Gr;
    Gr.k = k;
    std::memset(&Gr.a, 0, sizeof(Gr.a));
    Gr.a.js = (const cmtj_junction*)(arena + o_js);
    Gr.a.layers = (const cmtj_layer*)(arena + o_layers);
    Gr.a.order = (const uint32_t*)(arena + o_order[g]);
    Gr.a.dyn = (const DynRec*)(arena + o_dyn[g]);
    Gr.a.tables = (const double*)(arena + o_tables);
    Gr.a.n = (uint32_t)orders[g].size();
    Gr.a.flags = flags[g];
    Gr.a.log_f32 = log_dtype == CMTJ_LOG_F32;
    Gr.a.nl = j0.n_layers;
    Gr.a.log_mask = j0.log_mask;
    Gr.a.mr_mode = j0.mr_mode;
    const std::vector<Rec>& r0 = recs_of[orders[g][0]];
    Gr.a.ndyn = (int32_t)r0.size();
    for (int l = 0; l < j0.n_layers; ++l) Gr.a.lbits[l] = layer_bits(L0[l]);
    for (size_t q = 0; q < r0.size(); ++q) Gr.a.dmeta[q] = r0[q].meta;
    Gr.a.c = *c;
    Gr.blocks = (uint32_t)(((uint64_t)Gr.a.n * (uint64_t)k->lpi + 63) / 64);
    if (lanes == CMTJ_LANES_QUEUE && j0.solver == CMTJ_DORMAND_PRINCE && j0.n_layers == 1) {
      Gr.queue = (uint32_t*)(arena + o_queue) + g;
      Gr.blocks = (std::min<uint32_t>(Gr.a.n, queue_lanes()) + 63) / 64;
    }
    max_lpi = std::max(max_lpi, k->lpi);
    if (prec == CMTJ_FP32 && k->prec == CMTJ_FP64) ++P->info.fp64_groups;
    P->groups.push_back(Gr);
    const char* kname = Gr.queue ? "k_dp_q<double,1,F_ALL>" : k->name;
    if (names.find(kname) == std::string::npos) {
      if (!names.empty()) names += ";";
      names += kname;
    }
  }
  // host staging vectors die with this function: finish the copies first
  if (!host_view) HIP_TRY(hipStreamSynchronize(stream));
  P->info.n_instances = n;
  P->info.log_elements = extent;
  P->info.total_steps = total_steps;
  P->info.max_steps = max_steps;
  P->info.precision = prec;
  P->info.log_dtype = log_dtype;
  P->info.max_layers = max_layers;
  P->info.n_kernels = (int32_t)P->groups.size();
  P->info.lanes_per_instance = max_lpi;
  std::strncpy(P->info.kernel_name, names.c_str(), sizeof(P->info.kernel_name) - 1);
  P->log_extent = extent;
  return 0;
}

static int launch_plan(cmtj_plan* P, void* d_log, double* d_mag, cmtj_status* d_status, hipStream_t s,
                       cmtj_state* d_state = nullptr, uint32_t slice = 0, void* const* streams = nullptr,
                       int n_streams = 0) {
  g_last_kernels.clear();
  for (size_t q = 0; q < P->groups.size(); ++q) {
    Group& G = P->groups[q];
    g_last_kernels += (G.queue ? std::string("k_dp_q<double,1,F_ALL>") : std::string(G.k->name)) + ";";
    GroupArgs a = G.a;
    a.log = d_log;
    a.mag_out = d_mag;
    a.status = d_status;
    a.state = d_state;
    a.slice = slice;
    void* args[] = {&a};
    const hipStream_t st = n_streams > 0 ? (hipStream_t)streams[q % (size_t)n_streams] : s;
    if (G.queue) {
      uint32_t* ctr = G.queue;
      uint32_t one = 1;
      void* cargs[] = {&ctr, &one};
      HIP_TRY(hipLaunchKernel(clear_u32_kernel(), dim3(1), dim3(64), cargs, 0, st));
      void* qargs[] = {&a, &ctr};
      HIP_TRY(hipLaunchKernel(dp_queue_kernel(), dim3(G.blocks), dim3(G.threads), qargs, 0, st));
      continue;
    }
    HIP_TRY(hipLaunchKernel(G.k->fn, dim3(G.blocks), dim3(G.threads), args, 0, st));
  }
  return 0;
}

// ---------------------------------------------------------------------------
// the drop-in context: one stream, grow-only device buffers

namespace {
struct PlainCtx {
  std::mutex mu;
  hipStream_t stream = nullptr;
  int device = -1;
  void* arena = nullptr;
  size_t arena_bytes = 0;
  void* out = nullptr;
  size_t out_bytes = 0;
};
PlainCtx g_plain;

void* grow(void*& p, size_t& have, size_t need) {
  if (need <= have) return p;
  const size_t want = std::max(need, have * 2);
  if (p) (void)hipFree(p);
  p = nullptr;
  have = 0;
  if (hipMalloc(&p, want) != hipSuccess) {
    g_err = "hipMalloc failed for " + std::to_string(want) + " bytes";
    p = nullptr;
    return nullptr;
  }
  have = want;
  return p;
}

void* plain_arena(size_t bytes, void* ctx) {
  PlainCtx* c = (PlainCtx*)ctx;
  return grow(c->arena, c->arena_bytes, bytes);
}

struct FixedArena {
  void* p;
  size_t bytes;
};

void* fixed_arena(size_t bytes, void* ctx) {
  const FixedArena* a = (const FixedArena*)ctx;
  if (bytes > a->bytes) {
    g_err = "plan arena too small: " + std::to_string(bytes) + " bytes needed, " + std::to_string(a->bytes) +
            " given";
    return nullptr;
  }
  return a->p;
}

void* malloc_arena(size_t bytes, void*) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    g_err = "hipMalloc failed for " + std::to_string(bytes) + " bytes";
    return nullptr;
  }
  return p;
}
}  // namespace

// ---------------------------------------------------------------------------
// C ABI

// ---------------------------------------------------------------------------
// spin-diode mixing voltage: compute_sd (cmtj/utils/resistance.py:12-22) over
// many series, one lane per series.  x_k = -sin(2 pi f t_k) R_k with
// t_k = k dt, y = the IIR filter (b, a) of x in scipy.signal.lfilter's
// transposed direct form II (zero initial state), V = mean of y.  Operation
// order as in lfilter (no FMA contraction); the mean is a compensated sum.

constexpr int kMaxTaps = 9;
struct IirTaps {
  double b[kMaxTaps], a[kMaxTaps];
  int n;
};

__global__ __launch_bounds__(64) void k_vsd_mix(const double* __restrict__ R, uint64_t ld, uint32_t n, uint32_t len,
                                                const double* __restrict__ freq, double dt, IirTaps f,
                                                double* __restrict__ out) {
#pragma clang fp contract(off)
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const double* r = R + (uint64_t)q * ld;
  const double w = (2.0 * M_PI) * freq[q];
  double z[kMaxTaps - 1];
#pragma unroll
  for (int i = 0; i < kMaxTaps - 1; ++i) z[i] = 0.0;
  double sum = 0.0, comp = 0.0;
  const int nt = f.n;
  for (uint32_t k = 0; k < len; ++k) {
    const double x = -sin(w * ((double)k * dt)) * r[k];
    double y;
    if (nt > 1) {
      y = z[0] + f.b[0] * x;
#pragma unroll
      for (int i = 0; i < kMaxTaps - 2; ++i)
        if (i < nt - 2) z[i] = (z[i + 1] + x * f.b[i + 1]) - y * f.a[i + 1];
#pragma unroll
      for (int i = 0; i < kMaxTaps - 1; ++i)
        if (i == nt - 2) z[i] = x * f.b[i + 1] - y * f.a[i + 1];
    } else {
      y = x * f.b[0];
    }
    // Neumaier summation
    const double tsum = sum + y;
    comp += (fabs(sum) >= fabs(y)) ? (sum - tsum) + y : (y - tsum) + sum;
    sum = tsum;
  }
  out[q] = len ? (sum + comp) / (double)len : 0.0;
}

extern "C" {

int cmtj_abi(cmtj_abi_info* out) {
  if (!out) return fail(CMTJ_ERR_INVALID, "null pointer");
  out->version = CMTJ_ABI_VERSION;
  out->max_layers = CMTJ_MAX_LAYERS;
  out->nslot = CMTJ_NSLOT;
  out->sizeof_driver = (int32_t)sizeof(cmtj_driver);
  out->sizeof_layer = (int32_t)sizeof(cmtj_layer);
  out->sizeof_junction = (int32_t)sizeof(cmtj_junction);
  out->sizeof_constants = (int32_t)sizeof(cmtj_constants);
  out->sizeof_status = (int32_t)sizeof(cmtj_status);
  out->sizeof_state = (int32_t)sizeof(cmtj_state);
  out->sizeof_stack = (int32_t)sizeof(cmtj_stack);
  out->sizeof_group = (int32_t)sizeof(cmtj_group);
  out->offsetof_layer_drv = (int32_t)offsetof(cmtj_layer, drv);
  out->offsetof_layer_ref_type = (int32_t)offsetof(cmtj_layer, ref_type);
  out->offsetof_junction_dt = (int32_t)offsetof(cmtj_junction, dt);
  out->offsetof_junction_AHE = (int32_t)offsetof(cmtj_junction, AHE);
  return 0;
}

const char* cmtj_last_error(void) { return g_err.c_str(); }
const char* cmtj_last_kernels(void) { return g_last_kernels.c_str(); }

int cmtj_constants_get(cmtj_constants* out) {
  if (!out) return fail(CMTJ_ERR_INVALID, "null pointer");
  std::lock_guard<std::mutex> lk(g_consts_mu);
  *out = g_consts;
  return 0;
}

int cmtj_constants_set(const cmtj_constants* in) {
  if (!in) return fail(CMTJ_ERR_INVALID, "null pointer");
  std::lock_guard<std::mutex> lk(g_consts_mu);
  g_consts = *in;
  return 0;
}

int cmtj_constants_reset(void) {
  std::lock_guard<std::mutex> lk(g_consts_mu);
  g_consts = kDefaults;
  return 0;
}

int cmtj_driver_eval(const cmtj_driver* d, double t, double* out) {
  if (!d || !out) return fail(CMTJ_ERR_INVALID, "null pointer");
  if (d->type == CMTJ_TABLE) return fail(CMTJ_ERR_INVALID, "tabulated drivers have no closed form");
  if (d->type < 0 || d->type > CMTJ_GAUSSSTEP) return fail(CMTJ_ERR_INVALID, "unknown driver type");
  *out = driver_value(*d, t);
  return 0;
}

// The host-buffer entry points share one stream and buffers per process,
// made again when the current device changes; the caller holds g_plain.mu.
static int plain_bind() {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (g_plain.stream == nullptr || g_plain.device != dev) {
    if (g_plain.stream) (void)hipStreamDestroy(g_plain.stream);
    if (g_plain.arena) (void)hipFree(g_plain.arena);
    if (g_plain.out) (void)hipFree(g_plain.out);
    g_plain.arena = g_plain.out = nullptr;
    g_plain.arena_bytes = g_plain.out_bytes = 0;
    HIP_TRY(hipStreamCreateWithFlags(&g_plain.stream, hipStreamNonBlocking));
    g_plain.device = dev;
  }
  return 0;
}

// Doubles of mag_out_host that a one-junction run fills: the final m of
// every layer, then the PID state of every layer when a Dormand-Prince
// layer runs the PID controller (the kernels' mag_out block layout; at most
// 5 * 4 of its 3 * CMTJ_MAX_LAYERS doubles, since build_plan caps
// Dormand-Prince junctions at 4 layers).
static size_t run_mag_doubles(const cmtj_junction& j, const cmtj_layer* layers) {
  const size_t nl = (size_t)j.n_layers;
  if (j.solver == CMTJ_DORMAND_PRINCE)
    for (size_t l = 0; l < nl; ++l)
      if (layers[l].dp_use_pid) return 5 * nl;
  return 3 * nl;
}

int cmtj_run(const cmtj_junction* j, const cmtj_layer* layers, const cmtj_constants* c,
             const double* tables, size_t n_table, double* log_host, size_t log_cap,
             double* mag_out_host, cmtj_status* st) {
  if (!j || !layers || !c || !mag_out_host || !st) return fail(CMTJ_ERR_INVALID, "null pointer");
  std::lock_guard<std::mutex> lk(g_plain.mu);
  if (const int rb = plain_bind()) return rb;
  cmtj_junction jj = *j;
  jj.layer_offset = 0;
  jj.log_offset = 0;
  const size_t nch = (size_t)n_channels(jj.n_layers, jj.mr_mode, jj.log_mask);
  const size_t nlog = nch * jj.n_entries;
  if (nlog > 0 && (!log_host || log_cap < nlog)) return fail(CMTJ_ERR_INVALID, "log buffer too small");
  cmtj_plan P;
  P.owns = false;
  int r = build_plan(&P, &jj, 1, layers, (size_t)jj.n_layers, c, tables, n_table, CMTJ_FP64, CMTJ_LOG_F64,
                     CMTJ_LANES_AUTO, plain_arena, &g_plain, g_plain.stream);
  if (r) return r;
  const size_t o_mag = align_up(nlog * sizeof(double));
  const size_t o_st = align_up(o_mag + sizeof(double) * 3 * CMTJ_MAX_LAYERS);
  const size_t out_bytes = o_st + sizeof(cmtj_status);
  char* out = (char*)grow(g_plain.out, g_plain.out_bytes, out_bytes);
  if (!out) return CMTJ_ERR_HIP;
  r = launch_plan(&P, nlog ? out : nullptr, (double*)(out + o_mag), (cmtj_status*)(out + o_st), g_plain.stream);
  if (r) return r;
  if (nlog)
    HIP_TRY(hipMemcpyAsync(log_host, out, nlog * sizeof(double), hipMemcpyDeviceToHost, g_plain.stream));
  double mag[3 * CMTJ_MAX_LAYERS];
  HIP_TRY(hipMemcpyAsync(mag, out + o_mag, sizeof(mag), hipMemcpyDeviceToHost, g_plain.stream));
  HIP_TRY(hipMemcpyAsync(st, out + o_st, sizeof(cmtj_status), hipMemcpyDeviceToHost, g_plain.stream));
  HIP_TRY(hipStreamSynchronize(g_plain.stream));
  std::memcpy(mag_out_host, mag, sizeof(double) * run_mag_doubles(jj, layers));
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// the allocation-free drop-in: descriptors, log staging, final m and status
// of one junction live in one file-scope device array.  Host code never
// passes the array's name to a kernel or a copy: it uses the address
// hipGetSymbolAddress returns for the current device, taken once per device.

static __device__ __align__(256) char scratch[CMTJ_STATIC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_static_mu;
char* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Address of this device's copy of `scratch`; the caller holds g_static_mu.
static int static_scratch(char** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(CMTJ_ERR_INVALID, "device ordinal out of range");
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(CMTJ_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
    g_scratch_ptr[dev] = (char*)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

extern "C" {

int cmtj_run_noalloc(const cmtj_junction* j, const cmtj_layer* layers, const cmtj_constants* c,
                     const double* tables, size_t n_table, double* log_host, size_t log_cap,
                     double* mag_out_host, cmtj_status* st, void* stream) {
  if (!j || !layers || !c || !mag_out_host || !st) return fail(CMTJ_ERR_INVALID, "null pointer");
  std::lock_guard<std::mutex> lk(g_static_mu);
  char* base = nullptr;
  if (const int rs = static_scratch(&base)) return rs;
  const hipStream_t s = (hipStream_t)stream;
  cmtj_junction jj = *j;
  jj.layer_offset = 0;
  jj.log_offset = 0;
  const size_t nch = (size_t)n_channels(jj.n_layers, jj.mr_mode, jj.log_mask);
  const size_t nlog = nch * jj.n_entries;
  if (nlog > 0 && (!log_host || log_cap < nlog)) return fail(CMTJ_ERR_INVALID, "log buffer too small");
  // scratch layout: [plan arena][log nlog doubles][final m][status]
  cmtj_plan P;
  P.owns = false;
  size_t arena = 0;
  int r = build_plan(&P, &jj, 1, layers, (size_t)jj.n_layers, c, tables, n_table, CMTJ_FP64, CMTJ_LOG_F64,
                     CMTJ_LANES_AUTO, nullptr, nullptr, s, &arena);
  if (r) return r;
  const size_t o_log = align_up(arena);
  const size_t o_mag = align_up(o_log + nlog * sizeof(double));
  const size_t o_st = align_up(o_mag + sizeof(double) * 3 * CMTJ_MAX_LAYERS);
  const size_t need = o_st + sizeof(cmtj_status);
  if (need > (size_t)CMTJ_STATIC_SCRATCH_BYTES)
    return fail(CMTJ_ERR_INVALID, "run needs " + std::to_string(need) + " bytes of device scratch, more than the " +
                                      std::to_string((size_t)CMTJ_STATIC_SCRATCH_BYTES) +
                                      " of cmtj_run_noalloc (use cmtj_run)");
  FixedArena fa{base, arena};
  r = build_plan(&P, &jj, 1, layers, (size_t)jj.n_layers, c, tables, n_table, CMTJ_FP64, CMTJ_LOG_F64,
                 CMTJ_LANES_AUTO, fixed_arena, &fa, s);
  if (r) return r;
  r = launch_plan(&P, nlog ? base + o_log : nullptr, (double*)(base + o_mag), (cmtj_status*)(base + o_st), s);
  if (r) return r;
  if (nlog) HIP_TRY(hipMemcpyAsync(log_host, base + o_log, nlog * sizeof(double), hipMemcpyDeviceToHost, s));
  double mag[3 * CMTJ_MAX_LAYERS];
  HIP_TRY(hipMemcpyAsync(mag, base + o_mag, sizeof(mag), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(st, base + o_st, sizeof(cmtj_status), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  std::memcpy(mag_out_host, mag, sizeof(double) * run_mag_doubles(jj, layers));
  return 0;
}

int cmtj_run_stacks(const cmtj_stack* stacks, size_t n_stacks, const cmtj_junction* js, const cmtj_layer* layers,
                    size_t n_layer_rows, const cmtj_constants* c, const double* tables, size_t n_table,
                    double* log_host, size_t log_cap, double* mag_out_host, cmtj_status* st_host) {
  if (!stacks || !js || !layers || !c || !mag_out_host || !st_host) return fail(CMTJ_ERR_INVALID, "null pointer");
  if (n_stacks == 0) return 0;
  if (n_stacks > (1u << 24)) return fail(CMTJ_ERR_INVALID, "too many stacks");
  const int n = stacks[0].n_junctions;
  if (n < 2 || n > CMTJ_MAX_STACK)
    return fail(CMTJ_ERR_INVALID, "a stack has 2.." + std::to_string(CMTJ_MAX_STACK) + " junctions");
  for (size_t q = 0; q < n_stacks; ++q) {
    const cmtj_stack& s = stacks[q];
    if (s.n_junctions != n) return fail(CMTJ_ERR_INVALID, "every stack of a batch has the same member count");
    if (s.current.type == CMTJ_TABLE) {
      // a callback current, sampled once per step (KCL only)
      if (!s.kcl)
        return fail(CMTJ_ERR_INVALID, "a non-KCL stack needs a closed-form current (operator* drops a callback)");
      if (s.current.table_index < 0 || !tables ||
          (size_t)s.current.table_index + (size_t)js[0].n_steps > n_table)
        return fail(CMTJ_ERR_INVALID, "stack current table out of range");
    } else if (s.current.type < 0 || s.current.type > CMTJ_GAUSSSTEP) {
      return fail(CMTJ_ERR_INVALID, "unknown stack current driver kind");
    }
    if (!s.kcl && !s.delayed && !s.two_layer)
      return fail(CMTJ_ERR_INVALID, "non-delayed coupling is only supported for 2 layer stacks");
  }
  const size_t nj = n_stacks * (size_t)n;
  const cmtj_junction& j0 = js[0];
  if (j0.solver != CMTJ_RK4 && j0.solver != CMTJ_EULER_HEUN)
    return fail(CMTJ_ERR_INVALID, "stacks step with RK4 or Euler-Heun");
  size_t need = 0;
  for (size_t q = 0; q < nj; ++q) {
    const cmtj_junction& j = js[q];
    const cmtj_stack& s = stacks[q / (size_t)n];
    const int m = (int)(q % (size_t)n);
    // equal step counts keep the plan's instance order the caller's
    if (j.solver != j0.solver || j.n_steps != j0.n_steps || j.write_every != j0.write_every ||
        j.n_entries != j0.n_entries || j.dt != j0.dt)
      return fail(CMTJ_ERR_INVALID, "stack members must share solver, step count, time step and log cadence");
    if (j.n_layers < 1 || j.n_layers > 2) return fail(CMTJ_ERR_INVALID, "stack members have 1 or 2 layers");
    if (s.top[m] < 0 || s.top[m] >= j.n_layers || (s.two_layer && (s.bottom[m] < 0 || s.bottom[m] >= j.n_layers)))
      return fail(CMTJ_ERR_INVALID, "top/bottom layer index out of range");
    const size_t nch = (size_t)n_channels(j.n_layers, j.mr_mode, j.log_mask) + 2;
    need = std::max<size_t>(need, j.log_offset + nch * j.n_entries);
  }
  if (need > 0 && (!log_host || log_cap < need)) return fail(CMTJ_ERR_INVALID, "log buffer too small");
  uint32_t pad = 1;
  while (pad < (uint32_t)n) pad <<= 1;

  std::lock_guard<std::mutex> lk(g_plain.mu);
  if (const int rb = plain_bind()) return rb;
  cmtj_plan P;
  P.owns = false;
  int r = build_plan(&P, js, nj, layers, n_layer_rows, c, tables, n_table, CMTJ_FP64, CMTJ_LOG_F64,
                     CMTJ_LANES_ONE, plain_arena, &g_plain, g_plain.stream);
  if (r) return r;
  if (P.groups.size() != 1)
    return fail(CMTJ_ERR_INVALID,
                "the junctions of a stack must share one structure (layers, modes, driver kinds)");
  // the smallest stack kernel that covers the structure
  const KTable kt = kernels_stack();
  const KEntry* ke = nullptr;
  for (size_t q = 0; q < kt.n; ++q) {
    const KEntry& e = kt.e[q];
    if (e.solver == j0.solver && e.lm >= P.groups[0].a.nl &&
        covers(e, P.groups[0].a.flags, 0, P.groups[0].a.ndyn) && (!ke || rank(e) < rank(*ke)))
      ke = &e;
  }
  if (!ke) return fail(CMTJ_ERR_INVALID, "no stack kernel for this solver");
  // one lane per vector component (4 per layer) when every stack fits a
  // wave, none couples without delay (its members step one after another:
  // the one-lane form's sub-steps), and the stacks leave SIMDs free: a step
  // of a stack then takes the component kernel's latency; a sweep larger
  // than the chip keeps one lane per member (fewer instructions per step)
  bool any_nd = false;
  for (size_t q = 0; q < n_stacks; ++q) any_nd |= (!stacks[q].kcl && !stacks[q].delayed);
  if (!any_nd) {
    const KTable kc = kernels_stack_cl();
    const KEntry* kce = nullptr;
    for (size_t q = 0; q < kc.n; ++q) {
      const KEntry& e = kc.e[q];
      if (e.solver == j0.solver && e.lm >= P.groups[0].a.nl &&
          covers(e, P.groups[0].a.flags, 0, P.groups[0].a.ndyn) && (!kce || rank(e) < rank(*kce)))
        kce = &e;
    }
    if (kce) {
      uint32_t pc = 1;
      while (pc < (uint32_t)(n * kce->lpi)) pc <<= 1;
      if (pc <= 64 && (uint64_t)n_stacks * pc <= kSimds * 64) {
        ke = kce;
        pad = pc;
      }
    }
  }
  const size_t o_mag = align_up(need * sizeof(double));
  const size_t o_st = align_up(o_mag + sizeof(double) * 3 * CMTJ_MAX_LAYERS * nj);
  const size_t o_stk = align_up(o_st + sizeof(cmtj_status) * nj);
  const size_t out_bytes = o_stk + sizeof(cmtj_stack) * n_stacks;
  char* out = (char*)grow(g_plain.out, g_plain.out_bytes, out_bytes);
  if (!out) return CMTJ_ERR_HIP;
  HIP_TRY(hipMemcpyAsync(out + o_stk, stacks, sizeof(cmtj_stack) * n_stacks, hipMemcpyHostToDevice,
                         g_plain.stream));
  Group& G = P.groups[0];
  G.k = ke;
  G.blocks = (uint32_t)((n_stacks * pad + 63) / 64);
  G.a.stack = (const cmtj_stack*)(out + o_stk);
  G.a.stack_n = (uint32_t)n_stacks;
  G.a.stack_pad = pad;
  r = launch_plan(&P, need ? out : nullptr, (double*)(out + o_mag), (cmtj_status*)(out + o_st), g_plain.stream);
  if (r) return r;
  if (need) HIP_TRY(hipMemcpyAsync(log_host, out, need * sizeof(double), hipMemcpyDeviceToHost, g_plain.stream));
  HIP_TRY(hipMemcpyAsync(mag_out_host, out + o_mag, sizeof(double) * 3 * CMTJ_MAX_LAYERS * nj,
                         hipMemcpyDeviceToHost, g_plain.stream));
  HIP_TRY(hipMemcpyAsync(st_host, out + o_st, sizeof(cmtj_status) * nj, hipMemcpyDeviceToHost, g_plain.stream));
  HIP_TRY(hipStreamSynchronize(g_plain.stream));
  return 0;
}

int cmtj_run_groups(const cmtj_group* groups, size_t n_groups, const cmtj_junction* js, const cmtj_layer* layers,
                    size_t n_layer_rows, const cmtj_constants* c, const double* pair_terms, size_t n_pair_terms,
                    double* log_host, size_t log_cap, double* mag_out_host, cmtj_status* st_host) {
  if (!groups || !js || !layers || !c || !mag_out_host || !st_host) return fail(CMTJ_ERR_INVALID, "null pointer");
  if (n_groups == 0) return 0;
  if (n_groups > (1u << 24)) return fail(CMTJ_ERR_INVALID, "too many groups");
  const int n = groups[0].n_elements;
  if (n < 1 || n > CMTJ_MAX_GROUP)
    return fail(CMTJ_ERR_INVALID, "a group has 1.." + std::to_string(CMTJ_MAX_GROUP) + " elements");
  const size_t nj = n_groups * (size_t)n;
  const cmtj_junction& j0 = js[0];
  if (j0.solver != CMTJ_RK4 && j0.solver != CMTJ_EULER_HEUN)
    return fail(CMTJ_ERR_INVALID, "groups step with RK4 or Euler-Heun");
  std::vector<cmtj_layer> rows(layers, layers + n_layer_rows);
  size_t need = 0;
  for (size_t q = 0; q < n_groups; ++q) {
    const cmtj_group& G = groups[q];
    if (G.n_elements != n) return fail(CMTJ_ERR_INVALID, "every group of a batch has the same element count");
    if (G.interaction < CMTJ_INTERACTION_NULL || G.interaction > CMTJ_INTERACTION_TENSOR)
      return fail(CMTJ_ERR_INVALID, "unknown interaction");
    if (G.interaction == CMTJ_INTERACTION_TENSOR) {
      if (!pair_terms || n_pair_terms < n_groups * (size_t)n * 12 * (size_t)n)
        return fail(CMTJ_ERR_INVALID, "pair terms missing or too short for the groups");
      continue;  // the callable was evaluated (and checked) by the caller
    }
    if (G.interaction == CMTJ_INTERACTION_NULL || j0.n_steps == 0) continue;
    // computeDipoleInteraction / getDipoleTensorFromRelPositions throw on
    // the first step (core/reservoir.hpp:83-86, :39-42)
    const double lim = (G.interaction == CMTJ_INTERACTION_KANAO) ? 1e-10 : 1e-14;
    for (int a = 0; a < n; ++a)
      for (int b = 0; b < n; ++b) {
        if (a == b) continue;
        const double dx = G.pos[a][0] - G.pos[b][0], dy = G.pos[a][1] - G.pos[b][1], dz = G.pos[a][2] - G.pos[b][2];
        if (!(std::sqrt(dx * dx + dy * dy + dz * dz) >= lim))
          return fail(CMTJ_ERR_INVALID, "Points are too close for stable dipole calculation");
      }
  }
  for (size_t q = 0; q < nj; ++q) {
    const cmtj_junction& j = js[q];
    const cmtj_group& G = groups[q / (size_t)n];
    const int e = (int)(q % (size_t)n);
    // equal step counts keep the plan's instance order the caller's
    if (j.solver != j0.solver || j.n_steps != j0.n_steps || j.write_every != j0.write_every ||
        j.n_entries != j0.n_entries || j.dt != j0.dt)
      return fail(CMTJ_ERR_INVALID, "group elements must share solver, step count, time step and log cadence");
    if (j.n_layers < 1 || j.n_layers > 4) return fail(CMTJ_ERR_INVALID, "group elements have 1..4 layers");
    if (G.top[e] < 0 || G.top[e] >= j.n_layers) return fail(CMTJ_ERR_INVALID, "top layer index out of range");
    const size_t row = (size_t)j.layer_offset + (size_t)G.top[e];
    if (row >= n_layer_rows) return fail(CMTJ_ERR_INVALID, "layer_offset out of range");
    // setLayerReservedInteractionField replaces the top layer's driver
    // (reservoir.hpp:144-145); the kernel adds the group field itself
    for (int s = 0; s < 3; ++s) rows[row].drv[CMTJ_SLOT_HRES + s] = cmtj_driver{};
    const size_t nch = (size_t)n_channels(j.n_layers, j.mr_mode, j.log_mask);
    need = std::max<size_t>(need, j.log_offset + nch * j.n_entries);
  }
  if (need > 0 && (!log_host || log_cap < need)) return fail(CMTJ_ERR_INVALID, "log buffer too small");
  std::lock_guard<std::mutex> lk(g_plain.mu);
  if (const int rb = plain_bind()) return rb;
  cmtj_plan P;
  P.owns = false;
  int r = build_plan(&P, js, nj, rows.data(), rows.size(), c, nullptr, 0, CMTJ_FP64, CMTJ_LOG_F64, CMTJ_LANES_ONE,
                     plain_arena, &g_plain, g_plain.stream);
  if (r) return r;
  if (P.groups.size() != 1)
    return fail(CMTJ_ERR_INVALID,
                "the junctions of a group must share one structure (layers, modes, driver kinds)");
  // the smallest group kernel that covers the structure; four waves per
  // group while the groups leave SIMDs idle (the sum over the sources then
  // takes a quarter of the step), one wave each beyond
  bool any_field = false;
  for (size_t q = 0; q < n_groups; ++q) any_field |= groups[q].interaction != CMTJ_INTERACTION_NULL;
  const uint32_t gw = (any_field && n > 8 && n_groups * 4 <= 1024) ? 4u : 1u;
  const KTable kt = gw == 4 ? kernels_group4() : kernels_group();
  const KEntry* ke = nullptr;
  for (size_t q = 0; q < kt.n; ++q) {
    const KEntry& e = kt.e[q];
    if (e.solver == j0.solver && e.lm >= P.groups[0].a.nl &&
        covers(e, P.groups[0].a.flags, 0, P.groups[0].a.ndyn) && (!ke || rank(e) < rank(*ke)))
      ke = &e;
  }
  if (!ke) return fail(CMTJ_ERR_INVALID, "no group kernel for this solver and layer count");
  const size_t o_mag = align_up(need * sizeof(double));
  const size_t o_st = align_up(o_mag + sizeof(double) * 3 * CMTJ_MAX_LAYERS * nj);
  const size_t o_grp = align_up(o_st + sizeof(cmtj_status) * nj);
  bool any_tensor = false;
  for (size_t q = 0; q < n_groups; ++q) any_tensor |= groups[q].interaction == CMTJ_INTERACTION_TENSOR;
  const size_t n_pt = any_tensor ? n_groups * (size_t)n * 12 * (size_t)n : 0;
  const size_t o_pt = align_up(o_grp + sizeof(cmtj_group) * n_groups);
  const size_t out_bytes = o_pt + sizeof(double) * n_pt;
  char* out = (char*)grow(g_plain.out, g_plain.out_bytes, out_bytes);
  if (!out) return CMTJ_ERR_HIP;
  HIP_TRY(hipMemcpyAsync(out + o_grp, groups, sizeof(cmtj_group) * n_groups, hipMemcpyHostToDevice,
                         g_plain.stream));
  if (n_pt)
    HIP_TRY(hipMemcpyAsync(out + o_pt, pair_terms, sizeof(double) * n_pt, hipMemcpyHostToDevice, g_plain.stream));
  Group& G = P.groups[0];
  G.k = ke;
  G.blocks = (uint32_t)n_groups;  // group g on workgroup g, element e on lane e of its first wave
  G.threads = 64 * gw;
  G.a.group = (const cmtj_group*)(out + o_grp);
  G.a.group_n = (uint32_t)n_groups;
  G.a.pair = n_pt ? (const double*)(out + o_pt) : nullptr;
  r = launch_plan(&P, need ? out : nullptr, (double*)(out + o_mag), (cmtj_status*)(out + o_st), g_plain.stream);
  if (r) return r;
  if (need) HIP_TRY(hipMemcpyAsync(log_host, out, need * sizeof(double), hipMemcpyDeviceToHost, g_plain.stream));
  HIP_TRY(hipMemcpyAsync(mag_out_host, out + o_mag, sizeof(double) * 3 * CMTJ_MAX_LAYERS * nj,
                         hipMemcpyDeviceToHost, g_plain.stream));
  HIP_TRY(hipMemcpyAsync(st_host, out + o_st, sizeof(cmtj_status) * nj, hipMemcpyDeviceToHost, g_plain.stream));
  HIP_TRY(hipStreamSynchronize(g_plain.stream));
  return 0;
}

cmtj_plan* cmtj_plan_create(const cmtj_junction* js, size_t n, const cmtj_layer* layers, size_t n_layer_rows,
                            const cmtj_constants* c, const double* tables, size_t n_table, int32_t precision,
                            int32_t log_dtype, int32_t lanes_per_instance) {
  if (!js || !layers || !c) {
    fail(CMTJ_ERR_INVALID, "null pointer");
    return nullptr;
  }
  if (lanes_per_instance < CMTJ_LANES_AUTO || lanes_per_instance > CMTJ_LANES_QUEUE) {
    fail(CMTJ_ERR_INVALID,
         "lanes_per_instance must be 0 (auto), 1 (one lane per instance), 2 (one lane per layer), 3 (one lane "
         "per component) or 4 (a work queue)");
    return nullptr;
  }
  cmtj_plan* P = new cmtj_plan();
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
    fail(CMTJ_ERR_HIP, "hipStreamCreate failed");
    delete P;
    return nullptr;
  }
  const int r = build_plan(P, js, n, layers, n_layer_rows, c, tables, n_table, precision, log_dtype,
                           lanes_per_instance, malloc_arena, nullptr, s);
  (void)hipStreamDestroy(s);
  if (r) {
    if (P->arena) (void)hipFree(P->arena);
    delete P;
    return nullptr;
  }
  return P;
}

int cmtj_group_key(const cmtj_junction* j, const cmtj_layer* layers, uint64_t* key) {
  if (!j || !layers || !key) return fail(CMTJ_ERR_INVALID, "null pointer");
  cmtj_junction jj = *j;
  jj.layer_offset = 0;
  const int r = validate(jj, 0, layers, (size_t)std::max(jj.n_layers, 0), CMTJ_FP64, SIZE_MAX);
  if (r) return r;
  std::vector<Rec> recs;
  junction_records(jj, layers, recs);
  std::vector<int32_t> sig;
  signature(jj, layers, recs, sig);
  uint64_t h = 1469598103934665603ull;  // FNV-1a
  for (int32_t v : sig)
    for (int b = 0; b < 4; ++b) {
      h ^= (uint64_t)(((uint32_t)v >> (8 * b)) & 0xff);
      h *= 1099511628211ull;
    }
  *key = h;
  return 0;
}

int cmtj_plan_arena_size(const cmtj_junction* js, size_t n, const cmtj_layer* layers, size_t n_layer_rows,
                         size_t n_table, int32_t precision, int32_t log_dtype, int32_t lanes_per_instance,
                         size_t* bytes) {
  if (!js || !layers || !bytes) return fail(CMTJ_ERR_INVALID, "null pointer");
  cmtj_plan P;
  cmtj_constants c = kDefaults;
  return build_plan(&P, js, n, layers, n_layer_rows, &c, nullptr, n_table, precision, log_dtype,
                    lanes_per_instance, nullptr, nullptr, nullptr, bytes);
}

cmtj_plan* cmtj_plan_create_in(const cmtj_junction* js, size_t n, const cmtj_layer* layers, size_t n_layer_rows,
                               const cmtj_constants* c, const double* tables, size_t n_table, int32_t precision,
                               int32_t log_dtype, int32_t lanes_per_instance, void* d_arena, size_t arena_bytes,
                               void* stream, int32_t arena_is_host) {
  if (!js || !layers || !c || !d_arena) {
    fail(CMTJ_ERR_INVALID, "null pointer");
    return nullptr;
  }
  if (lanes_per_instance < CMTJ_LANES_AUTO || lanes_per_instance > CMTJ_LANES_QUEUE) {
    fail(CMTJ_ERR_INVALID,
         "lanes_per_instance must be 0 (auto), 1 (one lane per instance), 2 (one lane per layer), 3 (one lane "
         "per component) or 4 (a work queue)");
    return nullptr;
  }
  void* dev = d_arena;
  if (arena_is_host && hipHostGetDevicePointer(&dev, d_arena, 0) != hipSuccess) {
    fail(CMTJ_ERR_INVALID, "arena is not host memory mapped for the device (use cmtj_host_alloc)");
    return nullptr;
  }
  cmtj_plan* P = new cmtj_plan();
  P->owns = false;
  FixedArena fa{dev, arena_bytes};
  const int r = build_plan(P, js, n, layers, n_layer_rows, c, tables, n_table, precision, log_dtype,
                           lanes_per_instance, fixed_arena, &fa, (hipStream_t)stream, nullptr,
                           arena_is_host ? (char*)d_arena : nullptr);
  if (r) {
    delete P;
    return nullptr;
  }
  return P;
}

int cmtj_host_alloc(size_t bytes, void** host, void** device) {
  if (!host || !device || bytes == 0) return fail(CMTJ_ERR_INVALID, "bad host allocation request");
  *host = *device = nullptr;
  if (hipHostMalloc(host, bytes, hipHostMallocMapped | hipHostMallocPortable | hipHostMallocCoherent) !=
      hipSuccess)
    return fail(CMTJ_ERR_HIP, "hipHostMalloc failed for " + std::to_string(bytes) + " bytes");
  if (hipHostGetDevicePointer(device, *host, 0) != hipSuccess) {
    (void)hipHostFree(*host);
    *host = nullptr;
    return fail(CMTJ_ERR_HIP, "hipHostGetDevicePointer failed");
  }
  return 0;
}

int cmtj_host_free(void* host) {
  if (host && hipHostFree(host) != hipSuccess) return fail(CMTJ_ERR_HIP, "hipHostFree failed");
  return 0;
}

int cmtj_plan_execute(cmtj_plan* p, void* d_log, size_t log_capacity, double* d_mag_out,
                      cmtj_status* d_status, void* stream) {
  if (!p) return fail(CMTJ_ERR_INVALID, "null plan");
  if (d_log && log_capacity < p->log_extent)
    return fail(CMTJ_ERR_INVALID, "log buffer smaller than the plan's log extent (" +
                                      std::to_string(p->log_extent) + " elements)");
  return launch_plan(p, d_log, d_mag_out, d_status, (hipStream_t)stream);
}

int cmtj_plan_execute_slice(cmtj_plan* p, void* d_log, size_t log_capacity, double* d_mag_out,
                            cmtj_status* d_status, cmtj_state* d_state, uint32_t max_steps,
                            void* const* streams, int32_t n_streams) {
  if (!p) return fail(CMTJ_ERR_INVALID, "null plan");
  if (!d_state) return fail(CMTJ_ERR_INVALID, "a slice needs the instance state");
  if (max_steps == 0) return fail(CMTJ_ERR_INVALID, "max_steps must be positive");
  for (const Group& G : p->groups)
    if (G.k->solver == CMTJ_DORMAND_PRINCE)
      return fail(CMTJ_ERR_INVALID, "Dormand-Prince runs are not sliced (their step count is adaptive)");
  if (n_streams < 0 || (n_streams > 0 && !streams)) return fail(CMTJ_ERR_INVALID, "bad stream list");
  if (d_log && log_capacity < p->log_extent)
    return fail(CMTJ_ERR_INVALID, "log buffer smaller than the plan's log extent (" +
                                      std::to_string(p->log_extent) + " elements)");
  return launch_plan(p, d_log, d_mag_out, d_status, nullptr, d_state, max_steps, streams, n_streams);
}

int cmtj_vsd_mix(const double* d_R, size_t ld, size_t n, size_t len, const double* d_freq, double dt,
                 const double* b, const double* a, int32_t n_taps, double* d_out, void* stream) {
  if (n == 0) return 0;
  if (!d_R || !d_freq || !d_out || !b || !a) return fail(CMTJ_ERR_INVALID, "null pointer");
  if (n_taps < 1 || n_taps > kMaxTaps)
    return fail(CMTJ_ERR_INVALID, "n_taps must be 1.." + std::to_string(kMaxTaps));
  if (a[0] == 0.0) return fail(CMTJ_ERR_INVALID, "a[0] must be nonzero");
  if (len > 0xffffffffull || n > 0xffffffffull) return fail(CMTJ_ERR_INVALID, "series too long");
  if (ld < len) return fail(CMTJ_ERR_INVALID, "row stride shorter than a series");
  if (!(dt > 0.0)) return fail(CMTJ_ERR_INVALID, "dt must be positive");
  IirTaps f;
  std::memset(&f, 0, sizeof(f));
  f.n = n_taps;
  for (int i = 0; i < n_taps; ++i) {  // lfilter normalises by a[0]
    f.b[i] = b[i] / a[0];
    f.a[i] = a[i] / a[0];
  }
  const uint32_t blocks = (uint32_t)((n + 63) / 64);
  hipLaunchKernelGGL(k_vsd_mix, dim3(blocks), dim3(64), 0, (hipStream_t)stream, d_R, (uint64_t)ld, (uint32_t)n,
                     (uint32_t)len, d_freq, dt, f, d_out);
  HIP_TRY(hipGetLastError());
  return 0;
}

int cmtj_plan_query(const cmtj_plan* p, cmtj_plan_info* out) {
  if (!p || !out) return fail(CMTJ_ERR_INVALID, "null pointer");
  *out = p->info;
  return 0;
}

int cmtj_plan_graph_create(cmtj_plan* p, void* d_log, size_t log_capacity, double* d_mag_out,
                           cmtj_status* d_status) {
  if (!p) return fail(CMTJ_ERR_INVALID, "null plan");
  if (d_log && log_capacity < p->log_extent)
    return fail(CMTJ_ERR_INVALID, "log buffer smaller than the plan's log extent (" +
                                      std::to_string(p->log_extent) + " elements)");
  drop_graph(p);
  // the kernel arguments stay with the plan: the graph's nodes were built from them
  p->graph_args.resize(p->groups.size());
  HIP_TRY(hipGraphCreate(&p->graph, 0));
  for (size_t q = 0; q < p->groups.size(); ++q) {
    const Group& G = p->groups[q];
    GroupArgs& a = p->graph_args[q];
    a = G.a;
    a.log = d_log;
    a.mag_out = d_mag_out;
    a.status = d_status;
    a.state = nullptr;
    a.slice = 0;
    void* args[] = {&a};
    uint32_t* ctr = G.queue;
    uint32_t one = 1;
    void* cargs[] = {&ctr, &one};
    void* qargs[] = {&a, &ctr};
    hipKernelNodeParams kp;
    std::memset(&kp, 0, sizeof(kp));
    kp.func = const_cast<void*>(G.queue ? dp_queue_kernel() : G.k->fn);
    kp.gridDim = dim3(G.blocks);
    kp.blockDim = dim3(G.threads);
    kp.sharedMemBytes = 0;
    kp.kernelParams = G.queue ? qargs : args;
    kp.extra = nullptr;
    hipGraphNode_t clear = nullptr, node;
    hipError_t e = hipSuccess;
    if (G.queue) {  // the counter is cleared by a node the kernel node depends on
      hipKernelNodeParams cp;
      std::memset(&cp, 0, sizeof(cp));
      cp.func = const_cast<void*>(clear_u32_kernel());
      cp.gridDim = dim3(1);
      cp.blockDim = dim3(64);
      cp.kernelParams = cargs;
      e = hipGraphAddKernelNode(&clear, p->graph, nullptr, 0, &cp);
    }
    if (e == hipSuccess) e = hipGraphAddKernelNode(&node, p->graph, G.queue ? &clear : nullptr, G.queue ? 1 : 0, &kp);
    if (e != hipSuccess) {
      drop_graph(p);
      return fail(CMTJ_ERR_HIP, std::string("hipGraphAddKernelNode: ") + hipGetErrorString(e));
    }
  }
  const hipError_t e = hipGraphInstantiate(&p->graph_exec, p->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    drop_graph(p);
    return fail(CMTJ_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  return 0;
}

int cmtj_plan_graph_launch(cmtj_plan* p, void* stream) {
  if (!p) return fail(CMTJ_ERR_INVALID, "null plan");
  if (!p->graph_exec) return fail(CMTJ_ERR_INVALID, "the plan has no graph (cmtj_plan_graph_create)");
  HIP_TRY(hipGraphLaunch(p->graph_exec, (hipStream_t)stream));
  return 0;
}

int cmtj_plan_destroy(cmtj_plan* p) {
  if (!p) return 0;
  drop_graph(p);
  if (p->owns && p->arena) (void)hipFree(p->arena);
  delete p;
  return 0;
}

}  // extern "C"
