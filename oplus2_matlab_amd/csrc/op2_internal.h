// Internal state of the MI355X OP2 runtime (not part of the public API).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "op2_hip.h"

namespace op2 {

// Raised by fail() when the error mode is OP2_ERRORS_RETURN; caught at the
// C-ABI boundary.
struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Reports an error: prints and exits (reference behaviour) or throws Error.
[[noreturn]] void fail(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

#define OP2_HIP_CHECK(call)                                                     \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess)                                                       \
      ::op2::fail("%s(%d): HIP error %s in %s", __FILE__, __LINE__,             \
                  hipGetErrorString(e_), #call);                                \
  } while (0)

// A named constant from op_decl_const / op_decl_const2.
struct Const {
  std::string type;
  int dim = 0, size = 0;
  std::vector<char> bytes;
  double as_double(int i) const;
};

struct LoopConfig {
  int strategy = OP2_STRATEGY_AUTO;
  int part_size = 0;
  int block_size = 0;
};

// One staged map of a plan: the distinct target elements of each block.
struct PlanGroup {
  op_map map = nullptr;
  std::vector<int> off;   // [nblocks+1] into list
  std::vector<int> list;  // global target ids, ascending within a block
  int max_count = 0;      // largest per-block count
  int *d_off = nullptr, *d_list = nullptr;
};

// 16-bit position records of gather tiles: two windows of kIdSpan ids per
// map per tile; a record is window << 15 | offset in the window.
constexpr int kIdSpan = 1 << 15;

// Execution plan of one indirect loop on one set (this runtime's design:
// groups by map rather than by dat, packed 16-bit local indices per
// (map, idx) slot, elements sorted by thread colour inside each block).
struct Plan {
  // key
  std::string loop;
  op_set set = nullptr;
  int strategy = 0, part_size = 0;
  std::vector<op_map> maps;
  std::vector<int> idxs;
  std::vector<int> accs;
  std::vector<op_dat> dats;
  std::vector<op_arg> args;  // the request's arguments (a plan of another part size)

  // layout
  int nblocks = 0, max_block = 0;
  std::vector<int> blk_off;  // [nblocks+1] positions
  std::vector<int> perm;     // position -> element (empty: identity)
  bool need_perm = false;    // perm is uploaded for loops with direct dats
  std::vector<PlanGroup> groups;
  // slots: distinct (map, idx) of indirect args, in first-use order
  std::vector<std::pair<int, int>> slots;  // (group, idx)
  std::vector<int> arg_slot;               // per arg, -1 if direct / global
  int slot_stride = 0;                     // 1, 2 or 4 (padded)
  std::vector<uint16_t> loc;               // [set_size * slot_stride]
  // thread colours (positions of colour k of block b are
  // [col_off[col_base[b]+k], col_off[col_base[b]+k+1]))
  std::vector<int> col_off, col_base, blk_ncol;
  std::vector<int> thrcol;  // per position
  // block colours
  int ncolors = 0;
  std::vector<int> blkcol;      // per block
  std::vector<int> blkmap;      // blocks grouped by colour
  std::vector<int> colblk_off;  // [ncolors+1]
  long long verify_errors = -1;
  double mean_thread_colours = 0;
  // owner-compute layout (OP2_STRATEGY_OWNER): the targets of the OP_INC
  // slots are partitioned into tiles, one per block; a block holds every
  // element that increments one of its targets (positions may repeat an
  // element in two blocks) and, per owned target, the list of
  // (position << 2 | inc slot) entries that increment it.
  int owner_group = -1;
  std::vector<int> inc_slot_ids;   // slot id of each INC slot, in INC order
  std::vector<int> owned_off;      // [nblocks+1]
  std::vector<int> owned;          // global target ids
  std::vector<int> csr_off;        // [owned.size()+1]
  std::vector<uint16_t> csr;       // (position in block << 2) | inc slot index
  // the same lists padded to 4 entries per owned target (0xffff = none), when
  // no target has more than 4: one 8-byte record per owned target
  std::vector<uint16_t> ell;
  int max_positions = 0, max_owned = 0, max_incs = 0;
  bool owned_identity = false;  // owned[i] == i: tiles own consecutive targets
  bool airfoil_res = false;     // run by the res_calc kernels (graph stage fusion)

  // device copies
  int *d_blk_off = nullptr, *d_perm = nullptr, *d_col_off = nullptr,
      *d_col_base = nullptr, *d_blk_ncol = nullptr, *d_blkmap = nullptr;
  uint16_t *d_loc = nullptr;
  std::vector<int> ids;            // [positions * slot_stride] global targets (gather form)
  int *d_ids = nullptr;
  // the same targets as 16-bit records (four slots on two maps: slots 0, 1
  // through the first, 2, 3 through the second), read by the FP32 eager
  // res_calc; per tile {first window, second - first} of each map, the
  // target being first + offset (+ second - first for window 1); a tile
  // whose first base is negative is read through ids.  Empty when no tile
  // fits.
  std::vector<uint16_t> ids16;  // [positions * 4]
  std::vector<int> id_win;      // [nblocks][4]
  int short_tiles = 0;
  uint16_t *d_ids16 = nullptr;
  int *d_id_win = nullptr;
  int *d_owned_off = nullptr, *d_owned = nullptr, *d_csr_off = nullptr;
  uint16_t *d_csr = nullptr, *d_ell = nullptr;

  int count = 0;  // calls that used this plan

  bool uploaded = false;
  size_t device_bytes = 0;  // bytes of the device arrays (read once per call)
  void upload();  // no-op once uploaded
  // frees the device arrays, or moves their addresses to *keep when given
  void release(std::vector<void *> *keep = nullptr);
};

// An execution plan in the reference's own layout (common/op_lib_core.h:
// 122-150), built for callers of op_plan_core / op_plan_get: blocks of
// part_size consecutive elements; per indirect dataset (the caller's inds)
// the block's distinct targets in ascending order (ind_maps, ind_offs,
// ind_sizes) and every indirect argument's 16-bit local index per element
// (loc_maps); thread colours over the OP_INC arguments inside each block;
// block colours; shared-memory bytes and the transfer byte model.  The
// runtime's kernels do not use it (they run on Plan).
struct RefPlan {
  std::string name;
  op_plan view{};
  std::vector<op_map> maps;
  std::vector<op_dat> dats;
  std::vector<int> idxs, inds;
  std::vector<op_access> accs;
  std::vector<int> nthrcol, thrcol, offset, ind_offs, ind_sizes, nindirect, nelems, ncolblk,
      blkmap;
  std::vector<std::vector<int>> ind_maps;
  std::vector<int *> ind_map_ptrs;
  std::vector<std::vector<short>> loc_maps;
  std::vector<short *> loc_map_ptrs;
};

// Per-loop timing record (the op_kernel table of op_lib_core.c:381-427);
// OP_kernels[kernel] mirrors it (kernel numbers in first-call order).
struct LoopTiming {
  std::string name;
  int kernel = -1;
  bool indirect = false;  // has an argument through a map
  int count = 0;
  double seconds = 0;
  double bytes_per_call = 0;
  struct Pending {
    hipEvent_t start, stop;
  };
  std::vector<Pending> pending;
};

// Online autotuning state of one loop (op2_autotune_online).
struct OnlineTune {
  int reps = 3;
  bool started = false, done = false;
  std::vector<LoopConfig> cands;
  std::vector<double> us;  // mean device time per candidate (1e300: failed)
  std::vector<std::vector<std::pair<hipEvent_t, hipEvent_t>>> events;
  int call = 0;
  LoopConfig best;
  double best_us = 0;
};

struct Runtime {
  bool initialised = false;
  bool device_ready = false;
  int device = 0;
  hipStream_t stream = nullptr;
  int error_mode = OP2_ERRORS_EXIT;
  std::string last_error;
  std::map<std::string, Const> consts;
  std::map<std::string, LoopConfig> loop_configs;
  LoopConfig default_config;
  std::vector<std::unique_ptr<Plan>> plans;
  std::vector<std::unique_ptr<RefPlan>> ref_plans;  // op_plan_core views (OP_plans)
  std::deque<LoopTiming> timings;  // deque: OP_kernels[k].name points into it
  const Plan *last_plan = nullptr;  // the last plan get_plan returned
  std::vector<hipEvent_t> event_pool;
  // reduction scratch: per-block partials + ticket counter + result slots
  void *d_reduct = nullptr;
  size_t reduct_bytes = 0;
  double *h_result = nullptr;  // pinned
  bool timing_enabled = true;
  // deferred reductions: results appended to a device log instead of being
  // written back to the caller's variable
  bool deferred = false;
  double *d_log = nullptr;
  unsigned *d_log_count = nullptr;
  unsigned long long log_read = 0;
  // graph capture of loop sequences on the runtime stream; with
  // graph_fusion, adjacent loops that can share a pass are fused there
  bool capturing = false;
  int graph_fusion = 2;  // 0 off, 1 loop fusion, 2 also one kernel per Runge-Kutta stage
  // tiles of the fused stage kernels (op2_set_stage_config): cells per tile
  // and threads; 0 = 256 and 256, -1 = res_calc's own configuration
  int stage_part = 0, stage_block = 0;
  std::vector<hipGraphExec_t> graphs;
  std::vector<void *> retired;  // replaced scratch buffers that graphs may use
  std::map<std::string, OnlineTune> online;
  // kernel timing inside graphs (op2_set_graph_kernel_timing): every n-th
  // stage kernel of a capture (n = graph_kernel_timing, 0 = none; counted by
  // capture_stages) is bracketed by two event-record nodes; kind = the loops
  // it runs and its res traffic (op2_hip.h)
  struct KernelMark {
    hipEvent_t start, stop;
    int kind;
  };
  int graph_kernel_timing = 0;
  int capture_stages = 0;
  std::vector<hipEvent_t> mark_pool;             // recorded by no live graph
  std::vector<KernelMark> capture_marks;         // of the capture in progress
  std::map<int, std::vector<KernelMark>> graph_marks;  // per live graph id
  // renumbering (op2_renumber): new number of every element of a renumbered
  // set, the starts of its independently ordered pieces, and the
  // renumbered host copies of the maps (op_map::map_i points into them)
  std::map<op_set, std::vector<int>> set_perm;
  bool auto_renumber = false;  // renumber at the first indirect increment loop (pending)
  bool auto_renumber_default = false;  // op2_set_auto_renumber
  std::map<op_set, std::vector<int>> set_segments;
  std::map<op_map, std::vector<int>> map_internal;
  // generic loop path: loops and device constants registered by the
  // translation units that compile the user kernels
  std::map<std::string, op2_loop_fn> loops;
  std::map<std::string, op2_const_copy_fn> const_targets;
};

constexpr unsigned kLogCap = 1u << 16;  // power of two

Runtime &rt();
void ensure_device();
void *dev_alloc(size_t bytes);
void dev_free(void *p);
void release_exec_caches();
// buffers the stage kernels of a capture need (op2_graph_begin; a capture
// cannot allocate), and forgetting a plan's when the plan is dropped
void prepare_capture();
void forget_plan_caches(const Plan *p);
// graph kernel timing: around a kernel launched during capture, mark_begin
// records the start node (returns -1 when timing is off or not capturing)
// and mark_end the stop node
int mark_begin(int kind);
void mark_end(int mark);
// loop fusion during graph capture: launch, or forget, a held loop
void fusion_flush();
void fusion_drop();
// dats known to hold zeros: forgotten when work the runtime does not see
// may have written them (a graph replay, a restart, a renumbering); a dat
// whose device pointer the program has taken is never taken as zero
void forget_zero_dats();
void expose_dat(op_dat d);
void h2d(void *dst, const void *src, size_t bytes);
void d2h(void *dst, const void *src, size_t bytes);
void ensure_map_device(op_map m);

// checkpoint / fast-forward restart (op2_checkpoint.cpp): skip a loop call
// before a restart's checkpoint (handing back its recorded reductions);
// count / record a loop call that ran; forget everything (op_exit)
bool checkpoint_skip(const char *name, int nargs, const op_arg *args);
void checkpoint_after(const char *name, int nargs, const op_arg *args);
void checkpoint_reset();

// renumbering (op2_renumber.cpp)
inline const int *map_host(op_map m) { return m->map_i ? m->map_i : m->map; }
const std::vector<int> *set_new_of_old(op_set s);
void renumber(op_map base);
// with OP2_RENUMBER set: renumbers once, on the map of the first loop with
// an indirect OP_INC argument (for programs that do not call op2_renumber)
void auto_renumber(op_set set, int nargs, const op_arg *args);
void renumber_map(op_map m);  // (re)builds m's internal copy for the current numbering
void permute_to_internal(op_dat d, const char *user, char *internal);
void permute_to_user(op_dat d, const char *internal, char *user);

const Const &get_const(const char *name);
LoopConfig config_for(const std::string &loop);

// plan of the generic loop path (header op_hip_loop.h): coloured on every
// indirect write, no staging lists; not selectable by op2_set_loop_config
constexpr int kStrategyGeneric = OP2_PLAN_GENERIC;

// plan construction (op2_plan.cpp)
struct PlanRequest {
  std::string loop;
  op_set set;
  int nargs;
  const op_arg *args;
  int strategy;
  int part_size;
};
Plan *get_plan(const PlanRequest &req);
Plan *find_plan(const PlanRequest &req);  // null if not built (builds nothing)
long long verify_plan(const Plan &p);
// frees the op_plan_core views and the OP_plans array (op_exit)
void release_ref_plans();
// releases a plan's device arrays and forgets it (while a live graph may
// hold their addresses the arrays are kept until op_exit instead)
void drop_plan(Plan *p);
// true while an instantiated graph exists (op2_graph_end, not destroyed)
bool live_graphs();
// fails with "<what>: ..." while live_graphs(): the call would free or
// move device arrays that a captured graph reads
void refuse_with_live_graphs(const char *what);

// timing (op2_core.cpp)
LoopTiming &timing_for(const std::string &name, double bytes_per_call, bool indirect);
void timing_begin(LoopTiming &t, hipEvent_t *start);
void timing_end(LoopTiming &t, hipEvent_t start);
void timing_drain();

// scratch for reductions: at least `bytes` of device memory, 256-aligned
void *reduct_scratch(size_t bytes);
void ensure_reduction_log();

// loop executor (op2_exec.hip)
int par_loop(const char *name, op_set set, int nargs, op_arg *args);
int autotune(const char *name, op_set set, int nargs, op_arg *args, int reps, int *best_strategy,
             int *best_part, double *best_us);

}  // namespace op2
