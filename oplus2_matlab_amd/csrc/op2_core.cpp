// OP2 runtime core: registries, declarations, argument checks, constants,
// timing and the C ABI (see include/op2_hip.h for the entry points and the
// reference interfaces they replace).
#include <sys/time.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>

#include "op2_internal.h"

extern "C" {
int OP_diags = 0, OP_part_size = 0, OP_block_size = 256, OP_cache_line_size = 128;
int OP_set_index = 0, OP_set_max = 0, OP_map_index = 0, OP_map_max = 0,
    OP_dat_index = 0, OP_dat_max = 0, OP_plan_index = 0, OP_plan_max = 0,
    OP_kern_max = 0;
op_set *OP_set_list = nullptr;
op_map *OP_map_list = nullptr;
op_dat *OP_dat_list = nullptr;
op_kernel *OP_kernels = nullptr;
op_plan *OP_plans = nullptr;
}

namespace op2 {

Runtime &rt() {
  static Runtime r;
  return r;
}

void fail(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  Runtime &r = rt();
  r.last_error = buf;
  if (r.error_mode == OP2_ERRORS_RETURN) throw Error(buf);
  printf("%s\n", buf);
  fflush(stdout);
  exit(-1);
}

double Const::as_double(int i) const {
  if (i < 0 || i >= dim) fail("constant index %d out of range (dim %d)", i, dim);
  const char *p = bytes.data() + (size_t)i * size;
  if (type == "double") return *(const double *)p;
  if (type == "float") return *(const float *)p;
  if (type == "int") return *(const int *)p;
  fail("constant of type %s cannot be used by the device loops", type.c_str());
}

void ensure_device() {
  Runtime &r = rt();
  if (r.device_ready) return;
  int n = 0;
  OP2_HIP_CHECK(hipGetDeviceCount(&n));
  if (n == 0) fail("op2: no HIP device available");
  OP2_HIP_CHECK(hipSetDevice(r.device));
  OP2_HIP_CHECK(hipStreamCreateWithFlags(&r.stream, hipStreamNonBlocking));
  OP2_HIP_CHECK(hipHostMalloc((void **)&r.h_result, 64 * sizeof(double), hipHostMallocDefault));
  r.device_ready = true;
}

void *dev_alloc(size_t bytes) {
  ensure_device();
  if (rt().capturing)
    fail("op2: device allocation during graph capture (run every loop once before "
         "op2_graph_begin so that its plan and maps are on the device)");
  void *p = nullptr;
  OP2_HIP_CHECK(hipMalloc(&p, bytes ? bytes : 256));
  return p;
}

void dev_free(void *p) {
  if (p) (void)hipFree(p);
}

void h2d(void *dst, const void *src, size_t bytes) {
  if (!bytes) return;
  if (rt().capturing)
    fail("op2: host-to-device copy during graph capture (run every loop once before "
         "op2_graph_begin so that its plan and maps are on the device)");
  OP2_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, rt().stream));
  OP2_HIP_CHECK(hipStreamSynchronize(rt().stream));
}

void d2h(void *dst, const void *src, size_t bytes) {
  if (!bytes) return;
  if (rt().capturing) fail("op2: device-to-host copy during graph capture");
  OP2_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, rt().stream));
  OP2_HIP_CHECK(hipStreamSynchronize(rt().stream));
}

void ensure_map_device(op_map m) {
  if (m->map_d) return;
  size_t bytes = sizeof(int) * (size_t)m->dim * m->from->size;
  m->map_d = (int *)dev_alloc(bytes);
  h2d(m->map_d, map_host(m), bytes);
}

void *reduct_scratch(size_t bytes) {
  Runtime &r = rt();
  if (bytes > r.reduct_bytes) {
    if (r.capturing) fail("op2: reduction scratch grown during graph capture");
    if (r.d_reduct) {
      OP2_HIP_CHECK(hipStreamSynchronize(r.stream));
      // a captured graph may still use the old buffer: keep it until op_exit
      if (r.graphs.empty())
        dev_free(r.d_reduct);
      else
        r.retired.push_back(r.d_reduct);
    }
    r.reduct_bytes = ((bytes + 255) / 256) * 256 * 2;
    r.d_reduct = dev_alloc(r.reduct_bytes);
    OP2_HIP_CHECK(hipMemsetAsync(r.d_reduct, 0, r.reduct_bytes, r.stream));
  }
  return r.d_reduct;
}

void ensure_reduction_log() {
  Runtime &r = rt();
  if (r.d_log) return;
  if (r.capturing) fail("op2: reduction log created during graph capture");
  r.d_log = (double *)dev_alloc(sizeof(double) * kLogCap);
  r.d_log_count = (unsigned *)dev_alloc(256);
  OP2_HIP_CHECK(hipMemsetAsync(r.d_log_count, 0, 256, r.stream));
  OP2_HIP_CHECK(hipStreamSynchronize(r.stream));
  r.log_read = 0;
}

const Const &get_const(const char *name) {
  auto it = rt().consts.find(name);
  if (it == rt().consts.end())
    fail("op2: constant \"%s\" has not been declared (op_decl_const)", name);
  return it->second;
}

LoopConfig config_for(const std::string &loop) {
  Runtime &r = rt();
  LoopConfig c = r.default_config;
  auto it = r.loop_configs.find(loop);
  if (it != r.loop_configs.end()) {
    if (it->second.strategy) c.strategy = it->second.strategy;
    if (it->second.part_size) c.part_size = it->second.part_size;
    if (it->second.block_size) c.block_size = it->second.block_size;
  }
  return c;
}

// ---- timing -----------------------------------------------------------------

LoopTiming &timing_for(const std::string &name, double bytes_per_call, bool indirect) {
  Runtime &r = rt();
  for (auto &t : r.timings)
    if (t.name == name) {
      t.bytes_per_call = bytes_per_call;
      t.indirect = indirect;
      return t;
    }
  r.timings.emplace_back();
  LoopTiming &t = r.timings.back();
  t.name = name;
  t.bytes_per_call = bytes_per_call;
  t.indirect = indirect;
  // kernel numbers in first-call order, as op2.m numbers a program's loops
  t.kernel = (int)r.timings.size() - 1;
  op_timing_realloc(t.kernel);
  OP_kernels[t.kernel].name = t.name.c_str();
  return t;
}

static hipEvent_t take_event() {
  Runtime &r = rt();
  if (r.event_pool.empty()) {
    hipEvent_t e;
    OP2_HIP_CHECK(hipEventCreate(&e));
    return e;
  }
  hipEvent_t e = r.event_pool.back();
  r.event_pool.pop_back();
  return e;
}

int mark_begin(int kind) {
  Runtime &r = rt();
  if (!r.capturing || r.graph_kernel_timing <= 0) return -1;
  if (++r.capture_stages % r.graph_kernel_timing) return -1;
  Runtime::KernelMark m{nullptr, nullptr, kind};
  for (hipEvent_t *e : {&m.start, &m.stop}) {
    if (r.mark_pool.empty()) {
      OP2_HIP_CHECK(hipEventCreate(e));
    } else {
      *e = r.mark_pool.back();
      r.mark_pool.pop_back();
    }
  }
  r.capture_marks.push_back(m);
  OP2_HIP_CHECK(hipEventRecordWithFlags(m.start, r.stream, hipEventRecordExternal));
  return (int)r.capture_marks.size() - 1;
}

void mark_end(int mark) {
  Runtime &r = rt();
  if (mark < 0) return;
  OP2_HIP_CHECK(hipEventRecordWithFlags(r.capture_marks[mark].stop, r.stream,
                                        hipEventRecordExternal));
}

static void return_marks(std::vector<Runtime::KernelMark> &marks) {
  for (auto &m : marks) {
    rt().mark_pool.push_back(m.start);
    rt().mark_pool.push_back(m.stop);
  }
  marks.clear();
}

void timing_begin(LoopTiming &t, hipEvent_t *start) {
  (void)t;
  rt().last_plan = nullptr;
  *start = nullptr;
  if (!rt().timing_enabled || rt().capturing) return;
  *start = take_event();
  OP2_HIP_CHECK(hipEventRecord(*start, rt().stream));
}

void timing_end(LoopTiming &t, hipEvent_t start) {
  t.count++;
  // the reference's record (e.g. airfoil/res_calc_kernel.cpp:241-248):
  // transfer = bytes the loop uses (here each touched dat once per
  // direction and each map once); transfer2, for loops through maps only,
  // adds the bytes of the plan arrays the kernels read
  op_kernel &k = OP_kernels[t.kernel];
  k.count = t.count;
  k.transfer += (float)t.bytes_per_call;
  if (t.indirect) {
    const Plan *p = rt().last_plan;
    k.transfer2 += (float)(t.bytes_per_call + (p ? (double)p->device_bytes : 0.0));
  }
  if (!start) return;
  hipEvent_t stop = take_event();
  OP2_HIP_CHECK(hipEventRecord(stop, rt().stream));
  t.pending.push_back({start, stop});
  if (t.pending.size() >= 256) timing_drain();
}

void timing_drain() {
  Runtime &r = rt();
  if (!r.device_ready || r.capturing) return;
  OP2_HIP_CHECK(hipStreamSynchronize(r.stream));
  for (auto &t : r.timings) {
    for (auto &p : t.pending) {
      float ms = 0;
      OP2_HIP_CHECK(hipEventElapsedTime(&ms, p.start, p.stop));
      t.seconds += ms * 1e-3;
      r.event_pool.push_back(p.start);
      r.event_pool.push_back(p.stop);
    }
    t.pending.clear();
    if (t.kernel >= 0 && t.kernel < OP_kern_max) OP_kernels[t.kernel].time = (float)t.seconds;
  }
}

static void clear_all() {
  Runtime &r = rt();
  fusion_drop();
  if (r.capturing) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(r.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    r.capturing = false;
  }
  if (r.device_ready) (void)hipStreamSynchronize(r.stream);
  for (auto &g : r.graphs)
    if (g) (void)hipGraphExecDestroy(g);
  r.graphs.clear();
  return_marks(r.capture_marks);
  for (auto &gm : r.graph_marks) return_marks(gm.second);
  r.graph_marks.clear();
  for (void *p : r.retired) dev_free(p);
  r.retired.clear();
  r.deferred = false;
  for (auto &p : r.plans) p->release();
  r.plans.clear();
  release_ref_plans();
  release_exec_caches();
  for (int i = 0; i < OP_dat_index; i++) {
    dev_free(OP_dat_list[i]->data_d);
    free(OP_dat_list[i]);
  }
  for (int i = 0; i < OP_map_index; i++) {
    dev_free(OP_map_list[i]->map_d);
    free(OP_map_list[i]);
  }
  for (int i = 0; i < OP_set_index; i++) free(OP_set_list[i]);
  free(OP_dat_list);
  free(OP_map_list);
  free(OP_set_list);
  free(OP_kernels);
  OP_dat_list = nullptr;
  OP_map_list = nullptr;
  OP_set_list = nullptr;
  OP_kernels = nullptr;
  OP_set_index = OP_set_max = OP_map_index = OP_map_max = 0;
  OP_dat_index = OP_dat_max = OP_plan_index = OP_plan_max = OP_kern_max = 0;
  for (auto &t : r.timings)
    for (auto &p : t.pending) {
      r.event_pool.push_back(p.start);
      r.event_pool.push_back(p.stop);
    }
  r.timings.clear();
  r.consts.clear();
  r.set_perm.clear();
  r.auto_renumber = false;
  r.set_segments.clear();
  r.map_internal.clear();
  checkpoint_reset();
  for (auto &o : r.online)
    for (auto &ev : o.second.events)
      for (auto &e : ev) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
      }
  r.online.clear();
}

}  // namespace op2

using namespace op2;

#define API_BEGIN try {
#define API_END(ret)          \
  }                           \
  catch (const op2::Error &) { \
    return ret;               \
  }

template <class T>
static void grow(T **list, int *max, int index) {
  if (index == *max) {
    *max += 10;
    T *n = (T *)realloc(*list, sizeof(T) * (*max));
    if (!n) fail(" op2 error -- error reallocating memory");
    *list = n;
  }
}

extern "C" {

void op_init_core(int argc, char **argv, int diags) {
  API_BEGIN
  OP_diags = diags;
#ifdef OP_BLOCK_SIZE
  OP_block_size = OP_BLOCK_SIZE;
#endif
#ifdef OP_PART_SIZE
  OP_part_size = OP_PART_SIZE;
#endif
  for (int n = 1; n < argc; n++) {
    if (!argv || !argv[n]) continue;
    if (strncmp(argv[n], "OP_BLOCK_SIZE=", 14) == 0) {
      OP_block_size = atoi(argv[n] + 14);
      printf("\n OP_block_size = %d \n", OP_block_size);
    }
    if (strncmp(argv[n], "OP_PART_SIZE=", 13) == 0) {
      OP_part_size = atoi(argv[n] + 13);
      printf("\n OP_part_size  = %d \n", OP_part_size);
    }
    if (strncmp(argv[n], "OP_CACHE_LINE_SIZE=", 19) == 0) {
      OP_cache_line_size = atoi(argv[n] + 19);
      printf("\n OP_cache_line_size  = %d \n", OP_cache_line_size);
    }
  }
  // the kernels are instantiated for 2, 4 and 8 waves of 64 threads
  if (OP_block_size != 128 && OP_block_size != 256 && OP_block_size != 512)
    fail("op_init error -- OP_block_size %d is not supported (128, 256 or 512 threads)",
         OP_block_size);
  // renumber without a call in the program: OP2_RENUMBER=1 (auto, on) or
  // =0 (off, no) in the environment, else op2_set_auto_renumber's setting
  const char *rn = getenv("OP2_RENUMBER");
  if (rn && (!strcmp(rn, "1") || !strcmp(rn, "auto") || !strcmp(rn, "on")))
    rt().auto_renumber = true;
  else if (rn && (!strcmp(rn, "0") || !strcmp(rn, "off") || !strcmp(rn, "no")))
    rt().auto_renumber = false;
  else
    rt().auto_renumber = rt().auto_renumber_default;
  rt().initialised = true;
  API_END()
}

void op2_init(int argc, char **argv, int diags) {
  API_BEGIN
  op_init_core(argc, argv, diags);
  ensure_device();
  if (OP_diags > 0) {
    hipDeviceProp_t prop;
    OP2_HIP_CHECK(hipGetDeviceProperties(&prop, rt().device));
    printf("\n Using HIP device: %s (%s)\n", prop.name, prop.gcnArchName);
  }
  API_END()
}

op_set op_decl_set(int size, char const *name) {
  API_BEGIN
  if (size <= 0) fail(" op_decl_set error -- negative/zero size for set: %s", name);
  grow(&OP_set_list, &OP_set_max, OP_set_index);
  op_set s = (op_set)malloc(sizeof(op_set_core));
  s->index = OP_set_index;
  s->size = size;
  s->name = name;
  OP_set_list[OP_set_index++] = s;
  return s;
  API_END(nullptr)
}

op_map op_decl_map(op_set from, op_set to, int dim, int *imap, char const *name) {
  API_BEGIN
  if (!from) fail(" op_decl_map error -- invalid 'from' set for map %s", name);
  if (!to) fail("op_decl_map error -- invalid 'to' set for map %s", name);
  if (dim <= 0) fail("op_decl_map error -- negative/zero dimension for map %s", name);
  if (!imap) fail("op_decl_map error -- NULL data for map %s", name);
  for (int d = 0; d < dim; d++)
    for (int n = 0; n < from->size; n++) {
      int v = imap[d + (size_t)n * dim];
      if (v < 0 || v >= to->size)
        fail("op_decl_map error -- invalid data for map %s\nelement = %d, dimension = %d, map = %d",
             name, n, d, v);
    }
  grow(&OP_map_list, &OP_map_max, OP_map_index);
  op_map m = (op_map)malloc(sizeof(op_map_core));
  m->index = OP_map_index;
  m->from = from;
  m->to = to;
  m->dim = dim;
  m->map = imap;
  m->name = name;
  m->map_d = nullptr;  // copied to the device by the first loop that reads it
  m->map_i = nullptr;
  OP_map_list[OP_map_index++] = m;
  if (!rt().set_perm.empty()) renumber_map(m);
  return m;
  API_END(nullptr)
}

op_dat op_decl_dat_core(op_set set, int dim, char const *type, int size, char *data,
                        char const *name) {
  API_BEGIN
  if (!set) fail("op_decl_dat error -- invalid set for data: %s", name);
  if (dim <= 0) fail("op_decl_dat error -- negative/zero dimension for data: %s", name);
  grow(&OP_dat_list, &OP_dat_max, OP_dat_index);
  op_dat d = (op_dat)malloc(sizeof(op_dat_core));
  d->index = OP_dat_index;
  d->set = set;
  d->dim = dim;
  d->data = data;
  d->data_d = nullptr;
  d->name = name;
  d->type = type;
  d->size = dim * size;
  OP_dat_list[OP_dat_index++] = d;
  return d;
  API_END(nullptr)
}

op_dat op2_decl_dat_char(op_set set, int dim, char const *type, int size, char *data,
                         char const *name) {
  API_BEGIN
  op_dat d = op_decl_dat_core(set, dim, type, size, data, name);
  if (!d) return nullptr;
  size_t bytes = (size_t)d->size * set->size;
  d->data_d = (char *)dev_alloc(bytes);
  if (data && set_new_of_old(set)) {
    std::vector<char> internal(bytes);
    permute_to_internal(d, data, internal.data());
    h2d(d->data_d, internal.data(), bytes);
  } else if (data)
    h2d(d->data_d, data, bytes);
  else
    OP2_HIP_CHECK(hipMemsetAsync(d->data_d, 0, bytes, rt().stream));
  return d;
  API_END(nullptr)
}

void op2_decl_const_char(int dim, char const *type, int size, char *data, char const *name) {
  API_BEGIN
  if (!name || !*name) fail("op_decl_const error -- constant has no name");
  if (dim <= 0 || size <= 0 || !data) fail("op_decl_const error -- invalid constant %s", name);
  Const c;
  c.type = type ? type : "";
  c.dim = dim;
  c.size = size;
  c.bytes.assign(data, data + (size_t)dim * size);
  rt().consts[name] = c;
  auto it = rt().const_targets.find(name);
  if (it != rt().const_targets.end()) {
    ensure_device();
    it->second(c.bytes.data(), c.bytes.size(), (void *)rt().stream);
  }
  API_END()
}

void op2_fetch_data(op_dat dat) {
  API_BEGIN
  if (!dat || !dat->data_d || !dat->data) fail("op_fetch_data error -- dat has no device copy");
  const size_t bytes = (size_t)dat->size * dat->set->size;
  if (set_new_of_old(dat->set)) {
    std::vector<char> internal(bytes);
    d2h(internal.data(), dat->data_d, bytes);
    permute_to_user(dat, internal.data(), dat->data);
  } else {
    d2h(dat->data, dat->data_d, bytes);
  }
  API_END()
}

static void arg_err(const char *msg, int m, const char *name) {
  fail("error: arg %d in kernel \"%s\"\n%s ", m, name, msg);
}

void op_arg_check(op_set set, int m, op_arg arg, int *ninds, char const *name) {
  API_BEGIN
  if (arg.argtype == OP_ARG_DAT) {
    if (!set) arg_err("invalid set", m, name);
    if (!arg.dat) arg_err("invalid dataset", m, name);
    if (!arg.map && arg.dat->set != set) arg_err("dataset set does not match loop set", m, name);
    if (arg.map && (arg.map->from != set || arg.map->to != arg.dat->set))
      arg_err("mapping error", m, name);
    if ((!arg.map && arg.idx != -1) || (arg.map && (arg.idx < 0 || arg.idx >= arg.map->dim)))
      arg_err("invalid index", m, name);
    if (arg.dat->dim != arg.dim) arg_err("dataset dim does not match declared dim", m, name);
    if (strcmp(arg.dat->type, arg.type)) arg_err("dataset type does not match declared type", m, name);
    if (arg.idx >= 0) (*ninds)++;
  }
  if (arg.argtype == OP_ARG_GBL) {
    if (!strcmp(arg.type, "error")) arg_err("datatype does not match declared type", m, name);
    if (arg.dim <= 0) arg_err("dimension should be strictly positive", m, name);
    if (!arg.data) arg_err("NULL pointer for global data", m, name);
  }
  API_END()
}

op_arg op_arg_dat(op_dat dat, int idx, op_map map, int dim, char const *type, op_access acc) {
  op_arg a;
  memset(&a, 0, sizeof a);
  a.argtype = OP_ARG_DAT;
  a.dat = dat;
  a.map = map;
  a.dim = dim;
  a.idx = idx;
  if (dat) {
    a.size = dat->size;
    a.data = dat->data;
    a.data_d = dat->data_d;
  }
  a.type = type;
  a.acc = acc;
  return a;
}

op_arg op_arg_gbl_core(char *data, int dim, char const *type, op_access acc) {
  op_arg a;
  memset(&a, 0, sizeof a);
  a.argtype = OP_ARG_GBL;
  a.dat = nullptr;
  a.map = nullptr;
  a.dim = dim;
  a.idx = -1;
  a.size = 0;
  a.data = data;
  a.type = type;
  a.acc = acc;
  return a;
}

void op_diagnostic_output(void) {
  if (OP_diags <= 1) return;
  printf("\n  OP diagnostic output\n");
  printf(  "  --------------------\n");
  printf("\n       set       size\n");
  printf(  "  -------------------\n");
  for (int n = 0; n < OP_set_index; n++)
    printf("%10s %10d\n", OP_set_list[n]->name, OP_set_list[n]->size);
  printf("\n       map        dim       from         to\n");
  printf(  "  -----------------------------------------\n");
  for (int n = 0; n < OP_map_index; n++)
    printf("%10s %10d %10s %10s\n", OP_map_list[n]->name, OP_map_list[n]->dim,
           OP_map_list[n]->from->name, OP_map_list[n]->to->name);
  printf("\n       dat        dim        set\n");
  printf(  "  ------------------------------\n");
  for (int n = 0; n < OP_dat_index; n++)
    printf("%10s %10d %10s\n", OP_dat_list[n]->name, OP_dat_list[n]->dim,
           OP_dat_list[n]->set->name);
  printf("\n");
}

// The reference's table (common/op_lib_core.c:381-403): per loop its call
// count, seconds (device time of its kernels, HIP events) and GB/s of
// transfer and, for loops through maps, of transfer2.
void op_timing_output(void) {
  API_BEGIN
  timing_drain();
  if (OP_kern_max <= 0) return;
  printf("\n  count     time     GB/s     GB/s   kernel name ");
  printf("\n ----------------------------------------------- \n");
  for (int n = 0; n < OP_kern_max; n++) {
    const op_kernel &k = OP_kernels[n];
    if (k.count <= 0) continue;
    const float t = k.time > 0 ? k.time : 1e-30f;
    if (k.transfer2 == 0.0f)
      printf(" %6d  %8.4f %8.4f            %s \n", k.count, k.time, k.transfer / (1e9f * t),
             k.name);
    else
      printf(" %6d  %8.4f %8.4f %8.4f   %s \n", k.count, k.time, k.transfer / (1e9f * t),
             k.transfer2 / (1e9f * t), k.name);
  }
  API_END()
}

void op_timers(double *cpu, double *et) {
  (void)cpu;
  struct timeval t;
  gettimeofday(&t, (struct timezone *)0);
  *et = t.tv_sec + t.tv_usec * 1.0e-6;
}

void op_timing_realloc(int kernel) {
  API_BEGIN
  if (kernel >= OP_kern_max) {
    int nmax = kernel + 10;
    op_kernel *k = (op_kernel *)realloc(OP_kernels, nmax * sizeof(op_kernel));
    if (!k) fail(" op_timing_realloc error ");
    OP_kernels = k;
    for (int n = OP_kern_max; n < nmax; n++) {
      OP_kernels[n].count = 0;
      OP_kernels[n].time = 0.0f;
      OP_kernels[n].transfer = 0.0f;
      OP_kernels[n].transfer2 = 0.0f;
      OP_kernels[n].name = "unused";
    }
    OP_kern_max = nmax;
  }
  API_END()
}

void op_exit_core(void) {
  API_BEGIN
  clear_all();
  API_END()
}

void op2_exit(void) {
  API_BEGIN
  clear_all();
  API_END()
}

int op2_set_error_mode(int mode) {
  int old = rt().error_mode;
  rt().error_mode = mode;
  return old;
}

char const *op2_last_error(void) { return rt().last_error.c_str(); }

void op2_clear_error(void) { rt().last_error.clear(); }

int op2_set_loop_config(char const *name, int strategy, int part_size, int block_size) {
  API_BEGIN
  if (strategy < 0 || strategy > OP2_STRATEGY_OWNER_GATHER) fail("op2: unknown strategy %d", strategy);
  if (part_size < 0 || part_size > 65535) fail("op2: part size %d out of range", part_size);
  if (block_size != 0 && block_size != 128 && block_size != 256 && block_size != 512)
    fail("op2: block size %d is not supported (0 for the default, or 128, 256 or 512 "
         "threads)", block_size);
  LoopConfig c{strategy, part_size, block_size};
  if (name)
    rt().loop_configs[name] = c;
  else {
    rt().default_config = c;
    rt().loop_configs.clear();
  }
  return 0;
  API_END(-1)
}

int op2_get_loop_config(char const *name, int *strategy, int *part_size, int *block_size) {
  API_BEGIN
  if (!name) fail("op2_get_loop_config: invalid arguments");
  const LoopConfig c = config_for(name);
  if (strategy) *strategy = c.strategy;
  if (part_size) *part_size = c.part_size;
  if (block_size) *block_size = c.block_size ? c.block_size : OP_block_size;
  return 0;
  API_END(-1)
}

int op2_loop_stats(char const *name, int *count, double *seconds, double *bytes_per_call) {
  API_BEGIN
  timing_drain();
  for (auto &t : rt().timings)
    if (t.name == name) {
      if (count) *count = t.count;
      if (seconds) *seconds = t.seconds;
      if (bytes_per_call) *bytes_per_call = t.bytes_per_call;
      return 0;
    }
  return -1;
  API_END(-1)
}

void op2_timing_reset(void) {
  API_BEGIN
  timing_drain();
  for (auto &t : rt().timings) {
    t.count = 0;
    t.seconds = 0;
  }
  for (int n = 0; n < OP_kern_max; n++) {
    OP_kernels[n].count = 0;
    OP_kernels[n].time = OP_kernels[n].transfer = OP_kernels[n].transfer2 = 0.0f;
  }
  API_END()
}

void *op2_stream(void) {
  API_BEGIN
  ensure_device();
  return (void *)rt().stream;
  API_END(nullptr)
}

int op2_sync(void) {
  API_BEGIN
  ensure_device();
  OP2_HIP_CHECK(hipStreamSynchronize(rt().stream));
  return 0;
  API_END(-1)
}

void *op2_dat_device_ptr(op_dat dat) {
  if (!dat) return nullptr;
  expose_dat(dat);  // other code may now write it: never taken as known zero
  return (void *)dat->data_d;
}

int op2_set_deferred_reductions(int on) {
  API_BEGIN
  Runtime &r = rt();
  if (r.capturing) fail("op2: reduction mode cannot change during graph capture");
  int old = r.deferred ? 1 : 0;
  if (on) {
    ensure_device();
    ensure_reduction_log();
  }
  r.deferred = on != 0;
  return old;
  API_END(-1)
}

int op2_reduction_log(double *out, int max) {
  API_BEGIN
  Runtime &r = rt();
  if (r.capturing) fail("op2: reduction log read during graph capture");
  if (!r.d_log) return 0;
  unsigned cnt32 = 0;
  OP2_HIP_CHECK(hipStreamSynchronize(r.stream));
  OP2_HIP_CHECK(hipMemcpy(&cnt32, r.d_log_count, sizeof cnt32, hipMemcpyDeviceToHost));
  // the device counter is 32-bit; the host cursor carries the high bits
  unsigned long long cnt = (r.log_read & ~0xffffffffull) | cnt32;
  if (cnt < r.log_read) cnt += 1ull << 32;
  unsigned long long pending = cnt - r.log_read;
  if (pending > kLogCap)
    fail("op2: reduction log overflow (%llu results pending, capacity %u); read it more often",
         pending, kLogCap);
  int n = (int)std::min<unsigned long long>(pending, (unsigned long long)(max < 0 ? 0 : max));
  std::vector<double> all(kLogCap);
  if (n > 0) {
    OP2_HIP_CHECK(hipMemcpy(all.data(), r.d_log, sizeof(double) * kLogCap, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) out[i] = all[(r.log_read + i) & (kLogCap - 1)];
  }
  r.log_read += n;
  return n;
  API_END(-1)
}

int op2_graph_begin(void) {
  API_BEGIN
  Runtime &r = rt();
  ensure_device();
  if (r.capturing) fail("op2: graph capture already active");
  if (!r.deferred)
    fail("op2: graph capture needs deferred reductions (op2_set_deferred_reductions(1))");
  for (auto &o : r.online)
    if (!o.second.done)
      fail("op2: online autotuning of loop \"%s\" is unfinished; run the loop until it ends "
           "before graph capture", o.first.c_str());
  timing_drain();
  reduct_scratch(256 + sizeof(double) * 2048 * 4);
  forget_zero_dats();  // a replay may follow any eager work
  prepare_capture();
  return_marks(r.capture_marks);
  r.capture_stages = 0;
  OP2_HIP_CHECK(hipStreamSynchronize(r.stream));
  OP2_HIP_CHECK(hipStreamBeginCapture(r.stream, hipStreamCaptureModeThreadLocal));
  r.capturing = true;
  return 0;
  API_END(-1)
}

int op2_graph_end(void) {
  API_BEGIN
  Runtime &r = rt();
  if (!r.capturing) fail("op2: no graph capture active");
  hipGraph_t g = nullptr;
  try {
    fusion_flush();
  } catch (const Error &) {
    r.capturing = false;
    (void)hipStreamEndCapture(r.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    return_marks(r.capture_marks);
    throw;
  }
  r.capturing = false;
  const hipError_t ended = hipStreamEndCapture(r.stream, &g);
  if (ended != hipSuccess) return_marks(r.capture_marks);
  OP2_HIP_CHECK(ended);
  hipGraphExec_t ex = nullptr;
  OP2_HIP_CHECK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
  OP2_HIP_CHECK(hipGraphDestroy(g));
  // the graph's launch resources go to the device now, not at its first
  // launch
  OP2_HIP_CHECK(hipGraphUpload(ex, r.stream));
  r.graphs.push_back(ex);
  const int id = (int)r.graphs.size() - 1;
  r.graph_marks[id].swap(r.capture_marks);
  return_marks(r.capture_marks);
  return id;
  API_END(-1)
}

int op2_set_graph_fusion(int level) {
  Runtime &r = rt();
  const int prev = r.graph_fusion;
  if (r.capturing) return -1;
  r.graph_fusion = level <= 0 ? 0 : level == 1 ? 1 : 2;
  return prev;
}

int op2_set_stage_config(int part_size, int block_size) {
  Runtime &r = rt();
  if (r.capturing) return -1;
  if (part_size < 0) {
    r.stage_part = -1;
    r.stage_block = 0;
    return 0;
  }
  if (part_size == 0) {
    r.stage_part = r.stage_block = 0;
    return 0;
  }
  const int block = block_size ? block_size : 256;
  if ((block != 128 && block != 256 && block != 512) || part_size > 2 * block) return -1;
  r.stage_part = part_size;
  r.stage_block = block;
  return 0;
}

int op2_graph_launch(int id, int count) {
  API_BEGIN
  Runtime &r = rt();
  if (id < 0 || id >= (int)r.graphs.size() || !r.graphs[id]) fail("op2: no graph %d", id);
  for (int i = 0; i < count; i++) OP2_HIP_CHECK(hipGraphLaunch(r.graphs[id], r.stream));
  forget_zero_dats();  // the replays wrote dats
  return 0;
  API_END(-1)
}

int op2_graph_destroy(int id) {
  API_BEGIN
  Runtime &r = rt();
  if (id < 0 || id >= (int)r.graphs.size() || !r.graphs[id]) fail("op2: no graph %d", id);
  OP2_HIP_CHECK(hipStreamSynchronize(r.stream));
  OP2_HIP_CHECK(hipGraphExecDestroy(r.graphs[id]));
  r.graphs[id] = nullptr;
  auto gm = r.graph_marks.find(id);
  if (gm != r.graph_marks.end()) {
    return_marks(gm->second);
    r.graph_marks.erase(gm);
  }
  return 0;
  API_END(-1)
}

int op2_set_graph_kernel_timing(int every) {
  Runtime &r = rt();
  const int prev = r.graph_kernel_timing;
  if (r.capturing) return -1;
  r.graph_kernel_timing = std::max(every, 0);
  return prev;
}

int op2_graph_kernel_times(int id, int max, int *kinds, float *ms) {
  API_BEGIN
  Runtime &r = rt();
  if (id < 0 || id >= (int)r.graphs.size() || !r.graphs[id]) fail("op2: no graph %d", id);
  if (max < 0 || (max > 0 && (!kinds || !ms)))
    fail("op2_graph_kernel_times: invalid arguments");
  if (r.capturing) fail("op2_graph_kernel_times: not during graph capture");
  auto gm = r.graph_marks.find(id);
  if (gm == r.graph_marks.end()) return 0;
  OP2_HIP_CHECK(hipStreamSynchronize(r.stream));
  const auto &marks = gm->second;
  for (int i = 0; i < (int)marks.size() && i < max; i++) {
    kinds[i] = marks[i].kind;
    OP2_HIP_CHECK(hipEventElapsedTime(&ms[i], marks[i].start, marks[i].stop));
  }
  return (int)marks.size();
  API_END(-1)
}

int op2_autotune_loop(char const *name, op_set set, int nargs, op_arg *args, int reps,
                      int *best_strategy, int *best_part, double *best_us) {
  API_BEGIN
  if (!name || !set || nargs <= 0 || !args || reps <= 0) fail("op2_autotune_loop: invalid arguments");
  return autotune(name, set, nargs, args, reps, best_strategy, best_part, best_us);
  API_END(-1)
}

int op2_autotune_online(char const *name, int reps) {
  API_BEGIN
  if (!name || reps <= 0) fail("op2_autotune_online: invalid arguments");
  if (rt().capturing) fail("op2_autotune_online: not during graph capture");
  refuse_with_live_graphs("op2_autotune_online");
  OnlineTune t;
  t.reps = reps;
  rt().online[name] = t;
  return 0;
  API_END(-1)
}

int op2_autotune_online_result(char const *name, int *best_strategy, int *best_part,
                               double *best_us) {
  API_BEGIN
  if (!name) fail("op2_autotune_online_result: invalid arguments");
  auto it = rt().online.find(name);
  if (it == rt().online.end()) return -1;
  if (!it->second.done) return 0;
  if (best_strategy) *best_strategy = it->second.best.strategy;
  if (best_part) *best_part = it->second.best.part_size;
  if (best_us) *best_us = it->second.best_us;
  return 1;
  API_END(-1)
}

int op2_register_loop(char const *name, op2_loop_fn fn) {
  API_BEGIN
  if (!name) fail("op2_register_loop: invalid arguments");
  if (fn)
    rt().loops[name] = fn;
  else
    rt().loops.erase(name);
  return 0;
  API_END(-1)
}

int op2_register_const(char const *name, op2_const_copy_fn fn) {
  API_BEGIN
  if (!name || !fn) fail("op2_register_const: invalid arguments");
  Runtime &r = rt();
  r.const_targets[name] = fn;
  auto it = r.consts.find(name);
  if (it != r.consts.end()) {
    ensure_device();
    fn(it->second.bytes.data(), it->second.bytes.size(), (void *)r.stream);
  }
  return 0;
  API_END(-1)
}

int op2_generic_plan_get(char const *name, op_set set, int nargs, op_arg *args, int part_size,
                         op2_generic_plan *out) {
  API_BEGIN
  if (!name || !set || nargs <= 0 || !args || !out) fail("op2_generic_plan_get: invalid arguments");
  // block size: the argument, else the loop's configured part size, else 256
  int part = part_size > 0 ? part_size : config_for(name).part_size;
  if (part <= 0) part = 256;
  PlanRequest req{name, set, nargs, args, kStrategyGeneric, part};
  Plan *p = get_plan(req);
  p->upload();
  out->nblocks = p->nblocks;
  out->ncolors = p->ncolors;
  out->blk_off = p->d_blk_off;
  out->perm = p->d_perm;
  out->col_off = p->d_col_off;
  out->col_base = p->d_col_base;
  out->blk_ncol = p->d_blk_ncol;
  out->blkmap = p->d_blkmap;
  out->colblk_off = p->colblk_off.data();
  return 0;
  API_END(-1)
}

int op2_generic_owner_plan_get(char const *name, op_set set, int nargs, op_arg *args,
                               int part_size, op2_generic_owner_plan *out, int *arg_info) {
  API_BEGIN
  if (!name || !set || nargs <= 0 || !args || !out || !arg_info)
    fail("op2_generic_owner_plan_get: invalid arguments");
  // the form: indirect writes are OP_INC, all through one map; at most 4
  // distinct (map, index) pairs
  op_map inc_map = nullptr;
  std::vector<std::pair<op_map, int>> pairs;
  for (int a = 0; a < nargs; a++) {
    const op_arg &g = args[a];
    if (g.argtype != OP_ARG_DAT || !g.map) continue;
    if (g.acc != OP_READ && g.acc != OP_INC) return 1;
    if (g.acc == OP_INC) {
      if (inc_map && inc_map != g.map) return 1;
      inc_map = g.map;
    }
    if (std::find(pairs.begin(), pairs.end(), std::make_pair(g.map, g.idx)) == pairs.end())
      pairs.emplace_back(g.map, g.idx);
  }
  if (!inc_map || pairs.size() > 4) return 1;
  int part = part_size > 0 ? part_size : config_for(name).part_size;
  if (part <= 0) part = 128;
  // the gathering form of the owner plan: per position the target of each
  // (map, index) pair, so that a kernel reads no map
  PlanRequest req{name, set, nargs, args, OP2_STRATEGY_OWNER_GATHER, part};
  Plan *p = get_plan(req);
  p->upload();
  if (!p->d_perm) {  // (an owner plan of a loop without direct dats has no perm on the device)
    p->d_perm = (int *)dev_alloc(sizeof(int) * std::max<size_t>(p->perm.size(), 1));
    h2d(p->d_perm, p->perm.data(), sizeof(int) * p->perm.size());
    p->device_bytes += sizeof(int) * p->perm.size();
  }
  out->nblocks = p->nblocks;
  out->max_positions = p->max_positions;
  out->max_owned = p->max_owned;
  out->slot_stride = p->slot_stride;
  out->blk_off = p->d_blk_off;
  out->perm = p->d_perm;
  out->ids = p->d_ids;
  out->owned_off = p->d_owned_off;
  out->owned = p->owned_identity ? nullptr : p->d_owned;
  out->csr_off = p->d_csr_off;
  out->csr = p->d_csr;
  out->ell = p->d_ell;
  for (int a = 0; a < nargs; a++) {
    arg_info[a] = arg_info[nargs + a] = -1;
    if (args[a].argtype != OP_ARG_DAT || !args[a].map) continue;
    const int s = p->arg_slot[a];
    arg_info[nargs + a] = s;
    if (args[a].acc != OP_INC) continue;
    for (size_t k = 0; k < p->inc_slot_ids.size(); k++)
      if (p->inc_slot_ids[k] == s) arg_info[a] = (int)k;
  }
  return 0;
  API_END(-1)
}

int op2_report_error(char const *msg) {
  API_BEGIN
  fail("%s", msg ? msg : "op2: error");
  API_END(-1)
}

int const *op2_map_device_ptr(op_map map) {
  API_BEGIN
  if (!map) fail("op2_map_device_ptr: invalid map");
  ensure_map_device(map);
  return map->map_d;
  API_END(nullptr)
}

int op2_reduction_log_device(double **log, unsigned **count, unsigned *capacity) {
  Runtime &r = rt();
  if (!r.deferred || !r.d_log) return 0;
  if (log) *log = r.d_log;
  if (count) *count = r.d_log_count;
  if (capacity) *capacity = kLogCap;
  return 1;
}

void *op2_device_scratch(size_t bytes) {
  API_BEGIN
  return reduct_scratch(bytes);
  API_END(nullptr)
}

int op2_set_auto_renumber(int on) {
  const int old = rt().auto_renumber_default ? 1 : 0;
  rt().auto_renumber_default = on != 0;
  return old;
}

int op2_renumber(op_map base) {
  API_BEGIN
  renumber(base);
  return 0;
  API_END(-1)
}

int op2_renumber_export(op_set set, int *new_of_old) {
  API_BEGIN
  if (!set) fail("op2_renumber_export: invalid set");
  const std::vector<int> *p = set_new_of_old(set);
  if (!p) return 0;
  if (new_of_old) std::copy(p->begin(), p->end(), new_of_old);
  return 1;
  API_END(-1)
}

int op2_par_loop(char const *name, op_set set, int nargs, op_arg *args) {
  API_BEGIN
  return par_loop(name, set, nargs, args);
  API_END(-1)
}

}  // extern "C"
