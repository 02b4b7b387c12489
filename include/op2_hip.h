/*
 * C ABI of the MI355X OP2 runtime, libop2hip.so.
 *
 * Every entry point is extern "C" with plain pointers, sizes and the OP2
 * handle types of op_lib_core.h.  Each one names the reference interface it
 * replaces (file:line in the reference tree).  The C++ API of op_lib.h and
 * op_seq.h is a header-only layer over these symbols; a foreign-function
 * binding (ctypes, cgo, JNI) binds these directly.
 *
 * Errors: by default an error prints a message and calls exit(), as the
 * reference does (common/op_lib_core.c:260-264, common/op_lib.cu:60-80).
 * op2_set_error_mode(OP2_ERRORS_RETURN) makes entry points record the
 * message (op2_last_error) and return an error value instead.
 */
#ifndef OP2_HIP_H
#define OP2_HIP_H

#include "op_lib_core.h"

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * Reference C-linkage entry points (common/op_lib.h:101-153)
 */

/* common/op_lib_core.c:68-92: diagnostics level; argv OP_BLOCK_SIZE=,
 * OP_PART_SIZE=, OP_CACHE_LINE_SIZE=.  OP_block_size (threads per
 * workgroup of every loop without a block size of its own) must be 128, 256
 * (the default) or 512. */
void op_init_core(int argc, char **argv, int diags);

/* common/op_lib_core.c:94-119 (size must be > 0) */
op_set op_decl_set(int size, char const *name);

/* common/op_lib_core.c:121-169 (entries checked in range); the map is also
 * copied to the device */
op_map op_decl_map(op_set from, op_set to, int dim, int *imap, char const *name);

/* common/op_lib_core.c:171-207: registers the dat without a device copy */
op_dat op_decl_dat_core(op_set set, int dim, char const *type, int size,
                        char *data, char const *name);

/* common/op_lib_core.c:266-305 */
void op_arg_check(op_set set, int m, op_arg arg, int *ninds, char const *name);

/* common/op_lib_core.c:307-326 */
op_arg op_arg_dat(op_dat dat, int idx, op_map map, int dim, char const *type,
                  op_access acc);

/* common/op_lib_core.c:329-345 */
op_arg op_arg_gbl_core(char *data, int dim, char const *type, op_access acc);

/* common/op_lib_core.c:351-379 */
void op_diagnostic_output(void);

/* common/op_lib_core.c:381-403: count, seconds, GB/s per loop; seconds are
 * device time of the loop's kernels (HIP events) */
void op_timing_output(void);

/* common/op_lib_core.c:627-1089: the execution plan of a loop in the
 * reference's layout, built on the host and kept until op_exit: blocks of
 * part_size consecutive elements (0: the reference's size for 48 KB of
 * staged data per 64 elements); per indirect dataset inds[m] (0..ninds-1,
 * -1 for direct and global arguments) the ascending distinct targets of
 * each block (ind_maps, ind_offs, ind_sizes, nindirect) and each indirect
 * argument's 16-bit local index per element (loc_maps); thread colours over
 * the OP_INC arguments (thrcol, nthrcol); block colours (ncolors, ncolblk,
 * blkmap); nshared; the transfer / transfer2 byte model.  The runtime's
 * own kernels run on plans of their own layout (op2_plan_summary_get).
 * Returns NULL (with the error recorded) on invalid input. */
op_plan *op_plan_core(char const *name, op_set set, int part_size, int nargs,
                      op_arg *args, int ninds, int *inds);

/* common/op_lib_core.c:440-445 */
void op_timers(double *cpu, double *et);

/* common/op_lib_core.c:405-427 */
void op_timing_realloc(int kernel);

/* common/op_lib_core.c:209-254 */
void op_exit_core(void);

/* reference globals (common/op_lib.h:259-271) */
extern int OP_part_size, OP_block_size, OP_cache_line_size;
extern int OP_set_index, OP_set_max, OP_map_index, OP_map_max, OP_dat_index,
    OP_dat_max, OP_plan_index, OP_plan_max, OP_kern_max;
extern op_set *OP_set_list;
extern op_map *OP_map_list;
extern op_dat *OP_dat_list;
extern op_kernel *OP_kernels;
/* common/op_lib.h:85: the plans op_plan_core has built (OP_plan_index of
 * them), in the reference's layout */
extern op_plan *OP_plans;

/* ---------------------------------------------------------------------------
 * C-ABI forms of the reference's C++-linkage entry points
 */

/* op_init, common/op_lib.cu:128-143: op_init_core + HIP device 0, runtime
 * stream and events */
void op2_init(int argc, char **argv, int diags);

/* op_decl_dat_char, common/op_lib.cu:145-153: registers the dat and copies
 * it to device memory; the host array stays the user's */
op_dat op2_decl_dat_char(op_set set, int dim, char const *type, int size,
                         char *data, char const *name);

/* op_decl_const_char, airfoil/airfoil_kernels.cu:50-53: stores a named
 * constant (dim elements of size bytes) for the device loops */
void op2_decl_const_char(int dim, char const *type, int size, char *data,
                         char const *name);

/* op_fetch_data, common/op_lib.cu:116-121: device -> user's host array */
void op2_fetch_data(op_dat dat);

/* op_plan_get, common/op_lib.cu:155-184 */
op_plan *op2_plan_get(char const *name, op_set set, int part_size, int nargs,
                      op_arg *args, int ninds, int *inds);

/* op_exit, common/op_lib.cu:186-209: frees plans, device data, registries */
void op2_exit(void);

/* op_par_loop (common/op_seq.h:26-1340) and the generated per-loop stubs
 * op_par_loop_<name> (airfoil/airfoil_op.cpp:64-99, jac1/jac_op.cpp): runs
 * the loop `name` over `set` on the device.  The loop is identified by its
 * name and argument signature; loops this runtime has no device kernel for
 * are an error.  OP_GBL arguments are read on entry and, for OP_INC, OP_MIN
 * and OP_MAX, written back to the caller's variable before return (unless
 * reductions are deferred, see op2_set_deferred_reductions).
 * Returns 0 on success. */
int op2_par_loop(char const *name, op_set set, int nargs, op_arg *args);

/* No counterpart in the reference: op2_par_loop for callers that cannot
 * allocate device memory (code under graph capture by the caller, tight
 * serving loops, harnesses that forbid hipMalloc).  Runs save_soln,
 * adt_calc, and the airfoil and jac1 update loops (their OP_GBL reductions
 * included) without any device allocation: the reduction partials live in
 * a fixed device array of the library (at most 2 results of 2048 workgroup
 * partials), one per GPU, shared by every call on that GPU (calls that
 * overlap on one GPU race).  Needs op2_init first and, for adt_calc, its
 * map on the device (run the loop once through op2_par_loop); other loops
 * are refused with an error.  Not timed (op_timing_output) and not counted
 * by checkpoints.  Same results as op2_par_loop bit for bit.  Returns 0. */
int op2_par_loop_noalloc(char const *name, op_set set, int nargs, op_arg *args);

/* ---------------------------------------------------------------------------
 * Runtime controls (no counterpart in the reference)
 */

#define OP2_ERRORS_EXIT 0
#define OP2_ERRORS_RETURN 1
int op2_set_error_mode(int mode);
/* message of the last error ("" if none since op2_clear_error) */
char const *op2_last_error(void);
void op2_clear_error(void);

/* Execution strategies for indirect loops with OP_INC arguments.  AUTO (no
 * configuration) is OWNER_GATHER, i.e. OWNER for loops without a gathering
 * kernel. */
#define OP2_STRATEGY_AUTO 0
/* block colours (one launch per colour) and thread colours inside a block:
 * conflict-free, deterministic */
#define OP2_STRATEGY_COLOURED 1
/* one launch; LDS atomics inside a block, global atomics across blocks */
#define OP2_STRATEGY_ATOMIC 2
/* one launch, no plan: every element gathers straight from global memory
 * and increments with global atomics (loops that support it) */
#define OP2_STRATEGY_DIRECT 3
/* one launch, owner-compute: the OP_INC targets are cut into compact tiles,
 * one per block; a block evaluates every element that increments one of its
 * targets (elements between two tiles run in both) and adds each target's
 * increments in a fixed order with a plain read-modify-write: no atomics,
 * no colours, deterministic */
#define OP2_STRATEGY_OWNER 4
/* owner-compute tiles without LDS staging: elements gather their records
 * straight from global memory (loops that support it; others use OWNER) */
#define OP2_STRATEGY_OWNER_GATHER 5

/* Per-loop tuning: strategy, part size (elements per block, 0 = default)
 * and threads per workgroup of the loop's kernels (0 = OP_block_size, whose
 * default is 256; otherwise 128, 256 or 512, else an error) -- the
 * per-loop OP_BLOCK_SIZE_<k> of the reference's generated stubs
 * (airfoil/res_calc_kernel.cu:200-206).  name NULL applies to every loop.
 * Kernels that fuse several loops in a captured graph use the block size of
 * the loop that starts them (res_calc for a Runge-Kutta stage, update for
 * update + save_soln + adt_calc). */
int op2_set_loop_config(char const *name, int strategy, int part_size,
                        int block_size);

/* The configuration a loop runs with now (after op2_set_loop_config and
 * the autotuners): strategy (0 = the default), part size (0 = default) and
 * block size (threads per workgroup, resolved).  Returns 0. */
int op2_get_loop_config(char const *name, int *strategy, int *part_size, int *block_size);

/* Autotuner: runs the loop (with these arguments, on the live data) under
 * each candidate strategy and part size, then the best of them with each
 * block size (128, 256, 512 threads), reps times each, keeps the fastest as
 * the loop's configuration and reports it (the block size through
 * op2_get_loop_config).  Dats the loop writes are saved before and restored
 * after, so the state is unchanged.  Only indirect loops without OP_GBL
 * results (res_calc, bres_calc, jac1 res).  Returns 0 on success. */
int op2_autotune_loop(char const *name, op_set set, int nargs, op_arg *args, int reps,
                      int *best_strategy, int *best_part, double *best_us);

/* Online autotuner: from now on, the calls of loop `name` run each
 * candidate strategy, part size and (for res_calc's owner-gather tiles)
 * block size in turn, one untimed call then `reps`
 * timed calls per candidate, and every call is a real execution of the loop
 * (so the candidates are timed inside the program's own loop sequence, with
 * its cache state).  After the last candidate the fastest becomes the loop's
 * configuration and the losers' plans are freed.  Loops that cannot be tuned
 * (direct loops, loops with OP_GBL results) are refused at their first call.
 * Graph capture is refused while a tuning is unfinished.  Returns 0. */
int op2_autotune_online(char const *name, int reps);

/* Result of an online tuning: 1 when finished (outputs set), 0 while it
 * runs, -1 if none was started for `name`. */
int op2_autotune_online_result(char const *name, int *best_strategy, int *best_part,
                               double *best_us);

/* Renumbering for locality: the op_partition planned by the reference
 * (doc/user.tex:351-356, doc/dev.tex:191-206), done inside the runtime.
 * The target set of `base` (cells, for the edge -> cell map) is ordered by
 * recursive bisection of its adjacency graph (aligned runs of up to 512
 * elements are compact pieces); every set connected to it through the
 * declared maps follows (elements sorted by the first element of an ordered
 * set they reference, or numbered by first touch).  Device data and maps
 * are permuted; the user's arrays keep the user's numbering, op_fetch_data
 * returns data in it, and dats and maps declared later are translated.
 * Plans (and op_plan_get views) use the internal numbering.  Call before
 * capturing graphs; at most once per op_init.  Returns 0.
 * Programs that do not call it get the same automatically: the runtime
 * renumbers at the first loop with an indirect OP_INC argument, on that
 * argument's map, when op2_set_auto_renumber(1) was called before op_init
 * (the C++ API's op_init of op_lib.h does) or OP2_RENUMBER=auto (1, on) is
 * in the environment at op_init; OP2_RENUMBER=0 (off, no) keeps the
 * program's numbering. */
int op2_renumber(op_map base);

/* Whether op_init arms the automatic renumbering above when OP2_RENUMBER
 * is not in the environment (default 0).  Returns the previous setting. */
int op2_set_auto_renumber(int on);

/* The internal number of every element of `set` (new_of_old[set size]; may
 * be NULL): returns 1 if the set was renumbered, 0 if its numbering is the
 * user's. */
int op2_renumber_export(op_set set, int *new_of_old);

/* ---- generic loop path (no counterpart function in the reference; it
 * takes the place of the code common/op2.m:142-1141 generates per loop) ----
 * A translation unit compiled with hipcc instantiates the header-only
 * executor of include/op_hip_loop.h for the user's kernels and registers
 * each loop by name; op_par_loop / op2_par_loop then run a registered loop
 * whenever no built-in kernel matches the loop's name and arguments. */
typedef int (*op2_loop_fn)(char const *name, op_set set, int nargs, op_arg *args);
/* A registered loop runs the user's kernel even where a built-in kernel has
 * the same name and argument signature.  fn NULL removes the registration. */
int op2_register_loop(char const *name, op2_loop_fn fn);

/* Device copy of a constant: called with the constant's bytes and the
 * runtime stream whenever op_decl_const / op_decl_const2 declares `name`
 * (and at registration if it is already declared). */
typedef void (*op2_const_copy_fn)(void const *data, size_t bytes, void *stream);
int op2_register_const(char const *name, op2_const_copy_fn fn);

/* Execution arrays of the loop's generic plan (built and verified on first
 * use): contiguous blocks of `part_size` elements (0 = the loop's configured
 * part size, else 256) whose elements
 * are sorted by thread colour, thread colours and block colours such that
 * no two elements running together write the same indirect target (every
 * access other than OP_READ counts).  Pointers are device arrays except
 * colblk_off ([ncolors+1] on the host). */
typedef struct {
  int nblocks, ncolors;
  int const *blk_off, *perm, *col_off, *col_base, *blk_ncol, *blkmap;
  int const *colblk_off;
} op2_generic_plan;
int op2_generic_plan_get(char const *name, op_set set, int nargs, op_arg *args, int part_size,
                         op2_generic_plan *out);
/* strategy value under which op2_plan_summary_get reports the generic plan */
#define OP2_PLAN_GENERIC 16

/* Owner-compute plan of a generic loop whose indirect writes are all
 * OP_INC through one map, at most 4 (map, index) pairs in all: the
 * incremented targets cut into tiles of `part_size` (0 = the loop's
 * configured part size, else 128); tile b holds the positions
 * [blk_off[b], blk_off[b+1]) (perm: the element at each position, every
 * element incrementing one of the tile's targets, ascending; ids: per
 * position the target of each (map, index) pair, slot_stride ints), owns
 * the targets owned_off[b] .. owned_off[b+1]-1 of `owned` (owned = NULL:
 * the targets with those numbers), and each owned target i lists its
 * increments csr[csr_off[i] .. csr_off[i+1]-1] = position in the tile << 2
 * | k, in element order, k = the index of the increment's (map, index)
 * among the loop's OP_INC pairs (in argument order); when no target has
 * more than four increments, ell holds the same lists as one 8-byte record
 * of four entries per owned target (0xffff: none), else NULL.  arg_info[a] = k of
 * argument a (-1 if it is not an indirect OP_INC) and arg_info[nargs + a] =
 * its pair's place in an ids row (-1 if it is not indirect).  Device arrays
 * held by the runtime.  Returns 0, or 1 (nothing filled, no error) when the
 * loop does not have that form.  No counterpart in the reference (its
 * plans are coloured). */
typedef struct {
  int nblocks, max_positions, max_owned, slot_stride;
  int const *blk_off, *perm, *ids, *owned_off, *owned, *csr_off;
  unsigned short const *csr, *ell;
} op2_generic_owner_plan;
int op2_generic_owner_plan_get(char const *name, op_set set, int nargs, op_arg *args,
                               int part_size, op2_generic_owner_plan *out, int *arg_info);

/* Reports an error of a generic loop through the runtime's error mode
 * (prints and exits, or records it for op2_last_error); returns -1. */
int op2_report_error(char const *msg);

/* Device copy of a map in the runtime's numbering (uploaded on first use). */
int const *op2_map_device_ptr(op_map map);

/* Runtime-owned device scratch of at least `bytes` (valid until the next
 * call that asks for more). */
void *op2_device_scratch(size_t bytes);

/* With deferred reductions on (op2_set_deferred_reductions), the device
 * reduction log, its entry counter and its capacity (a power of two);
 * returns 1, else 0.  The generic path appends its OP_GBL results there, as
 * the built-in loops do, so that its loops can be captured into graphs.  No
 * counterpart in the reference. */
int op2_reduction_log_device(double **log, unsigned **count, unsigned *capacity);

/* Checkpoint and fast-forward restart (doc/dev.tex:1080-1087 describes
 * op_checkpoint; the reference does not implement it).
 * op2_checkpoint_record(path): from now on count loop calls and keep the
 * values of OP_GBL OP_INC/OP_MIN/OP_MAX arguments after each loop (deferred
 * reductions must be off).  op_checkpoint(): write every dat (in the user's
 * numbering), the counts and the kept reduction values to `path`; returns 0.
 * op2_checkpoint_restart(path), right after op_init in a new run of the
 * same program: loop calls before that checkpoint are skipped and hand back
 * their recorded reduction values; the op_checkpoint call that wrote the file
 * restores every dat and returns 1, and recording continues.  Returns -1 on
 * error. */
int op2_checkpoint_record(char const *path);
int op2_checkpoint_restart(char const *path);
int op_checkpoint(void);
/* 0 off, 1 recording, 2 fast-forwarding; loop calls and checkpoints so far */
int op2_checkpoint_state(unsigned long long *loops, unsigned long long *calls);

/* Device seconds, call count and algorithmic bytes per call of a loop
 * (device time from HIP events on the runtime stream).  Returns 0 if the
 * loop has run. */
int op2_loop_stats(char const *name, int *count, double *seconds,
                   double *bytes_per_call);
void op2_timing_reset(void);

/* The runtime's HIP stream (hipStream_t). */
void *op2_stream(void);
/* Waits for all work on the runtime stream. */
int op2_sync(void);

/* Deferred reductions: while on, OP_INC / OP_MIN / OP_MAX globals are not
 * written back to the caller's variable; each loop appends its results
 * (one double per global argument, in argument order) to a device log, and
 * loops never wait for the device.  Returns the previous setting. */
int op2_set_deferred_reductions(int on);
/* Waits for the stream, then copies up to max pending log entries (oldest
 * first) to out; returns how many were copied (-1 on error, e.g. more than
 * 65536 entries pending). */
int op2_reduction_log(double *out, int max);

/* Graph capture of a loop sequence on the runtime stream (hipGraph).
 * op2_graph_begin starts capturing (deferred reductions required; every
 * captured loop must have run once before, so that its plan and maps are
 * already on the device), op2_graph_end instantiates the graph and returns
 * its id, op2_graph_launch replays it count times.  OP_READ globals are
 * captured by value.  op2_exit destroys every graph. */
int op2_graph_begin(void);
int op2_graph_end(void);
int op2_graph_launch(int id, int count);
int op2_graph_destroy(int id);

/* Loop fusion during graph capture.  Level 1: a captured bres_calc joins
 * the update of the cells whose res it increments; an update joins the
 * save_soln that copies its q and the adt_calc that reads it, so q is read
 * once; a save_soln alone joins that adt_calc; the rms sum of an update is
 * folded by the next res_calc's kernel; and a res that an earlier captured
 * update zeroed is taken as zero (not read, and not stored when the next
 * owner-compute res_calc overwrites every cell).  Level 2 (the default)
 * adds one kernel per Runge-Kutta stage: a res_calc on an owner-gather plan
 * whose tiles own every cell, the bres_calc, update, next save_soln and
 * adt_calc that follow on the same cells and dats run in one kernel, each
 * tile finishing the cells it owns (res is not stored between stages; q and
 * adt alternate between the dat and a second buffer the runtime allocates
 * at op2_graph_begin, and are back in the dat at the end of the graph).
 * q, adt and res keep the bits of the separate loops; at level 2 the rms
 * sums are grouped per tile, so they agree to double rounding.  Fusion sees
 * only op_par_loop calls: other work captured on op2_stream that reads or
 * writes a dat calls for fusion off during that capture.  Returns the
 * previous level, or -1 during a capture (no change). */
int op2_set_graph_fusion(int level);

/* Tiles of the level-2 stage kernels: cells per tile and threads per
 * workgroup (128, 256 or 512; part_size at most twice the threads).  A
 * stage tile's plan is built at op2_graph_begin from the res_calc that ran
 * on owner-gather tiles before it, with this part size; res_calc's own
 * launches keep their configuration.  part_size 0: 256 cells on 256
 * threads (the default); -1: res_calc's configuration.  Returns 0, or -1 on
 * an invalid pair or during a capture (no change). */
int op2_set_stage_config(int part_size, int block_size);

/* Kernel timing inside graphs.  With every = n > 0, the stage kernels
 * (fusion level 2) of the captures that follow are counted from 1 in each
 * capture and those numbered n, 2n, 3n, ... are bracketed by two
 * event-record nodes of their graph (each node costs a few microseconds of
 * the replay, hence the sampling); 0 turns it off.
 * op2_graph_kernel_times waits for the runtime stream and returns
 * the number of bracketed kernels of graph id, filling up to max entries
 * (launch order) of kinds[] and of ms[] with each one's device time in its
 * graph's last launch.  kind: 1 the kernel runs bres_calc, 2 the next
 * save_soln, 4 the adt_calc that follows, 8 it reads res (else res is taken
 * as zero), 16 it stores res (zeros), 32 its adt_calc takes the cells' sides
 * from the tile's edges and a 2-byte recipe per cell (else it reads pcell
 * and the nodes), 64 its save_soln is its q store (the new q written over
 * qold), 128 the qold it reads is its q (left there by such a save_soln).
 * op2_set_graph_kernel_timing returns
 * the previous setting, or -1 during a capture (no change).  No counterpart
 * in the reference (its OP_kernels records time whole loops outside
 * graphs). */
int op2_set_graph_kernel_timing(int every);
int op2_graph_kernel_times(int id, int max, int *kinds, float *ms);

/* Device pointer of a dat (for zero-copy interop).  Other code may then
 * write the dat between loops: from this call on the runtime assumes
 * nothing about its contents (it otherwise knows, for example, that an
 * update left res zero, and the next res_calc does not read it). */
void *op2_dat_device_ptr(op_dat dat);

/* Plan introspection for the plan of the given strategy (COLOURED, ATOMIC,
 * OWNER or OWNER_GATHER) and part size (0 = 512; targets per tile for the
 * owner strategies): block count, block colours, largest block, largest
 * number of staged elements of one block (summed over its maps), mean
 * thread colours, the number of violations found by the exact plan
 * verifier (0 = every block and thread colouring is valid and every local
 * index maps back to the user's map), and for owner plans the largest
 * number of increments of one target (up to 4 they are packed in one
 * 8-byte record per target). */
typedef struct {
  int nblocks, ncolors, max_block, max_staged, strategy, part_size;
  double mean_thread_colours;
  long long verify_errors;
  int max_incs;
  int cached_plans; /* execution plans the runtime holds after this call */
  int short_ids;    /* OWNER_GATHER: tiles whose positions have 16-bit node and cell
                       records (each map's targets in two windows of 32768 ids) */
} op2_plan_summary;
int op2_plan_summary_get(char const *name, op_set set, int nargs, op_arg *args, int strategy,
                         int part_size, op2_plan_summary *out);

/* Plan arrays for independent checking: perm[set size] (element at each
 * position), thrcol[set size] (thread colour of each position), blk_off
 * [nblocks+1] (positions of each block) and blkcol[nblocks]; any pointer may
 * be NULL.  Returns nblocks, or -1 on error.  Plans are built and verified
 * on the host, so this needs no GPU. */
int op2_plan_export(char const *name, op_set set, int nargs, op_arg *args, int strategy,
                    int part_size, int *perm, int *thrcol, int *blk_off, int *blkcol);

#ifdef __cplusplus
}
#endif

#endif
