// Execution plans for indirect loops, built for gfx950 (64-wide waves,
// 160 KB LDS per CU).
//
// The reference plan (common/op_lib_core.c:627-1089) splits the set into
// blocks, renumbers each indirectly accessed dat per block for shared-memory
// staging, and colours threads inside a block and blocks across the grid so
// that no two concurrent elements increment the same target.  This builder
// keeps that structure with a different layout:
//   * staging is per map, not per (dat, access): q, adt and res of res_calc
//     all go through pecell and share one list of distinct cells per block;
//   * the local indices of an element are packed per (map, idx) slot into one
//     16-bit vector (res_calc: 4 x 16 bit = one 8-byte load per edge);
//   * inside a block the elements are sorted by thread colour, so a colour is
//     a contiguous range of positions and needs no per-element colour array;
//   * blocks are grouped by block colour for the coloured strategy, and the
//     atomic strategy needs no block colours at all;
//   * every plan is checked by an exact verifier before use.
#include <algorithm>
#include <cstring>
#include <numeric>

#include "op2_internal.h"

namespace op2 {

static bool same_key(const Plan &p, const PlanRequest &r) {
  if (p.loop != r.loop || p.set != r.set || p.strategy != r.strategy ||
      p.part_size != r.part_size || (int)p.maps.size() != r.nargs)
    return false;
  for (int a = 0; a < r.nargs; a++)
    if (p.maps[a] != r.args[a].map || p.idxs[a] != r.args[a].idx ||
        p.accs[a] != (int)r.args[a].acc || p.dats[a] != r.args[a].dat)
      return false;
  return true;
}

static inline int target(op_map m, int idx, int e) {
  return map_host(m)[(size_t)e * m->dim + idx];
}

void Plan::upload() {
  if (uploaded) return;
  uploaded = true;
  auto up = [](const void *src, size_t bytes) -> void * {
    void *d = dev_alloc(bytes);
    if (src) h2d(d, src, bytes);  // (an empty vector has no data: nothing to copy)
    return d;
  };
  d_blk_off = (int *)up(blk_off.data(), sizeof(int) * blk_off.size());
  if (need_perm) d_perm = (int *)up(perm.data(), sizeof(int) * perm.size());
  for (auto &g : groups) {
    g.d_off = (int *)up(g.off.data(), sizeof(int) * g.off.size());
    g.d_list = (int *)up(g.list.data(), sizeof(int) * std::max<size_t>(g.list.size(), 1));
  }
  if (!loc.empty()) d_loc = (uint16_t *)up(loc.data(), sizeof(uint16_t) * loc.size());
  if (!col_off.empty()) {
    d_col_off = (int *)up(col_off.data(), sizeof(int) * col_off.size());
    d_col_base = (int *)up(col_base.data(), sizeof(int) * col_base.size());
    d_blk_ncol = (int *)up(blk_ncol.data(), sizeof(int) * blk_ncol.size());
  }
  d_blkmap = (int *)up(blkmap.data(), sizeof(int) * blkmap.size());
  if (!ids.empty()) d_ids = (int *)up(ids.data(), sizeof(int) * ids.size());
  if (!owned_off.empty()) {
    d_owned_off = (int *)up(owned_off.data(), sizeof(int) * owned_off.size());
    d_owned = (int *)up(owned.data(), sizeof(int) * std::max<size_t>(owned.size(), 1));
    d_csr_off = (int *)up(csr_off.data(), sizeof(int) * csr_off.size());
    d_csr = (uint16_t *)up(csr.data(), sizeof(uint16_t) * std::max<size_t>(csr.size(), 1));
    if (!ell.empty()) d_ell = (uint16_t *)up(ell.data(), sizeof(uint16_t) * ell.size());
  }
}

void Plan::release() {
  for (void *p : {(void *)d_blk_off, (void *)d_perm, (void *)d_col_off, (void *)d_col_base,
                  (void *)d_blk_ncol, (void *)d_blkmap, (void *)d_loc, (void *)d_owned_off,
                  (void *)d_owned, (void *)d_csr_off, (void *)d_csr, (void *)d_ell,
                  (void *)d_ids})
    dev_free(p);
  d_ids = nullptr;
  d_owned_off = d_owned = d_csr_off = nullptr;
  d_csr = d_ell = nullptr;
  for (auto &g : groups) {
    dev_free(g.d_off);
    dev_free(g.d_list);
  }
  d_blk_off = d_perm = d_col_off = d_col_base = d_blk_ncol = d_blkmap = nullptr;
  d_loc = nullptr;
  uploaded = false;
}

// Greedy colouring with 32-bit masks, 32 colours per pass, over the INC
// targets of `items` (elements of one block, or whole blocks).  Each item
// lists its targets as (group, target) through `each`.
template <class Each>
static int greedy_colour(int nitems, std::vector<int> &colour,
                         std::vector<std::vector<uint32_t>> &mask, Each each) {
  colour.assign(nitems, -1);
  int base = 0, ncol = 0;
  bool pending = true;
  while (pending) {
    pending = false;
    // clear the masks of every target any item touches
    for (int i = 0; i < nitems; i++)
      each(i, [&](int g, int t) { mask[g][t] = 0; });
    for (int i = 0; i < nitems; i++) {
      if (colour[i] >= 0) continue;
      uint32_t used = 0;
      each(i, [&](int g, int t) { used |= mask[g][t]; });
      if (used == 0xffffffffu) {
        pending = true;
        continue;
      }
      int c = __builtin_ctz(~used);
      colour[i] = base + c;
      ncol = std::max(ncol, base + c + 1);
      each(i, [&](int g, int t) { mask[g][t] |= 1u << c; });
    }
    base += 32;
  }
  return ncol;
}


// Owner-compute layout.  The distinct targets of the OP_INC slots (all
// through one map) are cut into tiles of about `tile` targets: when an
// element increments two or more targets, tiles grow breadth-first over that
// adjacency (compact tiles, few cut elements); otherwise they are runs of
// ascending target ids.  Tiles are seeded in ascending id order, so tile
// numbers follow the user's numbering.
static void build_owner_layout(Plan &P, const std::vector<int> &inc_slots, int tile) {
  const int n = P.set->size;
  const int og = P.slots[inc_slots[0]].first;
  for (int s : inc_slots)
    if (P.slots[s].first != og)
      fail("op2: owner-compute needs every OP_INC argument of %s through one map", P.loop.c_str());
  P.owner_group = og;
  P.inc_slot_ids = inc_slots;
  op_map om = P.groups[og].map;
  const int nt = om->to->size;
  const int ni = (int)inc_slots.size();
  auto tgt = [&](int e, int k) { return target(om, P.slots[inc_slots[k]].second, e); };

  std::vector<int> tile_of(nt, -2);  // -2: never incremented
  for (int e = 0; e < n; e++)
    for (int k = 0; k < ni; k++) tile_of[tgt(e, k)] = -1;

  int ntiles = 0;
  auto segs = rt().set_segments.find(om->to);
  if (ni >= 2 && segs == rt().set_segments.end()) {
    // adjacency between targets incremented by one element
    std::vector<int> deg(nt + 1, 0);
    for (int e = 0; e < n; e++)
      for (int k = 0; k < ni; k++)
        for (int l = 0; l < ni; l++)
          if (k != l && tgt(e, k) != tgt(e, l)) deg[tgt(e, k) + 1]++;
    for (int t = 0; t < nt; t++) deg[t + 1] += deg[t];
    std::vector<int> adj(deg[nt]);
    std::vector<int> fill(deg.begin(), deg.end() - 1);
    for (int e = 0; e < n; e++)
      for (int k = 0; k < ni; k++)
        for (int l = 0; l < ni; l++)
          if (k != l && tgt(e, k) != tgt(e, l)) adj[fill[tgt(e, k)]++] = tgt(e, l);
    std::vector<int> queue(nt);
    for (int seed = 0; seed < nt; seed++) {
      if (tile_of[seed] != -1) continue;
      int head = 0, tail = 0, count = 1;
      tile_of[seed] = ntiles;
      queue[tail++] = seed;
      while (head < tail && count < tile) {
        const int u = queue[head++];
        for (int j = deg[u]; j < deg[u + 1] && count < tile; j++) {
          const int v = adj[j];
          if (tile_of[v] != -1) continue;
          tile_of[v] = ntiles;
          queue[tail++] = v;
          count++;
        }
      }
      ntiles++;
    }
  } else if (ni >= 2) {
    // renumbered targets: aligned runs of `tile` numbers inside each piece
    // of the renumbering are compact (one subtree of its bisection)
    const std::vector<int> &st = segs->second;
    size_t sg = 0;
    int base = 0, cur = -1;
    for (int t = 0; t < nt; t++) {
      while (sg < st.size() && st[sg] == t) base = st[sg++], cur = -1;
      if ((t - base) % tile == 0) cur = -1;
      if (tile_of[t] != -1) continue;
      if (cur < 0) cur = ntiles++;
      tile_of[t] = cur;
    }
  } else {
    int count = 0;
    for (int t = 0; t < nt; t++) {
      if (tile_of[t] != -1) continue;
      tile_of[t] = ntiles;
      if (++count == tile) {
        count = 0;
        ntiles++;
      }
    }
    if (count) ntiles++;
  }
  P.nblocks = ntiles;

  // owned targets per tile (ascending)
  P.owned_off.assign(ntiles + 1, 0);
  for (int t = 0; t < nt; t++)
    if (tile_of[t] >= 0) P.owned_off[tile_of[t] + 1]++;
  for (int b = 0; b < ntiles; b++) P.owned_off[b + 1] += P.owned_off[b];
  P.owned.resize(P.owned_off[ntiles]);
  std::vector<int> local(nt, -1);
  {
    std::vector<int> fill(P.owned_off.begin(), P.owned_off.end() - 1);
    for (int t = 0; t < nt; t++)
      if (tile_of[t] >= 0) {
        local[t] = fill[tile_of[t]] - P.owned_off[tile_of[t]];
        P.owned[fill[tile_of[t]]++] = t;
      }
  }
  for (int b = 0; b < ntiles; b++)
    P.max_owned = std::max(P.max_owned, P.owned_off[b + 1] - P.owned_off[b]);
  P.owned_identity = true;
  for (size_t i = 0; i < P.owned.size() && P.owned_identity; i++)
    P.owned_identity = P.owned[i] == (int)i;

  // elements of each tile: those incrementing one of its targets (ascending)
  std::vector<int> cnt(ntiles + 1, 0);
  for (int e = 0; e < n; e++)
    for (int k = 0; k < ni; k++) {
      const int b = tile_of[tgt(e, k)];
      bool first = true;
      for (int l = 0; l < k; l++) first &= tile_of[tgt(e, l)] != b;
      if (first) cnt[b + 1]++;
    }
  for (int b = 0; b < ntiles; b++) cnt[b + 1] += cnt[b];
  P.blk_off = cnt;
  P.perm.resize(cnt[ntiles]);
  {
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int e = 0; e < n; e++)
      for (int k = 0; k < ni; k++) {
        const int b = tile_of[tgt(e, k)];
        bool first = true;
        for (int l = 0; l < k; l++) first &= tile_of[tgt(e, l)] != b;
        if (first) P.perm[fill[b]++] = e;
      }
  }
  P.max_block = 0;
  for (int b = 0; b < ntiles; b++) P.max_block = std::max(P.max_block, cnt[b + 1] - cnt[b]);
  P.max_positions = P.max_block;
  if (P.max_block >= (1 << 14))
    fail("op2: owner tile of %s holds %d elements (limit 16383); reduce the part size",
         P.loop.c_str(), P.max_block);

  // per owned target: the (position, inc slot) pairs that increment it
  P.csr_off.assign(P.owned.size() + 1, 0);
  for (int b = 0; b < ntiles; b++)
    for (int pos = cnt[b]; pos < cnt[b + 1]; pos++)
      for (int k = 0; k < ni; k++) {
        const int t = tgt(P.perm[pos], k);
        if (tile_of[t] == b) P.csr_off[P.owned_off[b] + local[t] + 1]++;
      }
  for (size_t i = 0; i < P.owned.size(); i++) P.csr_off[i + 1] += P.csr_off[i];
  P.csr.resize(P.csr_off.back());
  {
    std::vector<int> fill(P.csr_off.begin(), P.csr_off.end() - 1);
    for (int b = 0; b < ntiles; b++)
      for (int pos = cnt[b]; pos < cnt[b + 1]; pos++)
        for (int k = 0; k < ni; k++) {
          const int t = tgt(P.perm[pos], k);
          if (tile_of[t] == b)
            P.csr[fill[P.owned_off[b] + local[t]]++] = (uint16_t)(((pos - cnt[b]) << 2) | k);
        }
  }
  // positions stay below 1 << 14, so no entry equals the 0xffff pad
  for (size_t i = 0; i < P.owned.size(); i++)
    P.max_incs = std::max(P.max_incs, P.csr_off[i + 1] - P.csr_off[i]);
  if (P.max_incs <= 4) {
    P.ell.assign(4 * P.owned.size(), 0xffff);
    for (size_t i = 0; i < P.owned.size(); i++)
      for (int j = P.csr_off[i]; j < P.csr_off[i + 1]; j++) P.ell[4 * i + j - P.csr_off[i]] = P.csr[j];
  }
}

static Plan *build(const PlanRequest &req) {
  auto p = std::make_unique<Plan>();
  Plan &P = *p;
  const int n = req.set->size;
  P.loop = req.loop;
  P.set = req.set;
  P.strategy = req.strategy;
  P.part_size = req.part_size;
  for (int a = 0; a < req.nargs; a++) {
    P.maps.push_back(req.args[a].map);
    P.idxs.push_back(req.args[a].idx);
    P.accs.push_back((int)req.args[a].acc);
    P.dats.push_back(req.args[a].dat);
  }

  // groups (distinct maps) and slots (distinct (map, idx))
  const bool generic = req.strategy == kStrategyGeneric;
  std::vector<int> inc_slots;
  P.arg_slot.assign(req.nargs, -1);
  bool has_direct_dat = false;
  for (int a = 0; a < req.nargs; a++) {
    const op_arg &arg = req.args[a];
    if (arg.argtype != OP_ARG_DAT) continue;
    if (!arg.map) {
      has_direct_dat = true;
      continue;
    }
    int g = -1;
    for (size_t k = 0; k < P.groups.size(); k++)
      if (P.groups[k].map == arg.map) g = (int)k;
    if (g < 0) {
      P.groups.emplace_back();
      P.groups.back().map = arg.map;
      g = (int)P.groups.size() - 1;
    }
    int s = -1;
    for (size_t k = 0; k < P.slots.size(); k++)
      if (P.slots[k].first == g && P.slots[k].second == arg.idx) s = (int)k;
    if (s < 0) {
      P.slots.emplace_back(g, arg.idx);
      s = (int)P.slots.size() - 1;
    }
    P.arg_slot[a] = s;
    // the generic path colours on every indirect write (its kernels write
    // to global memory directly), the device kernels on increments
    const bool writes = arg.acc == OP_INC || (generic && arg.acc != OP_READ);
    if (writes && std::find(inc_slots.begin(), inc_slots.end(), s) == inc_slots.end())
      inc_slots.push_back(s);
  }
  const int nslots = (int)P.slots.size();
  if (nslots == 0) fail("op2: plan requested for loop %s without indirect arguments", req.loop.c_str());
  if (nslots > 4 && !generic)
    fail("op2: loop %s has %d (map, index) pairs; at most 4 are supported", req.loop.c_str(), nslots);
  P.slot_stride = generic ? 0 : nslots == 1 ? 1 : nslots == 2 ? 2 : 4;

  const int bsize = req.part_size > 0 ? req.part_size : 512;
  const bool owner = req.strategy == OP2_STRATEGY_OWNER || req.strategy == OP2_STRATEGY_OWNER_GATHER;
  if (owner) {
    if (inc_slots.empty()) fail("op2: owner-compute plan for %s without OP_INC arguments", req.loop.c_str());
    build_owner_layout(P, inc_slots, bsize);
  } else {
    P.nblocks = (n + bsize - 1) / bsize;
    P.blk_off.resize(P.nblocks + 1);
    for (int b = 0; b <= P.nblocks; b++) P.blk_off[b] = std::min(n, b * bsize);
    P.max_block = std::min(n, bsize);
  }

  const int ng = (int)P.groups.size();
  std::vector<std::vector<uint32_t>> mask(ng);
  std::vector<std::vector<int>> local(ng);
  for (int g = 0; g < ng; g++) {
    mask[g].assign(P.groups[g].map->to->size, 0);
    local[g].assign(P.groups[g].map->to->size, -1);
    P.groups[g].off.assign(P.nblocks + 1, 0);
  }

  // thread colours inside each block, then positions sorted by colour
  const bool coloured =
      (req.strategy == OP2_STRATEGY_COLOURED || generic) && !inc_slots.empty();
  if (coloured) {
    P.perm.resize(n);
    P.thrcol.resize(n);
    P.col_base.resize(P.nblocks);
    P.blk_ncol.resize(P.nblocks);
    std::vector<int> colour;
    double total = 0;
    for (int b = 0; b < P.nblocks; b++) {
      const int lo = P.blk_off[b], cnt = P.blk_off[b + 1] - lo;
      auto each = [&](int i, auto &&fn) {
        for (int s : inc_slots) {
          op_map m = P.groups[P.slots[s].first].map;
          fn(P.slots[s].first, target(m, P.slots[s].second, lo + i));
        }
      };
      int ncol = greedy_colour(cnt, colour, mask, each);
      std::vector<int> order(cnt);
      std::iota(order.begin(), order.end(), 0);
      std::stable_sort(order.begin(), order.end(),
                       [&](int x, int y) { return colour[x] < colour[y]; });
      P.col_base[b] = (int)P.col_off.size();
      P.blk_ncol[b] = ncol;
      int k = 0;
      for (int c = 0; c < ncol; c++) {
        P.col_off.push_back(lo + k);
        while (k < cnt && colour[order[k]] == c) k++;
      }
      P.col_off.push_back(lo + cnt);
      for (int i = 0; i < cnt; i++) {
        P.perm[lo + i] = lo + order[i];
        P.thrcol[lo + i] = colour[order[i]];
      }
      total += ncol;
    }
    P.mean_thread_colours = total / P.nblocks;
  }
  auto elem = [&](int pos) { return P.perm.empty() ? pos : P.perm[pos]; };
  const size_t npos = P.blk_off[P.nblocks];

  // per block and group: distinct targets (ascending) and local indices
  P.loc.assign(npos * P.slot_stride, 0);
  std::vector<int> buf;
  for (int b = 0; b < P.nblocks && !generic; b++) {
    const int lo = P.blk_off[b], hi = P.blk_off[b + 1];
    for (int g = 0; g < ng; g++) {
      PlanGroup &G = P.groups[g];
      buf.clear();
      for (int s = 0; s < nslots; s++)
        if (P.slots[s].first == g)
          for (int pos = lo; pos < hi; pos++) buf.push_back(target(G.map, P.slots[s].second, elem(pos)));
      std::sort(buf.begin(), buf.end());
      buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
      G.off[b] = (int)G.list.size();
      for (size_t k = 0; k < buf.size(); k++) {
        local[g][buf[k]] = (int)k;
        G.list.push_back(buf[k]);
      }
      G.max_count = std::max(G.max_count, (int)buf.size());
      if (buf.size() > 65535)
        fail("op2: block of loop %s stages %zu elements; reduce the part size",
             req.loop.c_str(), buf.size());
      for (int s = 0; s < nslots; s++)
        if (P.slots[s].first == g)
          for (int pos = lo; pos < hi; pos++)
            P.loc[(size_t)pos * P.slot_stride + s] =
                (uint16_t)local[g][target(G.map, P.slots[s].second, elem(pos))];
    }
  }
  for (int g = 0; g < ng; g++) P.groups[g].off[P.nblocks] = (int)P.groups[g].list.size();

  // block colours (coloured strategy only)
  if (coloured) {
    auto each = [&](int b, auto &&fn) {
      for (int e = P.blk_off[b]; e < P.blk_off[b + 1]; e++)
        for (int s : inc_slots) {
          op_map m = P.groups[P.slots[s].first].map;
          fn(P.slots[s].first, target(m, P.slots[s].second, e));
        }
    };
    P.ncolors = greedy_colour(P.nblocks, P.blkcol, mask, each);
  } else {
    P.ncolors = 1;
    P.blkcol.assign(P.nblocks, 0);
  }
  P.colblk_off.assign(P.ncolors + 1, 0);
  for (int b = 0; b < P.nblocks; b++) P.colblk_off[P.blkcol[b] + 1]++;
  for (int c = 0; c < P.ncolors; c++) P.colblk_off[c + 1] += P.colblk_off[c];
  P.blkmap.resize(P.nblocks);
  {
    std::vector<int> fill(P.colblk_off.begin(), P.colblk_off.end() - 1);
    for (int b = 0; b < P.nblocks; b++) P.blkmap[fill[P.blkcol[b]]++] = b;
  }
  if (req.strategy == OP2_STRATEGY_OWNER_GATHER) {
    P.ids.assign(npos * P.slot_stride, 0);
    for (size_t pos = 0; pos < npos; pos++)
      for (int s = 0; s < nslots; s++)
        P.ids[pos * P.slot_stride + s] =
            target(P.groups[P.slots[s].first].map, P.slots[s].second, elem((int)pos));
  }
  // kernels read perm only for direct dats, and only if elements moved (the
  // generic kernels take every element through it)
  P.need_perm = (has_direct_dat || generic) && !P.perm.empty();
  if (owner) P.mean_thread_colours = 0;

  double tv0, tv1, cpu;
  op_timers(&cpu, &tv0);
  P.verify_errors = verify_plan(P);
  op_timers(&cpu, &tv1);
  if (OP_diags > 1) printf(" plan verification %.3f s\n", tv1 - tv0);
  if (P.verify_errors != 0)
    fail("op2: plan verification failed for loop %s (%lld violations)", req.loop.c_str(),
         P.verify_errors);

  // reference-shaped view (common/op_lib_core.h:122-150)
  P.view_offset.assign(P.blk_off.begin(), P.blk_off.end() - 1);
  P.view_nelems.resize(P.nblocks);
  for (int b = 0; b < P.nblocks; b++) P.view_nelems[b] = P.blk_off[b + 1] - P.blk_off[b];
  P.view_ncolblk.resize(P.ncolors);
  for (int c = 0; c < P.ncolors; c++) P.view_ncolblk[c] = P.colblk_off[c + 1] - P.colblk_off[c];
  P.view_nthrcol = P.blk_ncol.empty() ? std::vector<int>(P.nblocks, 1) : P.blk_ncol;
  op_plan &v = P.view;
  memset(&v, 0, sizeof v);
  v.set = P.set;
  v.nargs = req.nargs;
  v.ninds = ng;
  v.part_size = req.part_size;
  v.nblocks = P.nblocks;
  v.ncolors = P.ncolors;
  v.offset = P.view_offset.data();
  v.nelems = P.view_nelems.data();
  v.ncolblk = P.view_ncolblk.data();
  v.blkmap = P.blkmap.data();
  v.nthrcol = P.view_nthrcol.data();
  v.thrcol = P.thrcol.empty() ? nullptr : P.thrcol.data();
  return p.release();
}

// Exact verifier: every local index maps back to the user's map; positions
// of a block are a permutation of its elements; no two elements of one
// thread colour in a block, and no two blocks of one block colour, share an
// OP_INC target; colour ranges tile each block.
// Owner-compute plans: every increment (element, OP_INC slot) is applied
// exactly once, by the block that owns its target, and the owned sets
// partition the incremented targets.
static long long verify_owner(const Plan &P) {
  long long err = 0;
  const int n = P.set->size;
  const int nslots = (int)P.slots.size();
  const int ni = (int)P.inc_slot_ids.size();
  op_map om = P.groups[P.owner_group].map;
  if ((int)P.blk_off.size() != P.nblocks + 1 || P.blk_off[0] != 0) err++;
  if ((int)P.owned_off.size() != P.nblocks + 1) return err + 1;
  for (int b = 0; b < P.nblocks; b++) {
    for (int pos = P.blk_off[b]; pos < P.blk_off[b + 1]; pos++) {
      const int e = P.perm[pos];
      if (e < 0 || e >= n || (pos > P.blk_off[b] && e <= P.perm[pos - 1])) {
        err++;
        continue;
      }
      for (int s = 0; s < nslots; s++) {
        const PlanGroup &G = P.groups[P.slots[s].first];
        const int l = P.loc[(size_t)pos * P.slot_stride + s];
        if (l >= G.off[b + 1] - G.off[b]) {
          err++;
          continue;
        }
        err += G.list[G.off[b] + l] != target(G.map, P.slots[s].second, e);
      }
    }
    for (int i = P.owned_off[b] + 1; i < P.owned_off[b + 1]; i++) err += P.owned[i] <= P.owned[i - 1];
  }
  for (const auto &G : P.groups)
    for (int b = 0; b < P.nblocks; b++)
      for (int k = G.off[b] + 1; k < G.off[b + 1]; k++) err += G.list[k] <= G.list[k - 1];
  // owned sets partition the incremented targets
  std::vector<char> inc(om->to->size, 0), own(om->to->size, 0);
  for (int e = 0; e < n; e++)
    for (int k = 0; k < ni; k++) inc[target(om, P.slots[P.inc_slot_ids[k]].second, e)] = 1;
  for (int t : P.owned) {
    if (t < 0 || t >= om->to->size || own[t]) {
      err++;
      continue;
    }
    own[t] = 1;
  }
  for (int t = 0; t < om->to->size; t++) err += inc[t] != own[t];
  // every (element, inc slot) exactly once, at the owner of its target
  std::vector<long long> seen;
  seen.reserve((size_t)n * ni);
  for (int b = 0; b < P.nblocks; b++)
    for (int i = P.owned_off[b]; i < P.owned_off[b + 1]; i++)
      for (int j = P.csr_off[i]; j < P.csr_off[i + 1]; j++) {
        const int pos = P.blk_off[b] + (P.csr[j] >> 2), k = P.csr[j] & 3;
        if (pos >= P.blk_off[b + 1] || k >= ni) {
          err++;
          continue;
        }
        const int e = P.perm[pos];
        err += target(om, P.slots[P.inc_slot_ids[k]].second, e) != P.owned[i];
        seen.push_back((long long)e * ni + k);
      }
  std::sort(seen.begin(), seen.end());
  if ((long long)seen.size() != (long long)n * ni)
    err++;
  else
    for (size_t i = 0; i < seen.size(); i++) err += seen[i] != (long long)i;
  // packed records: the same entries in the same order, then padding
  if (!P.ell.empty()) {
    if (P.ell.size() != 4 * P.owned.size()) return err + 1;
    for (size_t i = 0; i < P.owned.size(); i++) {
      const int c0 = P.csr_off[i], nc = P.csr_off[i + 1] - c0;
      if (nc > 4) {
        err++;
        continue;
      }
      for (int s = 0; s < 4; s++) err += P.ell[4 * i + s] != (s < nc ? P.csr[c0 + s] : 0xffff);
    }
  }
  return err;
}

long long verify_plan(const Plan &P) {
  if (P.strategy == OP2_STRATEGY_OWNER || P.strategy == OP2_STRATEGY_OWNER_GATHER) {
    long long err = verify_owner(P);
    const int nslots = (int)P.slots.size();
    for (size_t pos = 0; pos < P.ids.size() / std::max(P.slot_stride, 1); pos++)
      for (int s = 0; s < nslots; s++)
        err += P.ids[pos * P.slot_stride + s] !=
               target(P.groups[P.slots[s].first].map, P.slots[s].second, P.perm[pos]);
    return err;
  }
  long long err = 0;
  const int n = P.set->size;
  const int nslots = (int)P.slots.size();
  if (P.blk_off.front() != 0 || P.blk_off.back() != n) err++;
  std::vector<char> seen(n, 0);
  for (int b = 0; b < P.nblocks; b++)
    for (int pos = P.blk_off[b]; pos < P.blk_off[b + 1]; pos++) {
      int e = P.perm.empty() ? pos : P.perm[pos];
      if (e < P.blk_off[b] || e >= P.blk_off[b + 1] || seen[e]) {
        err++;
        continue;
      }
      seen[e] = 1;
      for (int s = 0; s < nslots && P.strategy != kStrategyGeneric; s++) {
        const PlanGroup &G = P.groups[P.slots[s].first];
        int l = P.loc[(size_t)pos * P.slot_stride + s];
        int cnt = G.off[b + 1] - G.off[b];
        if (l >= cnt) {
          err++;
          continue;
        }
        err += G.list[G.off[b] + l] != target(G.map, P.slots[s].second, e);
      }
    }
  for (const auto &G : P.groups)
    for (int b = 0; b < P.nblocks; b++)
      for (int k = G.off[b] + 1; k < G.off[b + 1]; k++) err += G.list[k] <= G.list[k - 1];

  const bool generic = P.strategy == kStrategyGeneric;
  std::vector<int> inc_slots;
  for (size_t a = 0; a < P.maps.size(); a++)
    if ((P.accs[a] == OP_INC || (generic && P.accs[a] != OP_READ)) && P.arg_slot[a] >= 0 &&
        std::find(inc_slots.begin(), inc_slots.end(), P.arg_slot[a]) == inc_slots.end())
      inc_slots.push_back(P.arg_slot[a]);
  if ((P.strategy != OP2_STRATEGY_COLOURED && !generic) || inc_slots.empty()) return err;

  // (group, target, colour, owner) clash search
  struct Q {
    int g, t, c, o;
  };
  auto clashes = [](std::vector<Q> &v) {
    std::sort(v.begin(), v.end(), [](const Q &x, const Q &y) {
      if (x.g != y.g) return x.g < y.g;
      if (x.t != y.t) return x.t < y.t;
      if (x.c != y.c) return x.c < y.c;
      return x.o < y.o;
    });
    long long e = 0;
    for (size_t i = 1; i < v.size(); i++)
      e += v[i].g == v[i - 1].g && v[i].t == v[i - 1].t && v[i].c == v[i - 1].c &&
           v[i].o != v[i - 1].o;
    return e;
  };
  std::vector<Q> v;
  for (int b = 0; b < P.nblocks; b++) {
    v.clear();
    const int base = P.col_base[b], nc = P.blk_ncol[b];
    if (P.col_off[base] != P.blk_off[b] || P.col_off[base + nc] != P.blk_off[b + 1]) err++;
    for (int c = 0; c < nc; c++)
      for (int pos = P.col_off[base + c]; pos < P.col_off[base + c + 1]; pos++) {
        int e = P.perm[pos];
        err += P.thrcol[pos] != c;
        for (int s : inc_slots) {
          const PlanGroup &G = P.groups[P.slots[s].first];
          v.push_back({P.slots[s].first, target(G.map, P.slots[s].second, e), c, e});
        }
      }
    err += clashes(v);
  }
  v.clear();
  for (int b = 0; b < P.nblocks; b++)
    for (int e = P.blk_off[b]; e < P.blk_off[b + 1]; e++)
      for (int s : inc_slots) {
        const PlanGroup &G = P.groups[P.slots[s].first];
        v.push_back({P.slots[s].first, target(G.map, P.slots[s].second, e), P.blkcol[b], b});
      }
  err += clashes(v);
  return err;
}

void drop_plan(Plan *p) {
  Runtime &r = rt();
  for (size_t i = 0; i < r.plans.size(); i++)
    if (r.plans[i].get() == p) {
      if (r.device_ready) (void)hipStreamSynchronize(r.stream);
      p->release();
      r.plans.erase(r.plans.begin() + i);
      OP_plan_index = (int)r.plans.size();
      return;
    }
}

Plan *get_plan(const PlanRequest &req) {
  Runtime &r = rt();
  for (auto &p : r.plans)
    if (same_key(*p, req)) {
      p->count++;
      p->view.count = p->count;
      if (OP_diags > 3) printf(" old execution plan for kernel %s\n", req.loop.c_str());
      return p.get();
    }
  if (OP_diags > 1) printf(" new execution plan #%zu for kernel %s\n", r.plans.size(), req.loop.c_str());
  double t0, t1, cpu;
  op_timers(&cpu, &t0);
  Plan *p = build(req);
  op_timers(&cpu, &t1);
  if (OP_diags > 1) printf(" plan built and verified in %.3f s\n", t1 - t0);
  p->count = 1;
  p->view.count = 1;
  r.plans.emplace_back(p);
  p->view.name = p->loop.c_str();
  OP_plan_index = (int)r.plans.size();
  if (OP_diags > 1) {
    printf(" number of blocks       = %d \n", p->nblocks);
    printf(" number of block colors = %d \n", p->ncolors);
    printf(" maximum block size     = %d \n", p->max_block);
    printf(" average thread colors  = %.2f \n", p->mean_thread_colours);
  }
  return p;
}

}  // namespace op2

using namespace op2;

extern "C" {

op_plan *op_plan_core(char const *name, op_set set, int part_size, int nargs, op_arg *args,
                      int ninds, int *inds) {
  try {
    (void)ninds;
    (void)inds;
    LoopConfig c = config_for(name);
    int strategy = c.strategy ? c.strategy : OP2_STRATEGY_COLOURED;
    PlanRequest req{name, set, nargs, args, strategy, part_size ? part_size : c.part_size};
    return &get_plan(req)->view;
  } catch (const op2::Error &) {
    return nullptr;
  }
}

op_plan *op2_plan_get(char const *name, op_set set, int part_size, int nargs, op_arg *args,
                      int ninds, int *inds) {
  return op_plan_core(name, set, part_size, nargs, args, ninds, inds);
}

int op2_plan_summary_get(char const *name, op_set set, int nargs, op_arg *args, int strategy,
                         int part_size, op2_plan_summary *out) {
  try {
    if (strategy != OP2_STRATEGY_COLOURED && strategy != OP2_STRATEGY_ATOMIC &&
        strategy != OP2_STRATEGY_OWNER && strategy != OP2_STRATEGY_OWNER_GATHER &&
        strategy != OP2_PLAN_GENERIC)
      fail("op2: plan summary needs an explicit strategy");
    PlanRequest req{name, set, nargs, args, strategy, part_size};
    Plan *p = get_plan(req);
    out->nblocks = p->nblocks;
    out->ncolors = p->ncolors;
    out->max_block = p->max_block;
    out->strategy = p->strategy;
    out->part_size = p->part_size;
    out->mean_thread_colours = p->mean_thread_colours;
    out->verify_errors = verify_plan(*p);
    out->max_incs = p->max_incs;
    int staged = 0;
    for (auto &g : p->groups) staged += g.max_count;
    out->max_staged = staged;
    return 0;
  } catch (const op2::Error &) {
    return -1;
  }
}

int op2_plan_export(char const *name, op_set set, int nargs, op_arg *args, int strategy,
                    int part_size, int *perm, int *thrcol, int *blk_off, int *blkcol) {
  try {
    if (strategy != OP2_STRATEGY_COLOURED && strategy != OP2_STRATEGY_ATOMIC)
      fail("op2: plan export needs the coloured or atomic strategy");
    PlanRequest req{name, set, nargs, args, strategy, part_size};
    Plan *p = get_plan(req);
    const int n = set->size;
    for (int i = 0; i < n; i++) {
      if (perm) perm[i] = p->perm.empty() ? i : p->perm[i];
      if (thrcol) thrcol[i] = p->thrcol.empty() ? 0 : p->thrcol[i];
    }
    if (blk_off)
      for (int b = 0; b <= p->nblocks; b++) blk_off[b] = p->blk_off[b];
    if (blkcol)
      for (int b = 0; b < p->nblocks; b++) blkcol[b] = p->blkcol[b];
    return p->nblocks;
  } catch (const op2::Error &) {
    return -1;
  }
}

}  // extern "C"
