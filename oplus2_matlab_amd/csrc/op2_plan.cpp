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
#include <tuple>

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
  device_bytes = 0;
  auto up = [this](const void *src, size_t bytes) -> void * {
    device_bytes += bytes;
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
  if (!ids16.empty()) {
    d_ids16 = (uint16_t *)up(ids16.data(), sizeof(uint16_t) * ids16.size());
    d_id_win = (int *)up(id_win.data(), sizeof(int) * id_win.size());
  }
  if (!owned_off.empty()) {
    d_owned_off = (int *)up(owned_off.data(), sizeof(int) * owned_off.size());
    d_owned = (int *)up(owned.data(), sizeof(int) * std::max<size_t>(owned.size(), 1));
    d_csr_off = (int *)up(csr_off.data(), sizeof(int) * csr_off.size());
    d_csr = (uint16_t *)up(csr.data(), sizeof(uint16_t) * std::max<size_t>(csr.size(), 1));
    if (!ell.empty()) d_ell = (uint16_t *)up(ell.data(), sizeof(uint16_t) * ell.size());
  }
}

void Plan::release(std::vector<void *> *keep) {
  auto drop = [keep](void *p) {
    if (!p) return;
    if (keep)
      keep->push_back(p);
    else
      dev_free(p);
  };
  for (void *p : {(void *)d_blk_off, (void *)d_perm, (void *)d_col_off, (void *)d_col_base,
                  (void *)d_blk_ncol, (void *)d_blkmap, (void *)d_loc, (void *)d_owned_off,
                  (void *)d_owned, (void *)d_csr_off, (void *)d_csr, (void *)d_ell,
                  (void *)d_ids, (void *)d_ids16, (void *)d_id_win})
    drop(p);
  d_ids = nullptr;
  d_ids16 = nullptr;
  d_id_win = nullptr;
  d_owned_off = d_owned = d_csr_off = nullptr;
  d_csr = d_ell = nullptr;
  for (auto &g : groups) {
    drop(g.d_off);
    drop(g.d_list);
    g.d_off = g.d_list = nullptr;
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
  P.args.assign(req.args, req.args + req.nargs);

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
    // four slots on two maps (res_calc's nodes and cells): 16-bit records.
    // Each map's targets in a tile are covered greedily, in ascending order,
    // by at most two windows of kIdSpan ids; a tile that needs more keeps
    // its 32-bit ids.
    if (nslots == 4 && P.slots[0].first == P.slots[1].first &&
        P.slots[2].first == P.slots[3].first && P.slots[0].first != P.slots[2].first) {
      P.ids16.assign(npos * 4, 0);
      P.id_win.assign(4 * (size_t)P.nblocks, 0);
      std::vector<int> t;
      for (int b = 0; b < P.nblocks; b++) {
        int *w = &P.id_win[4 * (size_t)b];
        bool fits = true;
        for (int m = 0; m < 2 && fits; m++) {
          t.clear();
          for (int pos = P.blk_off[b]; pos < P.blk_off[b + 1]; pos++)
            for (int s = 2 * m; s < 2 * m + 2; s++)
              t.push_back(P.ids[(size_t)pos * P.slot_stride + s]);
          if (t.empty()) continue;
          const auto [lo, hi] = std::minmax_element(t.begin(), t.end());
          // the second window starts at the first target past the first one
          int w1 = *hi;
          for (int v : t)
            if (v - *lo >= kIdSpan) w1 = std::min(w1, v);
          fits = *hi - w1 < kIdSpan;
          if (!fits) break;
          w[2 * m] = *lo;
          w[2 * m + 1] = w1 - *lo;
          for (int pos = P.blk_off[b]; pos < P.blk_off[b + 1]; pos++)
            for (int s = 2 * m; s < 2 * m + 2; s++) {
              const int v = P.ids[(size_t)pos * P.slot_stride + s];
              const bool second = v - *lo >= kIdSpan;
              P.ids16[(size_t)pos * 4 + s] =
                  (uint16_t)(second ? 0x8000 | (v - w1) : v - *lo);
            }
        }
        if (fits) {
          P.short_tiles++;
        } else {
          w[0] = -1;
          w[1] = w[2] = w[3] = 0;
        }
      }
      if (P.short_tiles == 0) {
        P.ids16.clear();
        P.id_win.clear();
      }
    }
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
    // 16-bit records: window base + offset gives the same target
    if (!P.ids16.empty()) {
      if (P.ids16.size() != 4 * P.ids.size() / P.slot_stride ||
          P.id_win.size() != 4 * (size_t)P.nblocks)
        return err + 1;
      int short_tiles = 0;
      for (int b = 0; b < P.nblocks; b++) {
        const int *w = &P.id_win[4 * (size_t)b];
        if (w[0] < 0) continue;
        short_tiles++;
        for (int pos = P.blk_off[b]; pos < P.blk_off[b + 1]; pos++)
          for (int s = 0; s < 4; s++) {
            const unsigned h = P.ids16[(size_t)pos * 4 + s];
            const int m = s / 2;
            err += w[2 * m] + (int)(h & 0x7fffu) + ((h >> 15) ? w[2 * m + 1] : 0) !=
                   P.ids[(size_t)pos * P.slot_stride + s];
          }
      }
      err += short_tiles != P.short_tiles;
    }
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

bool live_graphs() {
  for (auto g : rt().graphs)
    if (g) return true;
  return false;
}

void refuse_with_live_graphs(const char *what) {
  if (live_graphs())
    fail("%s: captured graphs exist and may use the plans it would replace; destroy them "
         "(op2_graph_destroy) first", what);
}

void drop_plan(Plan *p) {
  Runtime &r = rt();
  for (size_t i = 0; i < r.plans.size(); i++)
    if (r.plans[i].get() == p) {
      if (r.device_ready) (void)hipStreamSynchronize(r.stream);
      p->release(live_graphs() ? &r.retired : nullptr);
      forget_plan_caches(p);
      if (r.last_plan == p) r.last_plan = nullptr;
      r.plans.erase(r.plans.begin() + i);
      return;
    }
}

Plan *find_plan(const PlanRequest &req) {
  for (auto &p : rt().plans)
    if (same_key(*p, req)) return p.get();
  return nullptr;
}

Plan *get_plan(const PlanRequest &req) {
  Runtime &r = rt();
  for (auto &p : r.plans)
    if (same_key(*p, req)) {
      r.last_plan = p.get();
      p->count++;
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
  r.plans.emplace_back(p);
  r.last_plan = p;
  if (OP_diags > 1) {
    printf(" number of blocks       = %d \n", p->nblocks);
    printf(" number of block colors = %d \n", p->ncolors);
    printf(" maximum block size     = %d \n", p->max_block);
    printf(" average thread colors  = %.2f \n", p->mean_thread_colours);
  }
  return p;
}

// ---- plans in the reference's layout (op_plan_core / op_plan_get) ----------

// the reference's key (common/op_lib_core.c:632-645) plus the dataset
// numbering inds, which shapes the plan too
static bool same_ref_key(const RefPlan &p, const char *name, op_set set, int part_size,
                         int nargs, const op_arg *args, int ninds, const int *inds) {
  if (p.name != name || p.view.set != set || p.view.part_size != part_size ||
      p.view.nargs != nargs || p.view.ninds != ninds)
    return false;
  for (int m = 0; m < nargs; m++)
    if (p.dats[m] != args[m].dat || p.maps[m] != args[m].map || p.idxs[m] != args[m].idx ||
        p.accs[m] != args[m].acc || p.inds[m] != (ninds > 0 ? inds[m] : -1))
      return false;
  return true;
}

// OP_plans holds a copy of every view (its arrays point into the RefPlans),
// refreshed whenever a view is added or its call count changes.
static void publish_ref_plans() {
  Runtime &r = rt();
  const int n = (int)r.ref_plans.size();
  if (n > OP_plan_max) {
    op_plan *np = (op_plan *)realloc(OP_plans, sizeof(op_plan) * (size_t)(n + 10));
    if (!np) fail(" op_plan error -- error reallocating memory for OP_plans");
    OP_plans = np;
    OP_plan_max = n + 10;
  }
  for (int i = 0; i < n; i++) OP_plans[i] = r.ref_plans[i]->view;
  OP_plan_index = n;
}

void release_ref_plans() {
  rt().ref_plans.clear();
  free(OP_plans);
  OP_plans = nullptr;
  OP_plan_index = OP_plan_max = 0;
}

// Exact check of a view: every local index maps back to the caller's map
// through ind_maps; per block and dataset the targets are distinct and
// ascending; blocks cover the set in order; blkmap is a permutation grouped
// by colour; no two elements of one thread colour in a block, and no two
// blocks of one colour, share an OP_INC target.
static long long check_ref_plan(const RefPlan &P) {
  const op_plan &v = P.view;
  const int n = v.set->size, nb = v.nblocks, ni = v.ninds;
  long long err = 0;
  int pos = 0;
  for (int b = 0; b < nb; b++) {
    err += P.offset[b] != pos;
    pos += P.nelems[b];
    for (int i = 0; i < ni; i++) {
      const int o = P.ind_offs[i + b * ni], sz = P.ind_sizes[i + b * ni];
      for (int k = o + 1; k < o + sz; k++) err += P.ind_maps[i][k] <= P.ind_maps[i][k - 1];
    }
  }
  err += pos != n;
  for (int m = 0; m < v.nargs; m++) {
    if (P.inds[m] < 0) continue;
    const int i = P.inds[m];
    const int *mp = map_host(P.maps[m]);
    for (int b = 0; b < nb; b++)
      for (int e = P.offset[b]; e < P.offset[b] + P.nelems[b]; e++) {
        const int l = P.loc_maps[m][e];
        if (l < 0 || l >= P.ind_sizes[i + b * ni]) {
          err++;
          continue;
        }
        err += P.ind_maps[i][P.ind_offs[i + b * ni] + l] != mp[(size_t)e * P.maps[m]->dim + P.idxs[m]];
      }
  }
  std::vector<int> seen(nb, 0);
  int k = 0;
  for (int c = 0; c < v.ncolors; c++)
    for (int j = 0; j < P.ncolblk[c]; j++, k++) {
      const int b = P.blkmap[k];
      if (b < 0 || b >= nb || seen[b]++) err++;
    }
  err += k != nb;
  // colour clashes: (dataset, target, colour) claimed by two owners
  std::vector<std::tuple<int, int, int, int>> t;
  auto clashes = [&t]() {
    std::sort(t.begin(), t.end());
    long long e = 0;
    for (size_t a = 1; a < t.size(); a++)
      e += std::get<0>(t[a]) == std::get<0>(t[a - 1]) && std::get<1>(t[a]) == std::get<1>(t[a - 1]) &&
           std::get<2>(t[a]) == std::get<2>(t[a - 1]) && std::get<3>(t[a]) != std::get<3>(t[a - 1]);
    t.clear();
    return e;
  };
  std::vector<int> blkcol(nb, 0);
  k = 0;
  for (int c = 0; c < v.ncolors; c++)
    for (int j = 0; j < P.ncolblk[c]; j++, k++)
      if (k < nb && P.blkmap[k] >= 0 && P.blkmap[k] < nb) blkcol[P.blkmap[k]] = c;
  for (int b = 0; b < nb; b++) {
    for (int e = P.offset[b]; e < P.offset[b] + P.nelems[b]; e++) {
      err += P.thrcol[e] < 0 || P.thrcol[e] >= P.nthrcol[b];
      for (int m = 0; m < v.nargs; m++)
        if (P.inds[m] >= 0 && P.accs[m] == OP_INC)
          t.emplace_back(P.inds[m], map_host(P.maps[m])[(size_t)e * P.maps[m]->dim + P.idxs[m]],
                         P.thrcol[e], e);
    }
    err += clashes();
  }
  for (int b = 0; b < nb; b++)
    for (int e = P.offset[b]; e < P.offset[b] + P.nelems[b]; e++)
      for (int m = 0; m < v.nargs; m++)
        if (P.inds[m] >= 0 && P.accs[m] == OP_INC)
          t.emplace_back(P.inds[m], map_host(P.maps[m])[(size_t)e * P.maps[m]->dim + P.idxs[m]],
                         blkcol[b], b);
  err += clashes();
  return err;
}

// Builds (or finds) the view of op_plan_core (common/op_lib_core.c:627-1089
// defines the layout and its fields): same inputs, same block partition
// (part_size, or with 0 the size that fits the staged data of 64 elements
// per 48 KB), same per-dataset ascending target lists, same greedy first-fit
// colouring in element order, 32 colours per pass (thread colours on the
// OP_INC arguments inside a block, then block colours), blkmap in block
// order within each colour, and the same nshared and float byte model.
op_plan *ref_plan(const char *name, op_set set, int part_size, int nargs, op_arg *args, int ninds,
                  int *inds) {
  Runtime &r = rt();
  if (!name || !set || nargs <= 0 || !args || ninds < 0 || (ninds > 0 && !inds))
    fail("op_plan_core error -- invalid arguments");
  for (auto &p : r.ref_plans)
    if (same_ref_key(*p, name, set, part_size, nargs, args, ninds, inds)) {
      if (OP_diags > 3) printf(" old execution plan for kernel %s\n", name);
      p->view.count++;
      publish_ref_plans();
      return &p->view;
    }
  if (OP_diags > 1) printf(" new execution plan #%zu for kernel %s\n", r.ref_plans.size(), name);

  // each dataset is one dat through one map; every dataset has an argument
  std::vector<int> first(ninds, -1);
  int maxbytes = 0;
  for (int m = 0; m < nargs; m++) {
    const int i = ninds > 0 ? inds[m] : -1;
    if (i < -1 || i >= ninds) fail("op_plan_core error -- inds[%d] = %d out of range", m, i);
    if (i < 0) continue;
    if (args[m].argtype != OP_ARG_DAT || !args[m].map || !args[m].dat || args[m].idx < 0 ||
        args[m].idx >= args[m].map->dim || args[m].map->from != set)
      fail("op_plan_core error -- argument %d of %s is not an indirect dat argument", m, name);
    if (first[i] < 0)
      first[i] = m;
    else if (args[first[i]].dat != args[m].dat || args[first[i]].map != args[m].map)
      fail("op_plan_core error -- indirect dataset %d of %s names two dats or maps", i, name);
    maxbytes += args[m].dat->size;
  }
  for (int i = 0; i < ninds; i++)
    if (first[i] < 0) fail("op_plan_core error -- indirect dataset %d of %s has no argument", i, name);
  int bsize = part_size;
  if (bsize == 0) {
    if (maxbytes == 0) fail("op_plan_core error -- %s has no indirect arguments", name);
    bsize = (48 * 1024 / (64 * maxbytes)) * 64;
  }
  if (bsize <= 0) fail("op_plan_core error -- block size %d for %s", bsize, name);
  const int n = set->size;
  const int nb = (n - 1) / bsize + 1;

  auto P = std::make_unique<RefPlan>();
  P->name = name;
  for (int m = 0; m < nargs; m++) {
    P->dats.push_back(args[m].dat);
    P->maps.push_back(args[m].map);
    P->idxs.push_back(args[m].idx);
    P->accs.push_back(args[m].acc);
    P->inds.push_back(ninds > 0 ? inds[m] : -1);
  }
  inds = P->inds.data();
  auto tgt = [&](int m, int e) {
    return map_host(args[m].map)[(size_t)e * args[m].map->dim + args[m].idx];
  };
  P->offset.resize(nb);
  P->nelems.resize(nb);
  P->nthrcol.resize(nb);
  P->thrcol.assign(n, -1);
  P->ind_offs.resize((size_t)nb * ninds);
  P->ind_sizes.resize((size_t)nb * ninds);
  P->nindirect.assign(ninds, 0);
  P->ind_maps.resize(ninds);
  P->loc_maps.resize(nargs);
  for (int m = 0; m < nargs; m++)
    if (inds[m] >= 0) P->loc_maps[m].assign(n, 0);
  // per dataset: a colour mask (or a local index) per target element
  std::vector<std::vector<uint32_t>> work(ninds);
  for (int i = 0; i < ninds; i++) work[i].assign(args[first[i]].map->to->size, 0);

  std::vector<int> list;
  double total_colours = 0;
  for (int b = 0; b < nb; b++) {
    const int e0 = b * bsize, bs = std::min(bsize, n - e0);
    P->offset[b] = e0;
    P->nelems[b] = bs;
    for (int i = 0; i < ninds; i++) {
      list.clear();
      for (int m = 0; m < nargs; m++)
        if (inds[m] == i)
          for (int e = e0; e < e0 + bs; e++) list.push_back(tgt(m, e));
      std::sort(list.begin(), list.end());
      list.erase(std::unique(list.begin(), list.end()), list.end());
      if (list.size() > 32767)
        fail("op_plan_core error -- block of %s stages %zu elements of dataset %d (16-bit local "
             "indices); reduce the part size", name, list.size(), i);
      const int o = P->nindirect[i];
      for (size_t k = 0; k < list.size(); k++) {
        P->ind_maps[i].push_back(list[k]);
        work[i][list[k]] = (uint32_t)k;
      }
      P->nindirect[i] += (int)list.size();
      P->ind_offs[i + (size_t)b * ninds] = o;
      P->ind_sizes[i + (size_t)b * ninds] = (int)list.size();
      for (int m = 0; m < nargs; m++)
        if (inds[m] == i)
          for (int e = e0; e < e0 + bs; e++) P->loc_maps[m][e] = (short)work[i][tgt(m, e)];
    }
    // thread colours
    int base = 0, ncol = 0;
    for (bool again = true; again; base += 32) {
      again = false;
      for (int m = 0; m < nargs; m++)
        if (inds[m] >= 0)
          for (int e = e0; e < e0 + bs; e++) work[inds[m]][tgt(m, e)] = 0;
      for (int e = e0; e < e0 + bs; e++) {
        if (P->thrcol[e] >= 0) continue;
        uint32_t mask = 0;
        for (int m = 0; m < nargs; m++)
          if (inds[m] >= 0 && args[m].acc == OP_INC) mask |= work[inds[m]][tgt(m, e)];
        if (mask == 0xffffffffu) {
          again = true;
          continue;
        }
        const int c = __builtin_ctz(~mask);
        P->thrcol[e] = base + c;
        ncol = std::max(ncol, base + c + 1);
        for (int m = 0; m < nargs; m++)
          if (inds[m] >= 0 && args[m].acc == OP_INC) work[inds[m]][tgt(m, e)] |= 1u << c;
      }
    }
    P->nthrcol[b] = ncol;
    total_colours += ncol;
  }

  // block colours
  std::vector<int> blkcol(nb, -1);
  int ncolors = 0;
  for (int base = 0, again = 1; again; base += 32) {
    again = 0;
    for (int i = 0; i < ninds; i++) std::fill(work[i].begin(), work[i].end(), 0u);
    for (int b = 0; b < nb; b++) {
      if (blkcol[b] >= 0) continue;
      const int e0 = P->offset[b], e1 = e0 + P->nelems[b];
      uint32_t mask = 0;
      for (int m = 0; m < nargs; m++)
        if (inds[m] >= 0 && args[m].acc == OP_INC)
          for (int e = e0; e < e1; e++) mask |= work[inds[m]][tgt(m, e)];
      if (mask == 0xffffffffu) {
        again = 1;
        continue;
      }
      const int c = __builtin_ctz(~mask);
      blkcol[b] = base + c;
      ncolors = std::max(ncolors, base + c + 1);
      for (int m = 0; m < nargs; m++)
        if (inds[m] >= 0 && args[m].acc == OP_INC)
          for (int e = e0; e < e1; e++) work[inds[m]][tgt(m, e)] |= 1u << c;
    }
  }
  P->ncolblk.assign(std::max(ncolors, 1), 0);
  for (int b = 0; b < nb; b++) P->ncolblk[blkcol[b]]++;
  {
    std::vector<int> start(ncolors + 1, 0);
    for (int c = 0; c < ncolors; c++) start[c + 1] = start[c] + P->ncolblk[c];
    P->blkmap.resize(nb);
    for (int b = 0; b < nb; b++) P->blkmap[start[blkcol[b]]++] = b;
  }

  // shared memory per block and the transfer byte model, accumulated in
  // float in the order of the reference's loops
  int nshared = 0;
  float transfer = 0, transfer2 = 0;
  const int cl = OP_cache_line_size > 0 ? OP_cache_line_size : 128;
  for (int b = 0; b < nb; b++) {
    int bytes = 0;
    for (int i = 0; i < ninds; i++)
      bytes += ROUND_UP(P->ind_sizes[i + (size_t)b * ninds] * args[first[i]].dat->size);
    nshared = std::max(nshared, bytes);
  }
  for (int b = 0; b < nb; b++) {
    for (int m = 0; m < nargs; m++) {
      if (inds[m] < 0) {
        const float fac = args[m].acc == OP_READ ? 1.0f : 2.0f;
        if (args[m].dat) {
          transfer += fac * P->nelems[b] * args[m].dat->size;
          transfer2 += fac * P->nelems[b] * args[m].dat->size;
        }
      } else {
        transfer += (float)(P->nelems[b] * sizeof(short));
        transfer2 += (float)(P->nelems[b] * sizeof(short));
      }
    }
    for (int i = 0; i < ninds; i++) {
      const op_arg &a = args[first[i]];
      const long long sz = a.dat->size;
      float fac = a.acc == OP_READ ? 1.0f : 2.0f;
      const int o = P->ind_offs[i + (size_t)b * ninds], cnt = P->ind_sizes[i + (size_t)b * ninds];
      transfer += fac * cnt * a.dat->size;
      // cache lines the block's targets occupy
      long long l_old = -1;
      for (int e = o; e < o + cnt; e++) {
        const long long t = P->ind_maps[i][e];
        long long l_new = (t * sz) / cl;
        if (l_new > l_old) transfer2 += fac * cl;
        l_old = l_new;
        l_new = ((t + 1) * sz - 1) / cl;
        transfer2 += fac * (l_new - l_old) * cl;
        l_old = l_new;
      }
      // the mapping that loads / stores the staged data
      fac = a.acc == OP_RW ? 2.0f : 1.0f;
      transfer += fac * cnt * sizeof(int);
      transfer2 += fac * cnt * sizeof(int);
    }
  }

  for (int i = 0; i < ninds; i++) P->ind_map_ptrs.push_back(P->ind_maps[i].data());
  for (int m = 0; m < nargs; m++)
    P->loc_map_ptrs.push_back(P->loc_maps[m].empty() ? nullptr : P->loc_maps[m].data());
  op_plan &v = P->view;
  v.name = P->name.c_str();
  v.set = set;
  v.nargs = nargs;
  v.ninds = ninds;
  v.part_size = part_size;
  v.maps = P->maps.data();
  v.dats = P->dats.data();
  v.idxs = P->idxs.data();
  v.accs = P->accs.data();
  v.nthrcol = P->nthrcol.data();
  v.thrcol = P->thrcol.data();
  v.offset = P->offset.data();
  v.ind_maps = P->ind_map_ptrs.data();
  v.ind_offs = P->ind_offs.data();
  v.ind_sizes = P->ind_sizes.data();
  v.nindirect = P->nindirect.data();
  v.loc_maps = P->loc_map_ptrs.data();
  v.nblocks = nb;
  v.nelems = P->nelems.data();
  v.ncolors = ncolors;
  v.ncolblk = P->ncolblk.data();
  v.blkmap = P->blkmap.data();
  v.nshared = nshared;
  v.transfer = transfer;
  v.transfer2 = transfer2;
  v.count = 1;
  if (OP_diags > 1) {
    printf(" number of blocks       = %d \n", nb);
    printf(" number of block colors = %d \n", ncolors);
    printf(" maximum block size     = %d \n", bsize);
    printf(" average thread colors  = %.2f \n", total_colours / nb);
    printf(" shared memory required = %.2f KB \n", nshared / 1024.0f);
    printf(" data transfer (used)   = %.2f MB \n", transfer / (1024.0f * 1024.0f));
    printf(" data transfer (total)  = %.2f MB \n\n", transfer2 / (1024.0f * 1024.0f));
  }
  const long long errs = check_ref_plan(*P);
  if (errs) fail("op_plan_core error -- plan for %s fails its check (%lld violations)", name, errs);
  r.ref_plans.push_back(std::move(P));
  publish_ref_plans();
  return &r.ref_plans.back()->view;
}

}  // namespace op2

using namespace op2;

extern "C" {

op_plan *op_plan_core(char const *name, op_set set, int part_size, int nargs, op_arg *args,
                      int ninds, int *inds) {
  try {
    return ref_plan(name, set, part_size, nargs, args, ninds, inds);
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
    out->cached_plans = (int)rt().plans.size();
    out->short_ids = p->ids16.empty() ? 0 : p->short_tiles;
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
