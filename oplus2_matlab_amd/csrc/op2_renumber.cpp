// Renumbering of the sets for locality: the op_partition the reference plans
// (doc/user.tex:351-356, doc/dev.tex:191-206: "reorder/renumber all sets to
// improve data locality, so that neighboring nodes or edges have indices,
// and therefore memory locations, which are close to each other").
//
// The runtime renumbers internally: dats and maps on the device (and the
// host copies of the maps that plans are built from) use the new numbers;
// the user's arrays keep theirs, and op_fetch_data puts the data back in the
// user's order.  The ordering:
//   * the target set P of the base map (cells, for pecell) is ordered by
//     recursive bisection of its graph (two targets are adjacent when one
//     element of the map references both): each connected component in
//     turn (components keep their original order), split in two halves of
//     its breadth-first order from a pseudo-peripheral vertex, recursively.
//     The split points are multiples of a power of two, so every aligned run
//     of 2^k elements (k <= 9) is one subtree: a compact piece of the mesh;
//   * a set with a map into an ordered set (edges, bedges -> cells) is sorted
//     by the smallest new number its elements reference (stable);
//   * a set that an ordered set maps into (nodes, from cells) is numbered by
//     first touch in the new order of the ordered set;
//   * repeated over the maps until no more sets are reached; sets that are
//     never reached keep their numbers.
#include <sys/time.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "op2_internal.h"

namespace op2 {

namespace {

constexpr int kAlignMax = 512;  // largest aligned run that stays one subtree

struct Graph {
  std::vector<int> off, adj;
};

Graph target_graph(op_map m) {
  const int n = m->from->size, nt = m->to->size, dim = m->dim;
  const int *mm = map_host(m);
  Graph g;
  g.off.assign(nt + 1, 0);
  for (int e = 0; e < n; e++)
    for (int a = 0; a < dim; a++)
      for (int b = 0; b < dim; b++) {
        const int u = mm[(size_t)e * dim + a], v = mm[(size_t)e * dim + b];
        if (a != b && u != v) g.off[u + 1]++;
      }
  for (int t = 0; t < nt; t++) g.off[t + 1] += g.off[t];
  g.adj.resize(g.off[nt]);
  std::vector<int> fill(g.off.begin(), g.off.end() - 1);
  for (int e = 0; e < n; e++)
    for (int a = 0; a < dim; a++)
      for (int b = 0; b < dim; b++) {
        const int u = mm[(size_t)e * dim + a], v = mm[(size_t)e * dim + b];
        if (a != b && u != v) g.adj[fill[u]++] = v;
      }
  return g;
}

// Recursive bisection of the vertices vs[lo, hi) (all carrying label `lab`
// in part[]); writes the final order back into vs[lo, hi).
// The two halves of a split touch disjoint vertices, so the upper levels
// run their halves on separate threads (sharing part[] and dist[]).
struct Bisector {
  const Graph &g;
  std::vector<int> &part, &dist;
  std::atomic<int> &next_label;
  std::vector<int> queue;

  Bisector(const Graph &gr, std::vector<int> &p, std::vector<int> &d, std::atomic<int> &nl)
      : g(gr), part(p), dist(d), next_label(nl) {}

  // breadth-first order of the vertices labelled `lab` from `seed`,
  // continuing from the next unvisited one of vs when a piece is exhausted;
  // returns the last vertex reached from `seed`
  int bfs(int *vs, int n, int lab, int seed, std::vector<int> &out) {
    out.clear();
    for (int i = 0; i < n; i++) dist[vs[i]] = -1;
    int last = seed, k = 0;
    auto run = [&](int s) {
      size_t head = out.size();
      dist[s] = 0;
      out.push_back(s);
      while (head < out.size()) {
        const int u = out[head++];
        for (int j = g.off[u]; j < g.off[u + 1]; j++) {
          const int v = g.adj[j];
          if (part[v] != lab || dist[v] >= 0) continue;
          dist[v] = dist[u] + 1;
          out.push_back(v);
        }
      }
    };
    run(seed);
    last = out.back();
    while ((int)out.size() < n) {
      while (dist[vs[k]] >= 0) k++;
      run(vs[k]);
    }
    return last;
  }

  void order(int *vs, int n, int lab, int align, int threads = 1) {
    if (n <= 1) return;
    // pseudo-peripheral start: the far end of a sweep from the first vertex
    int seed = *std::min_element(vs, vs + n);
    seed = bfs(vs, n, lab, seed, queue);
    bfs(vs, n, lab, seed, queue);
    std::copy(queue.begin(), queue.end(), vs);
    if (n <= 32) return;
    // split at a multiple of the largest power of two not above n / 2 (at
    // most kAlignMax): the halves then start on aligned positions
    int gsz = 1;
    while (gsz * 2 <= n / 2 && gsz * 2 <= std::min(align, kAlignMax)) gsz *= 2;
    int h = (int)(((long long)n / 2 + gsz / 2) / gsz) * gsz;
    h = std::max(gsz, std::min(h, n - 1));
    const int la = next_label++, lb = next_label++;
    for (int i = 0; i < h; i++) part[vs[i]] = la;
    for (int i = h; i < n; i++) part[vs[i]] = lb;
    if (threads > 1 && n > 65536) {
      std::thread t([&] {
        Bisector other(g, part, dist, next_label);
        other.order(vs + h, n - h, lb, gsz, threads / 2);
      });
      order(vs, h, la, gsz, threads - threads / 2);
      t.join();
    } else {
      order(vs, h, la, gsz);
      order(vs + h, n - h, lb, gsz);
    }
  }
};

// new_of_old of the base map's target set; seg gets the start of every
// connected component in the new order
std::vector<int> bisection_order(op_map base, std::vector<int> &seg) {
  const Graph g = target_graph(base);
  const int nt = base->to->size;
  // connected components, in order of their smallest element
  std::vector<int> comp(nt, -1), verts;
  verts.reserve(nt);
  std::vector<int> comp_off{0};
  for (int s = 0; s < nt; s++) {
    if (comp[s] >= 0) continue;
    const int c = (int)comp_off.size() - 1;
    size_t head = verts.size();
    comp[s] = c;
    verts.push_back(s);
    while (head < verts.size()) {
      const int u = verts[head++];
      for (int j = g.off[u]; j < g.off[u + 1]; j++)
        if (comp[g.adj[j]] < 0) {
          comp[g.adj[j]] = c;
          verts.push_back(g.adj[j]);
        }
    }
    comp_off.push_back((int)verts.size());
  }
  // the bisection runs on the graph relabelled in this breadth-first order,
  // where a vertex's neighbours sit close to it in memory
  std::vector<int> lab_of(nt);
  for (int i = 0; i < nt; i++) lab_of[verts[i]] = i;
  Graph h;
  h.off.assign(nt + 1, 0);
  h.adj.resize(g.adj.size());
  for (int i = 0; i < nt; i++) {
    const int u = verts[i];
    h.off[i + 1] = h.off[i] + (g.off[u + 1] - g.off[u]);
    for (int j = g.off[u], k = h.off[i]; j < g.off[u + 1]; j++, k++) h.adj[k] = lab_of[g.adj[j]];
  }
  std::vector<int> part(nt), order(nt);
  for (size_t c = 0; c + 1 < comp_off.size(); c++)
    for (int i = comp_off[c]; i < comp_off[c + 1]; i++) part[i] = (int)c + 1;  // labels from 1
  std::iota(order.begin(), order.end(), 0);
  std::vector<int> dist(nt, -1);
  std::atomic<int> next_label((int)comp_off.size() + 1);
  Bisector bis(h, part, dist, next_label);
  const int threads = (int)std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
  for (size_t c = 0; c + 1 < comp_off.size(); c++)
    bis.order(order.data() + comp_off[c], comp_off[c + 1] - comp_off[c], (int)c + 1, kAlignMax,
              threads);
  std::vector<int> new_of_old(nt);
  for (int i = 0; i < nt; i++) new_of_old[verts[order[i]]] = i;
  seg.assign(comp_off.begin(), comp_off.end() - 1);
  return new_of_old;
}

double now() {
  timeval t;
  gettimeofday(&t, nullptr);
  return t.tv_sec + 1e-6 * t.tv_usec;
}

bool is_bijection(const std::vector<int> &p) {
  std::vector<char> seen(p.size(), 0);
  for (int v : p) {
    if (v < 0 || v >= (int)p.size() || seen[v]) return false;
    seen[v] = 1;
  }
  return true;
}

}  // namespace

const std::vector<int> *set_new_of_old(op_set s) {
  auto it = rt().set_perm.find(s);
  return it != rt().set_perm.end() ? &it->second : nullptr;
}

// internal copy of map m under the current set permutations (none needed:
// the entry is removed)
void renumber_map(op_map m) {
  Runtime &r = rt();
  const std::vector<int> *pf = set_new_of_old(m->from), *pt = set_new_of_old(m->to);
  r.map_internal.erase(m);
  m->map_i = nullptr;
  if (m->map_d) {
    if (r.device_ready) OP2_HIP_CHECK(hipStreamSynchronize(r.stream));
    dev_free(m->map_d);
    m->map_d = nullptr;
  }
  if (!pf && !pt) return;
  const int n = m->from->size, dim = m->dim;
  std::vector<int> out((size_t)n * dim);
  for (int e = 0; e < n; e++) {
    const int ne = pf ? (*pf)[e] : e;
    for (int k = 0; k < dim; k++) {
      const int v = m->map[(size_t)e * dim + k];
      out[(size_t)ne * dim + k] = pt ? (*pt)[v] : v;
    }
  }
  std::vector<int> &kept = r.map_internal[m];
  kept = std::move(out);
  m->map_i = kept.data();
}

// bytes of a dat, in the user's order -> internal order
void permute_to_internal(op_dat d, const char *user, char *internal) {
  const std::vector<int> *p = set_new_of_old(d->set);
  const size_t es = d->size, n = d->set->size;
  if (!p) {
    memcpy(internal, user, es * n);
    return;
  }
  for (size_t e = 0; e < n; e++) memcpy(internal + (size_t)(*p)[e] * es, user + e * es, es);
}

void permute_to_user(op_dat d, const char *internal, char *user) {
  const std::vector<int> *p = set_new_of_old(d->set);
  const size_t es = d->size, n = d->set->size;
  if (!p) {
    memcpy(user, internal, es * n);
    return;
  }
  for (size_t e = 0; e < n; e++) memcpy(user + e * es, internal + (size_t)(*p)[e] * es, es);
}

void renumber(op_map base) {
  Runtime &r = rt();
  if (!base) fail("op2_renumber: invalid map");
  if (r.capturing) fail("op2_renumber: not during graph capture");
  if (!r.graphs.empty())
    fail("op2_renumber: renumber before capturing graphs (captured graphs hold the old layout)");
  if (!r.set_perm.empty()) fail("op2_renumber: the sets are already renumbered");
  for (auto &o : r.online)
    if (o.second.started && !o.second.done)
      fail("op2_renumber: an online autotuning is unfinished");

  // orders: the base map's target set, then along the maps
  const double t0 = now();
  std::map<op_set, std::vector<int>> perm;
  std::vector<int> seg;
  perm[base->to] = bisection_order(base, seg);
  const double t1 = now();
  for (bool grew = true; grew;) {
    grew = false;
    // the base map first, then the others in declaration order
    std::vector<op_map> maps{base};
    for (int i = 0; i < OP_map_index; i++)
      if (OP_map_list[i] != base) maps.push_back(OP_map_list[i]);
    for (op_map m : maps) {
      const bool hf = perm.count(m->from) > 0, ht = perm.count(m->to) > 0;
      if (hf == ht) continue;
      const int n = m->from->size, dim = m->dim;
      const int *mm = m->map;
      if (ht) {
        // from-set sorted by the smallest new target it references
        // (a counting sort: keys are target numbers)
        const std::vector<int> &pt = perm[m->to];
        const int nt = m->to->size;
        std::vector<int> key(n), cnt(nt + 1, 0);
        for (int e = 0; e < n; e++) {
          int k = pt[mm[(size_t)e * dim]];
          for (int a = 1; a < dim; a++) k = std::min(k, pt[mm[(size_t)e * dim + a]]);
          key[e] = k;
          cnt[k + 1]++;
        }
        for (int t = 0; t < nt; t++) cnt[t + 1] += cnt[t];
        std::vector<int> nof(n);
        for (int e = 0; e < n; e++) nof[e] = cnt[key[e]]++;
        perm[m->from] = std::move(nof);
      } else {
        // to-set numbered by first touch in the from-set's new order
        const std::vector<int> &pf = perm[m->from];
        std::vector<int> old_of_new(n);
        for (int e = 0; e < n; e++) old_of_new[pf[e]] = e;
        const int nt = m->to->size;
        std::vector<int> nof(nt, -1);
        int next = 0;
        for (int i = 0; i < n; i++)
          for (int a = 0; a < dim; a++) {
            const int t = mm[(size_t)old_of_new[i] * dim + a];
            if (nof[t] < 0) nof[t] = next++;
          }
        for (int t = 0; t < nt; t++)
          if (nof[t] < 0) nof[t] = next++;
        perm[m->to] = std::move(nof);
      }
      grew = true;
    }
  }
  for (auto &kv : perm)
    if (!is_bijection(kv.second)) fail("op2_renumber: internal error (not a permutation)");
  const double t2 = now();

  // plans and cached layouts refer to the old numbers
  if (r.device_ready) OP2_HIP_CHECK(hipStreamSynchronize(r.stream));
  while (!r.plans.empty()) drop_plan(r.plans.back().get());
  release_exec_caches();

  // drop identity permutations
  for (auto &kv : perm) {
    bool id = true;
    for (size_t i = 0; i < kv.second.size() && id; i++) id = kv.second[i] == (int)i;
    if (!id) r.set_perm[kv.first] = std::move(kv.second);
  }
  if (r.set_perm.count(base->to)) r.set_segments[base->to] = seg;
  const double t3 = now();
  for (int i = 0; i < OP_map_index; i++) renumber_map(OP_map_list[i]);
  const double t4 = now();
  // device data of every dat on a renumbered set
  for (int i = 0; i < OP_dat_index; i++) {
    op_dat d = OP_dat_list[i];
    if (!d->data_d || !set_new_of_old(d->set)) continue;
    const size_t bytes = (size_t)d->size * d->set->size;
    std::vector<char> user(bytes), internal(bytes);
    d2h(user.data(), d->data_d, bytes);
    permute_to_internal(d, user.data(), internal.data());
    h2d(d->data_d, internal.data(), bytes);
  }
  forget_zero_dats();
  if (OP_diags > 1)
    printf(" renumbering: bisection %.2f s, other sets %.2f s, plans %.2f s, maps %.2f s, "
           "dats %.2f s\n", t1 - t0, t2 - t1, t3 - t2, t4 - t3, now() - t4);
}

// The first loop with an indirect OP_INC argument names the mesh's main
// connectivity (airfoil's res_calc: edges to cells; jac: edges to nodes), so
// its map is the base.  Loops before it ran on the declared numbering; their
// plans are dropped with the others.  Nothing happens once graphs exist or
// the program has renumbered itself.
void auto_renumber(op_set set, int nargs, const op_arg *args) {
  Runtime &r = rt();
  if (r.capturing || !r.graphs.empty() || !r.set_perm.empty()) {
    r.auto_renumber = false;
    return;
  }
  for (auto &o : r.online)
    if (o.second.started && !o.second.done) return;  // after the tuning
  for (int m = 0; m < nargs; m++) {
    const op_arg &a = args[m];
    if (a.argtype != OP_ARG_DAT || a.acc != OP_INC || !a.map || a.map->from != set ||
        a.map->to == set)
      continue;
    r.auto_renumber = false;
    if (OP_diags > 1) printf(" renumbering on map %s (automatic)\n", a.map->name);
    renumber(a.map);
    return;
  }
}

}  // namespace op2
