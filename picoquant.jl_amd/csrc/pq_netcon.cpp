// Optimal pairwise contraction order for small tensor networks: the search
// PicoQuant's netcon (src/layer2/netcon_contraction.jl:14-41) obtains from
// TensorOperations.optimaltree.
//
// Cost of a tree = sum over its pairwise contractions of the number of
// complex multiply-adds, prod(dims of the indices of A union B) -- the MAC
// count of the reference's Metrics.flops (src/backends.jl:37-38).
//
// Search: breadth-first over the number of tensors in a subnetwork with a
// cost cap, after Pfeifer, Haegeman and Verstraete, "Faster identification
// of optimal contraction sequences for tensor networks", PRE 90, 033315
// (2014).  Level c holds every subnetwork of c tensors reachable below the
// cap, each with its cheapest tree; level c is built from pairs of levels
// d and c - d that share at least one index.  When the whole network is not
// reached the cap is raised to the cheapest cost the round rejected (at
// least doubled) and the search repeats.  Outer products are taken only
// between connected components, smallest first.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "pq_hip.h"

namespace {

constexpr int MAX_TENSORS = 64;

struct Net {
  int n = 0;
  int words = 1;                            // 64-bit words per index bitset
  std::vector<double> dim;                  // per index id: extent
  std::vector<std::vector<uint64_t>> open;  // per tensor: its index bitset
};

struct Entry {
  double cost = 0;     // total MACs of the best tree for this subnetwork
  uint64_t left = 0;   // best split (0 for a single tensor)
  uint64_t right = 0;
  int64_t open_at = 0; // offset of the subnetwork's open-index bitset in the pool
};

struct Search {
  const Net& net;
  std::unordered_map<uint64_t, Entry> best;
  std::vector<std::vector<uint64_t>> levels;
  std::vector<uint64_t> pool;
  int64_t work = 0, max_work;
  explicit Search(const Net& n, int64_t mw) : net(n), max_work(mw) {}

  const uint64_t* open_of(const Entry& e) const { return pool.data() + e.open_at; }

  // MACs of contracting two subnetworks (product of the extents of their
  // open indices, shared ones once), and whether they share an index.
  // Indices join at most two tensors, so the union's open set is the
  // symmetric difference of the parts' open sets.
  double pair_macs(const uint64_t* a, const uint64_t* b, bool& shared) const {
    double s = 1;
    shared = false;
    for (int w = 0; w < net.words; ++w) {
      uint64_t u = a[w] | b[w];
      if (a[w] & b[w]) shared = true;
      while (u) {
        const int bit = __builtin_ctzll(u);
        u &= u - 1;
        s *= net.dim[(size_t)w * 64 + bit];
      }
    }
    return s;
  }

  // One breadth-first pass with cap mu over the tensors in `members`.
  // Returns the cheapest rejected cost (infinity if nothing was rejected).
  double pass(const std::vector<int>& members, double mu, bool& over) {
    best.clear();
    pool.clear();
    const int m = (int)members.size();
    levels.assign(m + 1, {});
    for (int t : members) {
      Entry e;
      e.open_at = (int64_t)pool.size();
      pool.insert(pool.end(), net.open[t].begin(), net.open[t].end());
      best[1ull << t] = e;
      levels[1].push_back(1ull << t);
    }
    double rejected = INFINITY;
    std::vector<uint64_t> tmp(net.words);
    for (int c = 2; c <= m; ++c) {
      for (int d = 1; d <= c / 2; ++d) {
        const auto& L1 = levels[d];
        const auto& L2 = levels[c - d];
        for (size_t i = 0; i < L1.size(); ++i) {
          const uint64_t s1 = L1[i];
          const Entry e1 = best[s1];
          for (size_t j = (d == c - d) ? i + 1 : 0; j < L2.size(); ++j) {
            const uint64_t s2 = L2[j];
            if (s1 & s2) continue;
            if (++work > max_work) {
              over = true;
              return rejected;
            }
            const Entry& e2 = best[s2];
            bool shared;
            const double macs = pair_macs(open_of(e1), open_of(e2), shared);
            if (!shared) continue;
            const double cost = e1.cost + e2.cost + macs;
            if (cost > mu) {
              rejected = std::min(rejected, cost);
              continue;
            }
            const uint64_t s = s1 | s2;
            auto it = best.find(s);
            if (it == best.end()) {
              const uint64_t* a = open_of(e1);
              const uint64_t* b = open_of(best[s2]);
              for (int w = 0; w < net.words; ++w) tmp[w] = a[w] ^ b[w];
              Entry e;
              e.cost = cost;
              e.left = s1;
              e.right = s2;
              e.open_at = (int64_t)pool.size();
              pool.insert(pool.end(), tmp.begin(), tmp.end());
              best[s] = e;
              levels[c].push_back(s);
            } else if (cost < it->second.cost) {
              it->second.cost = cost;
              it->second.left = s1;
              it->second.right = s2;
            }
          }
        }
      }
    }
    return rejected;
  }
};

// Post-order emission of the tree of subnetwork s; returns its id (tensor
// index for leaves, n + step for intermediates).
int emit(const std::unordered_map<uint64_t, Entry>& best, uint64_t s, int n, std::vector<int32_t>& plan) {
  const Entry& e = best.at(s);
  if (e.left == 0) return __builtin_ctzll(s);
  const int a = emit(best, e.left, n, plan);
  const int b = emit(best, e.right, n, plan);
  plan.push_back(a);
  plan.push_back(b);
  return n + (int)plan.size() / 2 - 1;
}

}  // namespace

extern "C" int pq_netcon(int n_tensors, const int32_t* rank, const int64_t* labels, const int64_t* dims,
                         int64_t max_work, int32_t* plan_out, double* cost_out) {
  if (n_tensors < 1 || n_tensors > MAX_TENSORS || !rank || !plan_out || !cost_out) return PQ_ERR_ARG;
  // Dense index ids; each label must join at most two tensors with one extent.
  std::unordered_map<int64_t, int> id;
  std::vector<int> seen;
  std::vector<double> dimv;
  std::vector<std::vector<int>> tidx(n_tensors);
  int64_t pos = 0;
  for (int t = 0; t < n_tensors; ++t) {
    if (rank[t] < 0) return PQ_ERR_ARG;
    for (int k = 0; k < rank[t]; ++k, ++pos) {
      if (dims[pos] < 1) return PQ_ERR_ARG;
      auto it = id.find(labels[pos]);
      int x;
      if (it == id.end()) {
        x = (int)dimv.size();
        id[labels[pos]] = x;
        dimv.push_back((double)dims[pos]);
        seen.push_back(0);
      } else {
        x = it->second;
        if ((double)dims[pos] != dimv[x]) return PQ_ERR_ARG;
      }
      if (++seen[x] > 2) return PQ_ERR_ARG;
      tidx[t].push_back(x);
    }
  }
  for (int t = 0; t < n_tensors; ++t) {
    std::vector<int> s = tidx[t];
    std::sort(s.begin(), s.end());
    if (std::adjacent_find(s.begin(), s.end()) != s.end()) return PQ_ERR_ARG;   // trace within a tensor
  }
  Net net;
  net.n = n_tensors;
  net.words = std::max<int>(1, ((int)dimv.size() + 63) / 64);
  net.dim = dimv;
  net.dim.resize((size_t)net.words * 64, 1.0);
  net.open.assign(n_tensors, std::vector<uint64_t>(net.words, 0));
  for (int t = 0; t < n_tensors; ++t)
    for (int x : tidx[t]) net.open[t][x / 64] |= 1ull << (x % 64);
  // Connected components (tensors sharing an index).
  std::vector<int> comp(n_tensors, -1);
  int ncomp = 0;
  for (int t = 0; t < n_tensors; ++t) {
    if (comp[t] >= 0) continue;
    std::vector<int> stack{t};
    comp[t] = ncomp;
    while (!stack.empty()) {
      const int u = stack.back();
      stack.pop_back();
      for (int v = 0; v < n_tensors; ++v) {
        if (comp[v] >= 0) continue;
        bool share = false;
        for (int w = 0; w < net.words && !share; ++w) share = (net.open[u][w] & net.open[v][w]) != 0;
        if (share) {
          comp[v] = ncomp;
          stack.push_back(v);
        }
      }
    }
    ++ncomp;
  }
  std::vector<int32_t> plan;
  double total = 0;
  struct Part { int id; double size; };
  std::vector<Part> parts;
  Search srch(net, max_work > 0 ? max_work : (int64_t)1 << 32);
  for (int cc = 0; cc < ncomp; ++cc) {
    std::vector<int> members;
    uint64_t full = 0;
    for (int t = 0; t < n_tensors; ++t)
      if (comp[t] == cc) {
        members.push_back(t);
        full |= 1ull << t;
      }
    double mu = 1;
    for (int t : members) {
      bool dummy;
      mu = std::max(mu, srch.pair_macs(net.open[t].data(), net.open[t].data(), dummy));
    }
    for (;;) {
      bool over = false;
      const double rejected = srch.pass(members, mu, over);
      if (over) return PQ_ERR_UNSUPPORTED;
      if (srch.best.count(full)) break;
      if (!std::isfinite(rejected)) return PQ_ERR_UNSUPPORTED;
      mu = std::max(2 * mu, rejected);
    }
    total += srch.best[full].cost;
    const int root = emit(srch.best, full, n_tensors, plan);
    bool dummy;
    const Entry& e = srch.best[full];
    const uint64_t* o = srch.pool.data() + e.open_at;
    parts.push_back({root, srch.pair_macs(o, o, dummy)});
  }
  // Disconnected components: outer products, smallest first.
  while (parts.size() > 1) {
    std::stable_sort(parts.begin(), parts.end(), [](const Part& a, const Part& b) { return a.size < b.size; });
    plan.push_back(parts[0].id);
    plan.push_back(parts[1].id);
    const double sz = parts[0].size * parts[1].size;
    total += sz;
    Part p{n_tensors + (int)plan.size() / 2 - 1, sz};
    parts.erase(parts.begin(), parts.begin() + 2);
    parts.push_back(p);
  }
  std::memcpy(plan_out, plan.data(), plan.size() * sizeof(int32_t));
  *cost_out = total;
  return PQ_OK;
}
