// mpi_exp -- exponential contraction over P MPI ranks: the host-CPU baseline.
// TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg runs it (mpirun -np P)
// on the GPU box's host cores beside the GPU measurement; the product
// (libkacomp) never calls it.
//
// A restatement, in its own code, of the reference's distributed default path
// (app/exp_contraction.cpp:75-113, ExponentialContraction::FindComponents,
// lib/components/exponential_contraction.h:60-64,218-317):
//   * the input ids are split over the ranks in blocks of ceil(n/P)
//     (StaticGraph's block distribution, static_graph.h:224-271); degree-0
//     vertices are inactive (graph_io.h:378-381) and rng_offset is the global
//     count of the active ones (exponential_contraction.h:61-64);
//   * local phase (FindLocalComponents :392-411, CAGBuilder::
//     BuildLocalComponentGraph cag_builder.h:160-230): the components of the
//     edges inside a rank's block are contracted, except the vertices with a
//     neighbour on another rank (interface vertices), which stay themselves;
//   * decomposition (RunContractionLP :439-628): deviates of
//     exponential_distribution(beta) over minstd_rand0 seeded with
//     gid + seed + iteration*(rng_offset + seed) (:467-471), then in-place
//     sweeps over the local rows with the ghost keys exchanged after every
//     sweep (only the changed ones, as SendAndReceiveGhostVertices sends only
//     changed payloads) and an all-reduce of the change flag (:558-563);
//   * contraction (DynamicContraction::ExponentialContraction,
//     dynamic_contraction.h:62-136): every cluster edge (label(u), label(w))
//     goes to the rank owning label(u), which deduplicates it; the centres
//     with an edge form the next level, numbered in label order;
//   * cut-off (ComputeSequentialLimit :1287-1292, relative limit m/P * seq
//     or a fixed one); then the remaining graph is gathered on rank 0 and its
//     components found there (RunSequentialCC :908-922, AllReduce);
//   * undo (UndoContraction dynamic_contraction.h:568-658) level by level, and
//     the labels of the input vertices (ApplyToLocalComponents :861-906),
//     made canonical (minimum id per component) so that they can be compared
//     with union-find and with the GPU's labels.
// A level whose first sweep changes no key contracts nothing; it is skipped
// (same graph, next iteration) instead of being rebuilt as the identity.
//
// Timing: MPI_Wtime around the FindComponents equivalent on every rank, max
// over the ranks (exp_contraction.cpp:103-113), mean and deviation over -i
// iterations with seed = seed + i + 1000 (:75-77).
//
// Input: a CSR file "<n u64><2m u64><offsets u64[n+1]><targets u32[2m]>"
// (both directions of every edge).  Output: the reference's GLOBAL RESULT
// line with levels and sweeps added; -labels <file> writes the canonical
// labels of the last iteration (n u32) from rank 0.
#include <mpi.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

namespace {

using u32 = uint32_t;
using u64 = uint64_t;
constexpr u32 kNone = 0xFFFFFFFFu;
constexpr u64 kHop = 1ull << 32;

int g_rank = 0, g_size = 1;
bool g_verbose = false;

// -v: rank 0 prints the time since the previous mark
void mark(const char *what) {
  static double last = 0;
  if (!g_verbose) return;
  const double t = MPI_Wtime();
  if (g_rank == 0 && last > 0) std::fprintf(stderr, "[mpi_exp] %-24s %8.3f s\n", what, t - last);
  last = t;
}

[[noreturn]] void die(const char *msg) {
  std::fprintf(stderr, "mpi_exp rank %d: %s\n", g_rank, msg);
  MPI_Abort(MPI_COMM_WORLD, 1);
  std::exit(1);
}

// The deviate of exponential_contraction.h:467-471 restated: minstd_rand0,
// generate_canonical<double, 53> from two 31-bit outputs (fused sum), then
// -log(1 - u) / beta truncated.
u32 exp_deviate(u32 gid, int32_t seed, u32 iteration, u64 rng_offset, double beta) {
  const u32 s = static_cast<u32>(static_cast<u64>(gid) + static_cast<u64>(static_cast<int64_t>(seed)) +
                                 static_cast<u64>(iteration) * (rng_offset + static_cast<u64>(static_cast<int64_t>(seed))));
  const u64 m = 2147483647ull;
  u64 x = s % m;
  if (x == 0) x = 1;
  x = (x * 16807ull) % m;
  const double lo = static_cast<double>(x - 1);
  x = (x * 16807ull) % m;
  const double hi = static_cast<double>(x - 1);
  const double r = static_cast<double>(m - 1);
  double u = std::fma(hi, r, lo) / (r * r);
  if (u >= 1.0) u = std::nextafter(1.0, 0.0);
  return static_cast<u32>(static_cast<u64>(-std::log(1.0 - u) / beta));
}

// All-to-all of per-peer vectors of a trivially copyable type.  Counts and
// displacements are in elements of a contiguous datatype of sizeof(T) bytes
// (MPI's int counts would cap a rank's bytes at 2 GB; the rank that owns the
// hubs of a skewed graph receives more than that).
template <class T>
std::vector<T> all_to_all(const std::vector<std::vector<T>> &out, std::vector<u64> *from = nullptr) {
  const int P = g_size;
  MPI_Datatype elem;
  MPI_Type_contiguous(static_cast<int>(sizeof(T)), MPI_BYTE, &elem);
  MPI_Type_commit(&elem);
  std::vector<int> sc(P), rc(P), sd(P), rd(P);
  u64 stot = 0;
  for (int q = 0; q < P; ++q) {
    if (stot + out[q].size() > static_cast<u64>(INT_MAX)) die("exchange of more than 2^31 elements per rank");
    sc[q] = static_cast<int>(out[q].size());
    sd[q] = static_cast<int>(stot);
    stot += out[q].size();
  }
  MPI_Alltoall(sc.data(), 1, MPI_INT, rc.data(), 1, MPI_INT, MPI_COMM_WORLD);
  u64 rtot = 0;
  for (int q = 0; q < P; ++q) {
    if (rtot + static_cast<u64>(rc[q]) > static_cast<u64>(INT_MAX)) die("exchange of more than 2^31 elements per rank");
    rd[q] = static_cast<int>(rtot);
    rtot += static_cast<u64>(rc[q]);
  }
  std::vector<T> sbuf(std::max<u64>(stot, 1));
  for (int q = 0; q < P; ++q)
    if (!out[q].empty()) std::memcpy(sbuf.data() + sd[q], out[q].data(), out[q].size() * sizeof(T));
  std::vector<T> in(std::max<u64>(rtot, 1));
  MPI_Alltoallv(sbuf.data(), sc.data(), sd.data(), elem, in.data(), rc.data(), rd.data(), elem, MPI_COMM_WORLD);
  MPI_Type_free(&elem);
  in.resize(rtot);
  if (from) {
    from->assign(P + 1, 0);
    for (int q = 0; q < P; ++q) (*from)[q + 1] = (*from)[q] + static_cast<u64>(rc[q]);
  }
  return in;
}

int owner_of(const std::vector<u64> &starts, u64 x) {
  return static_cast<int>(std::upper_bound(starts.begin(), starts.end(), x) - starts.begin()) - 1;
}

// One contraction level of this rank: the owned vertices [lo, lo + nloc) of
// a contiguous split (starts), their rows of global level indices, their ids.
struct Level {
  std::vector<u64> starts;
  u64 lo = 0, nloc = 0;
  std::vector<u64> off;
  std::vector<u32> adj;
  std::vector<u32> gid;
  u64 m2_global = 0;
};

// What the undo keeps of a level.
struct Record {
  std::vector<u64> starts;
  u64 lo = 0;
  std::vector<u32> lab, newidx, gid;
};

// Rows re-indexed to [0, nloc) for owned and [nloc, nloc + ghosts) for the
// others, and the exchange plan of the ghost keys.
struct Halo {
  u64 ghosts = 0;
  std::vector<u32> adj_l;                 // local index per edge slot
  std::vector<std::vector<u32>> recvslot; // ghost slot of the i-th key peer q sends
  std::vector<std::vector<u32>> sendidx;  // owned indices peer q holds as ghosts
};

Halo build_halo(const Level &L) {
  Halo h;
  const int P = g_size;
  const u64 hi = L.lo + L.nloc;
  std::vector<u32> slot(L.starts[P], kNone);  // global index -> ghost slot
  std::vector<std::vector<u32>> req(P);
  h.recvslot.assign(P, {});
  h.adj_l.resize(L.adj.size());
  for (u64 e = 0; e < L.adj.size(); ++e) {
    const u32 w = L.adj[e];
    if (w >= L.lo && w < hi) {
      h.adj_l[e] = static_cast<u32>(w - L.lo);
      continue;
    }
    if (slot[w] == kNone) {
      slot[w] = static_cast<u32>(h.ghosts++);
      const int q = owner_of(L.starts, w);
      req[q].push_back(w);
      h.recvslot[q].push_back(slot[w]);
    }
    h.adj_l[e] = static_cast<u32>(L.nloc + slot[w]);
  }
  std::vector<u64> from;
  std::vector<u32> in = all_to_all(req, &from);
  h.sendidx.assign(P, {});
  for (int q = 0; q < P; ++q)
    for (u64 i = from[q]; i < from[q + 1]; ++i) h.sendidx[q].push_back(static_cast<u32>(in[i] - L.lo));
  return h;
}

struct KeyMsg {
  u32 pos;
  u32 pad;
  u64 key;
};

// Sends the owned keys peers hold as ghosts (all of them, or only the
// changed ones) and stores what arrives in the ghost slots.
void exchange_keys(const Halo &h, u64 nloc, std::vector<u64> &key, const std::vector<uint8_t> *dirty) {
  const int P = g_size;
  std::vector<std::vector<KeyMsg>> out(P);
  for (int q = 0; q < P; ++q)
    for (u32 i = 0; i < h.sendidx[q].size(); ++i) {
      const u32 v = h.sendidx[q][i];
      if (!dirty || (*dirty)[v]) out[q].push_back({i, 0, key[v]});
    }
  std::vector<u64> from;
  std::vector<KeyMsg> in = all_to_all(out, &from);
  for (int q = 0; q < P; ++q)
    for (u64 i = from[q]; i < from[q + 1]; ++i) key[nloc + h.recvslot[q][in[i].pos]] = in[i].key;
}

// Asks the owners of the sorted, distinct global indices `want` (split by
// starts) for one value each; answer(local index) gives it.
template <class F>
std::vector<u32> ask_owners(const std::vector<u64> &starts, u64 lo, const std::vector<u32> &want, F answer) {
  const int P = g_size;
  std::vector<std::vector<u32>> req(P);
  for (u32 w : want) req[owner_of(starts, w)].push_back(w);
  std::vector<u64> from;
  std::vector<u32> in = all_to_all(req, &from);
  std::vector<std::vector<u32>> rep(P);
  for (int q = 0; q < P; ++q)
    for (u64 i = from[q]; i < from[q + 1]; ++i) rep[q].push_back(answer(in[i] - lo));
  std::vector<u32> back = all_to_all(rep);  // in the order of want (grouped by owner, owners ascending)
  return back;
}

// The next level from the labels (global level index of each owned vertex's
// centre; ghost labels in lab_all[nloc..]): cluster edges routed to the
// owner of their first label, deduplicated there.
void contract(const Level &L, const Halo &h, const std::vector<u32> &lab_all, Record &rec, Level &next) {
  const int P = g_size;
  std::vector<std::vector<u64>> out(P);
  for (u64 v = 0; v < L.nloc; ++v) {
    const u32 lu = lab_all[v];
    const int q = owner_of(L.starts, lu);
    for (u64 e = L.off[v]; e < L.off[v + 1]; ++e) {
      const u32 lw = lab_all[h.adj_l[e]];
      if (lw != lu) out[q].push_back((static_cast<u64>(lu) << 32) | lw);
    }
  }
  mark("  cluster edges");
  std::vector<u64> pairs = all_to_all(out);
  out.clear();
  out.shrink_to_fit();
  mark("  routed");
  // the routed edges bucketed by their owned centre (counting sort), then
  // every centre's far ends sorted and deduplicated
  std::vector<u64> bstart(L.nloc + 1, 0);
  for (u64 p : pairs) bstart[(p >> 32) - L.lo + 1]++;
  for (u64 v = 0; v < L.nloc; ++v) bstart[v + 1] += bstart[v];
  std::vector<u32> far_of(pairs.size());
  {
    std::vector<u64> pos(bstart.begin(), bstart.end() - 1);
    for (u64 p : pairs) far_of[pos[(p >> 32) - L.lo]++] = static_cast<u32>(p);
  }
  pairs.clear();
  pairs.shrink_to_fit();
  std::vector<u32> centres;
  std::vector<u64> uoff(1, 0);
  u64 kept = 0;
  for (u64 v = 0; v < L.nloc; ++v) {
    u32 *b = far_of.data() + bstart[v], *e = far_of.data() + bstart[v + 1];
    if (b == e) continue;
    std::sort(b, e);
    e = std::unique(b, e);
    const u64 k = static_cast<u64>(e - b);
    std::memmove(far_of.data() + kept, b, k * sizeof(u32));
    kept += k;
    centres.push_back(static_cast<u32>(L.lo + v));
    uoff.push_back(kept);
  }
  far_of.resize(kept);
  mark("  sorted");
  u64 cnt = centres.size(), base = 0;
  MPI_Exscan(&cnt, &base, 1, MPI_UINT64_T, MPI_SUM, MPI_COMM_WORLD);
  if (g_rank == 0) base = 0;
  std::vector<u64> counts(P);
  MPI_Allgather(&cnt, 1, MPI_UINT64_T, counts.data(), 1, MPI_UINT64_T, MPI_COMM_WORLD);
  next.starts.assign(P + 1, 0);
  for (int q = 0; q < P; ++q) next.starts[q + 1] = next.starts[q] + counts[q];
  next.lo = base;
  next.nloc = cnt;
  rec.newidx.assign(L.nloc, kNone);
  next.gid.resize(cnt);
  for (u64 i = 0; i < cnt; ++i) {
    rec.newidx[centres[i] - L.lo] = static_cast<u32>(base + i);
    next.gid[i] = L.gid[centres[i] - L.lo];
  }
  // far ends: their next-level index from their owner, through a dense map
  std::vector<u32> slot(L.starts[P], kNone), far;
  for (u32 w : far_of)
    if (slot[w] == kNone) {
      slot[w] = 0;
      far.push_back(w);
    }
  std::sort(far.begin(), far.end());
  const std::vector<u32> far_new = ask_owners(L.starts, L.lo, far, [&](u64 i) { return rec.newidx[i]; });
  for (u64 i = 0; i < far.size(); ++i) slot[far[i]] = far_new[i];
  mark("  far ends");
  next.off.assign(uoff.begin(), uoff.end());
  next.adj.resize(kept);
  for (u64 i = 0; i < kept; ++i) next.adj[i] = slot[far_of[i]];
  u64 m2 = kept;
  MPI_Allreduce(&m2, &next.m2_global, 1, MPI_UINT64_T, MPI_SUM, MPI_COMM_WORLD);
  rec.starts = L.starts;
  rec.lo = L.lo;
  rec.gid = L.gid;
}

struct RunStats {
  u32 levels = 0;
  u64 sweeps = 0;
};

// FindComponents: rep (input id of the component's representative) of every
// owned input vertex.
std::vector<u32> find_components(const Level &input, u64 n_input, int32_t seed, double beta, u64 seq_limit,
                                 bool local_phase, RunStats &st) {
  std::vector<Record> records;
  u64 active = 0, active_global = 0;
  for (u64 v = 0; v < input.nloc; ++v) active += input.off[v + 1] > input.off[v];
  MPI_Allreduce(&active, &active_global, 1, MPI_UINT64_T, MPI_SUM, MPI_COMM_WORLD);
  const u64 rng_offset = active_global;
  Level cur = input;
  bool first = true;
  u32 iteration = 1;
  std::vector<u32> top;  // rep of the owned vertices of the top level (base case)
  bool have_top = false;
  while (cur.m2_global > 0) {
    const bool pre = first && local_phase;
    first = false;
    if (!pre && cur.m2_global / 2 <= seq_limit) {
      // RunSequentialCC: the remaining graph on rank 0
      std::vector<u32> edges;
      for (u64 v = 0; v < cur.nloc; ++v)
        for (u64 e = cur.off[v]; e < cur.off[v + 1]; ++e)
          if (cur.adj[e] > cur.lo + v) {
            edges.push_back(static_cast<u32>(cur.lo + v));
            edges.push_back(cur.adj[e]);
          }
      const u64 ntot = cur.starts[g_size];
      int cnt = static_cast<int>(edges.size()), gcnt = static_cast<int>(cur.nloc);
      std::vector<int> cnts(g_size), gcnts(g_size), displ(g_size), gdispl(g_size);
      MPI_Gather(&cnt, 1, MPI_INT, cnts.data(), 1, MPI_INT, 0, MPI_COMM_WORLD);
      MPI_Gather(&gcnt, 1, MPI_INT, gcnts.data(), 1, MPI_INT, 0, MPI_COMM_WORLD);
      u64 tot = 0;
      for (int q = 0; q < g_size; ++q) {
        displ[q] = static_cast<int>(tot);
        tot += static_cast<u64>(cnts[q]);
        gdispl[q] = static_cast<int>(cur.starts[q]);
      }
      if (g_rank == 0 && tot > static_cast<u64>(INT_MAX)) die("base case above 2^31 words");
      std::vector<u32> all(g_rank == 0 ? std::max<u64>(tot, 1) : 1), gids(g_rank == 0 ? std::max<u64>(ntot, 1) : 1);
      MPI_Gatherv(edges.data(), cnt, MPI_UINT32_T, all.data(), cnts.data(), displ.data(), MPI_UINT32_T, 0, MPI_COMM_WORLD);
      MPI_Gatherv(cur.gid.data(), gcnt, MPI_UINT32_T, gids.data(), gcnts.data(), gdispl.data(), MPI_UINT32_T, 0,
                  MPI_COMM_WORLD);
      std::vector<u32> rep(g_rank == 0 ? std::max<u64>(ntot, 1) : 1);
      if (g_rank == 0) {
        std::vector<u32> p(ntot);
        std::iota(p.begin(), p.end(), 0u);
        auto find = [&](u32 x) {
          while (p[x] != x) x = p[x] = p[p[x]];
          return x;
        };
        for (u64 i = 0; i + 1 < tot; i += 2) {
          const u32 a = find(all[i]), b = find(all[i + 1]);
          if (a != b) p[std::max(a, b)] = std::min(a, b);
        }
        for (u64 v = 0; v < ntot; ++v) rep[v] = gids[find(static_cast<u32>(v))];
      }
      top.resize(std::max<u64>(cur.nloc, 1));
      MPI_Scatterv(rep.data(), gcnts.data(), gdispl.data(), MPI_UINT32_T, top.data(), gcnt, MPI_UINT32_T, 0,
                   MPI_COMM_WORLD);
      top.resize(cur.nloc);
      have_top = true;
      mark("base case");
      break;
    }
    mark("level start");
    Halo h = build_halo(cur);
    mark("halo");
    const u64 nall = cur.nloc + h.ghosts;
    std::vector<u64> key(nall);
    std::vector<u32> lab_all(nall);
    if (pre) {
      // components of the edges inside the block; interface vertices stay
      std::vector<u32> p(cur.nloc);
      std::iota(p.begin(), p.end(), 0u);
      std::vector<uint8_t> inner(cur.nloc, 1);
      auto find = [&](u32 x) {
        while (p[x] != x) x = p[x] = p[p[x]];
        return x;
      };
      for (u64 v = 0; v < cur.nloc; ++v)
        for (u64 e = cur.off[v]; e < cur.off[v + 1]; ++e) {
          const u32 w = h.adj_l[e];
          if (w >= cur.nloc) {
            inner[v] = 0;
            continue;
          }
          const u32 a = find(static_cast<u32>(v)), b = find(w);
          if (a != b) p[std::max(a, b)] = std::min(a, b);
        }
      for (u64 v = 0; v < cur.nloc; ++v)
        key[v] = cur.lo + (inner[v] ? find(static_cast<u32>(v)) : v);
      exchange_keys(h, cur.nloc, key, nullptr);
    } else {
      for (u64 v = 0; v < cur.nloc; ++v)
        key[v] = (static_cast<u64>(exp_deviate(cur.gid[v], seed, iteration, rng_offset, beta)) << 32) | (cur.lo + v);
      exchange_keys(h, cur.nloc, key, nullptr);
      std::vector<uint8_t> dirty(cur.nloc, 0);
      u64 sweeps = 0;
      bool first_change = false;
      while (true) {
        int changed = 0;
        for (u64 v = 0; v < cur.nloc; ++v) {
          u64 best = key[v];
          for (u64 e = cur.off[v]; e < cur.off[v + 1]; ++e) {
            const u64 c = key[h.adj_l[e]] + kHop;
            best = c < best ? c : best;
          }
          if (best < key[v]) {
            key[v] = best;
            dirty[v] = 1;
            changed = 1;
          }
        }
        exchange_keys(h, cur.nloc, key, &dirty);
        std::fill(dirty.begin(), dirty.end(), 0);
        int any = 0;
        MPI_Allreduce(&changed, &any, 1, MPI_INT, MPI_MAX, MPI_COMM_WORLD);
        ++sweeps;
        if (sweeps == 1) first_change = any != 0;
        if (!any) break;
      }
      st.sweeps += sweeps;
      ++st.levels;
      if (!first_change) {  // every vertex is its own centre: the contraction is the identity
        ++iteration;
        continue;
      }
    }
    for (u64 i = 0; i < nall; ++i) lab_all[i] = static_cast<u32>(key[i]);
    key.clear();
    key.shrink_to_fit();
    Record rec;
    rec.lab.assign(lab_all.begin(), lab_all.begin() + static_cast<int64_t>(cur.nloc));
    Level next;
    mark("sweeps");
    contract(cur, h, lab_all, rec, next);
    mark("contract");
    records.push_back(std::move(rec));
    cur = std::move(next);
    if (!pre) ++iteration;
  }
  mark("levels done");
  // undo, top down: rep_l[v] = rep_{l+1}[newidx[lab[v]]], or gid[lab[v]]
  std::vector<u32> up = have_top ? top : std::vector<u32>();
  for (int l = static_cast<int>(records.size()) - 1; l >= 0; --l) {
    const Record &r = records[l];
    const u64 upper_lo = (l + 1 < static_cast<int>(records.size())) ? records[l + 1].lo : cur.lo;
    std::vector<u32> want(r.lab);
    std::sort(want.begin(), want.end());
    want.erase(std::unique(want.begin(), want.end()), want.end());
    const std::vector<u32> got = ask_owners(r.starts, r.lo, want, [&](u64 i) {
      const u32 a2 = r.newidx[i];
      return a2 != kNone ? up[a2 - upper_lo] : r.gid[i];
    });
    std::vector<u32> down(r.lab.size());
    for (u64 v = 0; v < r.lab.size(); ++v)
      down[v] = got[static_cast<u64>(std::lower_bound(want.begin(), want.end(), r.lab[v]) - want.begin())];
    up.swap(down);
  }
  mark("undo");
  // canonical: minimum input id per representative, asked of the representative's owner
  // (no level and no base case: the input has no edge, every vertex is alone)
  std::vector<u32> rep(input.nloc);
  for (u64 v = 0; v < input.nloc; ++v) rep[v] = up.empty() ? static_cast<u32>(input.lo + v) : up[v];
  std::vector<std::vector<u64>> out(g_size);
  for (u64 v = 0; v < input.nloc; ++v)
    out[owner_of(input.starts, rep[v])].push_back((static_cast<u64>(rep[v]) << 32) | (input.lo + v));
  std::vector<u64> in = all_to_all(out);
  std::vector<u32> minid(input.nloc, kNone);
  for (u64 x : in) {
    const u64 r = (x >> 32) - input.lo;
    minid[r] = std::min(minid[r], static_cast<u32>(x));
  }
  std::vector<u32> want(rep);
  std::sort(want.begin(), want.end());
  want.erase(std::unique(want.begin(), want.end()), want.end());
  const std::vector<u32> got = ask_owners(input.starts, input.lo, want, [&](u64 i) { return minid[i]; });
  std::vector<u32> labels(input.nloc);
  for (u64 v = 0; v < input.nloc; ++v)
    labels[v] = got[static_cast<u64>(std::lower_bound(want.begin(), want.end(), rep[v]) - want.begin())];
  (void)n_input;
  return labels;
}

// The owned block of the input CSR file.
Level read_block(const char *path, u64 &n) {
  FILE *f = std::fopen(path, "rb");
  if (!f) die("cannot open the CSR file");
  u64 head[2];
  if (std::fread(head, 8, 2, f) != 2) die("short CSR header");
  n = head[0];
  const u64 m2 = head[1];
  const u64 bs = (n + g_size - 1) / (g_size ? g_size : 1);
  Level L;
  L.starts.resize(g_size + 1);
  for (int q = 0; q <= g_size; ++q) L.starts[q] = std::min<u64>(n, static_cast<u64>(q) * bs);
  L.lo = L.starts[g_rank];
  L.nloc = L.starts[g_rank + 1] - L.lo;
  L.off.resize(L.nloc + 1);
  if (std::fseek(f, static_cast<long>(16 + 8 * L.lo), SEEK_SET) ||
      std::fread(L.off.data(), 8, L.nloc + 1, f) != L.nloc + 1)
    die("short CSR offsets");
  const u64 e0 = L.off[0], e1 = L.off[L.nloc];
  L.adj.resize(e1 - e0);
  if (std::fseek(f, static_cast<long>(16 + 8 * (n + 1) + 4 * e0), SEEK_SET) ||
      (e1 > e0 && std::fread(L.adj.data(), 4, e1 - e0, f) != e1 - e0))
    die("short CSR targets");
  std::fclose(f);
  for (u64 &o : L.off) o -= e0;
  L.gid.resize(L.nloc);
  for (u64 v = 0; v < L.nloc; ++v) L.gid[v] = static_cast<u32>(L.lo + v);
  L.m2_global = m2;
  return L;
}

}  // namespace

int main(int argc, char **argv) {
  MPI_Init(&argc, &argv);
  MPI_Comm_rank(MPI_COMM_WORLD, &g_rank);
  MPI_Comm_size(MPI_COMM_WORLD, &g_size);
  const char *path = nullptr, *labels_path = nullptr;
  int seed = 1, iters = 1, local_phase = 1;
  double beta = 0.1;
  u64 seq = 1;
  bool fixed = false;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> const char * {
      if (i + 1 >= argc) die("missing flag value");
      return argv[++i];
    };
    if (a == "-seed") seed = std::atoi(next());
    else if (a == "-beta") beta = std::atof(next());
    else if (a == "-i") iters = std::atoi(next());
    else if (a == "-seq") seq = std::strtoull(next(), nullptr, 10);
    else if (a == "-fixed-seq") fixed = true;
    else if (a == "-no-local") local_phase = 0;
    else if (a == "-labels") labels_path = next();
    else if (a == "-v") g_verbose = true;
    else if (a[0] != '-') path = argv[i];
    else die("unknown flag");
  }
  if (!path) die("usage: mpi_exp <csr file> [-seed s] [-beta b] [-i iters] [-seq k] [-fixed-seq] [-no-local] [-labels out]");
  u64 n = 0;
  const Level input = read_block(path, n);
  const u64 m = input.m2_global / 2;
  // ComputeSequentialLimit (exponential_contraction.h:1287-1292), in undirected edges
  const u64 seq_limit = fixed ? seq : (m / static_cast<u64>(g_size)) * seq;
  double sum = 0, sum2 = 0;
  RunStats last;
  std::vector<u32> labels;
  for (int it = 0; it < iters; ++it) {
    const int round_seed = seed + it + 1000;
    RunStats st;
    MPI_Barrier(MPI_COMM_WORLD);
    const double t0 = MPI_Wtime();
    labels = find_components(input, n, round_seed, beta, seq_limit, local_phase != 0, st);
    const double local = MPI_Wtime() - t0;
    double t = 0;
    MPI_Reduce(&local, &t, 1, MPI_DOUBLE, MPI_MAX, 0, MPI_COMM_WORLD);
    sum += t;
    sum2 += t * t;
    last = st;
  }
  if (labels_path) {
    std::vector<int> cnts(g_size), displ(g_size);
    for (int q = 0; q < g_size; ++q) {
      cnts[q] = static_cast<int>(input.starts[q + 1] - input.starts[q]);
      displ[q] = static_cast<int>(input.starts[q]);
    }
    std::vector<u32> all(g_rank == 0 ? std::max<u64>(n, 1) : 1);
    MPI_Gatherv(labels.data(), cnts[g_rank], MPI_UINT32_T, all.data(), cnts.data(), displ.data(), MPI_UINT32_T, 0,
                MPI_COMM_WORLD);
    if (g_rank == 0) {
      FILE *f = std::fopen(labels_path, "wb");
      if (!f || std::fwrite(all.data(), 4, n, f) != n) die("cannot write the labels");
      std::fclose(f);
    }
  }
  if (g_rank == 0) {
    const double avg = sum / iters, var = iters > 1 ? std::max(0.0, (sum2 - iters * avg * avg) / (iters - 1)) : 0.0;
    std::printf("GLOBAL RESULT runner=exp time=%.6f stddev=%.6f comm_time=0 stddev=0 send_volume=0 stddev=0 "
                "recv_volume=0 stddev=0 iterations=%d ranks=%d n=%llu m=%llu seq_limit=%llu local_phase=%d "
                "levels=%u sweeps=%llu\n",
                avg, std::sqrt(var), iters, g_size, static_cast<unsigned long long>(n), static_cast<unsigned long long>(m),
                static_cast<unsigned long long>(seq_limit), local_phase, last.levels,
                static_cast<unsigned long long>(last.sweeps));
  }
  MPI_Finalize();
  return 0;
}
