// Exact maximum clique with the output of FMC::maxClique (findClique.cpp:31-145,
// wrapper findMaxClique src/utils/GraphUtils.cpp:9-17, lower bound 0):
// vertices are tried as seeds from n-1 down to 0; each seed's clique grows
// only with lower-index neighbours, by depth-first search that always
// extends with the highest candidate and prunes (size + |U| <= best, and
// candidates of degree < best).  The returned clique is the first
// maximum-size clique of that search order, listed deepest vertex first and
// the seed last.
//
// That output is fixed by three facts, computed here in three phases:
//   A. the maximum clique size omega: parallel branch and bound, one
//      wavefront per seed (persistent wavefronts pull seeds from a counter),
//      candidate sets as bitsets with one word per lane, the bound shared
//      through a device atomic and seeded with the greedy clique size;
//   B. the highest seed s whose search holds an omega-clique: the same
//      search in existence mode (prune size + |U| < omega, stop at the first);
//   C. the first omega-clique of seed s's search, replayed by one wavefront
//      with best = omega - 1 (branches that cannot reach omega are the only
//      ones this bound removes, so the first omega-clique met is the one the
//      sequential search records).
// Besides FMC's own bounds every node is also bounded by a greedy colouring of
// its candidate set: a clique holds at most one vertex of each colour class,
// so a node whose clique plus its class count cannot beat the bound is
// pruned.  Like FMC's bounds this removes only branches that cannot hold a
// larger (phase A) or an omega-sized (phases B, C) clique, so the search
// order, omega and the first omega-clique met are unchanged; what it changes
// is the size of the tree on dense inlier blocks, where degrees bound
// nothing and FMC's search visits every way of dropping one endpoint of each
// missing edge.
// A work budget (nodes plus colouring steps) and a time limit per phase
// (the device's constant-rate wall clock) bound the search; exceeding either
// is reported as an error.  Every wavefront checks both every 4,096 units
// of work, so each phase ends within about that much past its limit.
#include <algorithm>
#include <climits>
#include <vector>

#include "rpgo_internal.h"

namespace rpgo {

struct ExactCtl {
  unsigned long long best;  // (size << 32) | seed of the best clique so far
  int next_seed;            // seeds still to claim, counted down
  int aborted;              // 1: work budget exceeded, 2: time limit
  unsigned long long nodes;
  unsigned long long found;  // phase B: highest seed with an omega-clique (+1), atomicMax
};

constexpr int kExactWaves = 256;
constexpr int kExactPhaseSeconds = 20;  // device time limit of each phase

// wavefront sum and max by DPP (whole wavefronts only: every call site runs
// with all 64 lanes active)
__device__ __forceinline__ int wsum_i(int v) { return __ockl_wfred_add_i32(v); }
__device__ __forceinline__ int wmax_i(int v) { return __ockl_wfred_max_i32(v); }

constexpr int kExWPL = 8;  // words per lane of a candidate set (n <= 32768)

// Greedy colouring of the candidate set U (W words, word w on lane w % 64):
// each class is built from the lowest uncoloured vertex up, dropping the
// neighbours of every vertex it takes, so it is an independent set.  Returns
// true when U fits in `limit` classes; stops with false as soon as it needs
// more.  *steps counts the vertices coloured (the work).
__device__ bool colour_fits(const unsigned long long* __restrict__ U, int W,
                            const unsigned long long* __restrict__ adj, int lane, int limit, int* steps) {
  if (limit <= 0) return false;
  unsigned long long P[kExWPL];
#pragma unroll
  for (int k = 0; k < kExWPL; ++k) {
    const int w = lane + 64 * k;
    P[k] = w < W ? U[w] : 0ull;
  }
  int classes = 0, st = 0;
  for (;;) {
    bool left = false;
#pragma unroll
    for (int k = 0; k < kExWPL; ++k) left |= P[k] != 0ull;
    if (__ballot(left) == 0ull) {
      *steps += st;
      return true;
    }
    if (classes == limit) {
      *steps += st;
      return false;
    }
    ++classes;
    unsigned long long Q[kExWPL];
#pragma unroll
    for (int k = 0; k < kExWPL; ++k) Q[k] = P[k];
    for (;;) {
      // lowest vertex of Q: word w = lane + 64 k, so the lowest k with a
      // non-empty word, its lowest such lane (one ballot per k), that lane's
      // lowest bit by v_readlane
      int v = INT_MAX;
#pragma unroll
      for (int k = 0; k < kExWPL; ++k) {
        const unsigned long long lanes = __ballot(Q[k] != 0ull);
        if (lanes) {
          v = __builtin_amdgcn_readlane((lane + 64 * k) * 64 + __builtin_ctzll(Q[k]), __builtin_ctzll(lanes));
          break;
        }
      }
      if (v == INT_MAX) break;
      ++st;
      const int wv = v >> 6;
      const unsigned long long bit = 1ull << (v & 63);
      const unsigned long long* __restrict__ row = adj + (size_t)v * W;
#pragma unroll
      for (int k = 0; k < kExWPL; ++k) {
        const int w = lane + 64 * k;
        if (w == wv) {
          P[k] &= ~bit;
          Q[k] &= ~bit;
        }
        if (w < W && Q[k]) Q[k] &= ~row[w];  // empty words stay empty: their row words are not read
      }
    }
  }
}

// bits u < seed of word w
__device__ __forceinline__ unsigned long long below(int seed, int w) {
  const int lo = w * 64;
  if (lo + 64 <= seed) return ~0ull;
  if (lo >= seed) return 0ull;
  return (1ull << (seed - lo)) - 1ull;
}

// MODE 0: bound search (phase A); MODE 1: existence of an om-clique (phase B);
// MODE 2: replay of one seed recording the clique (phase C, one wavefront).
template <int MODE>
__global__ void __launch_bounds__(256) k_exact(int n, int W, const unsigned long long* __restrict__ adj,
                                               const int* __restrict__ deg,
                                               const unsigned long long* __restrict__ dmask, int om,
                                               int seed_lo, ExactCtl* ctl, unsigned long long* __restrict__ stacks,
                                               int depth_cap, int* __restrict__ path, int* __restrict__ out,
                                               long long node_budget, long long clock_limit) {
  const int lane = threadIdx.x & 63;
  const unsigned long long t_start = wall_clock64();
  const int wid = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  unsigned long long* stk = stacks + (size_t)wid * depth_cap * W;
  int* pth = path + (size_t)wid * depth_cap;
  long long local_nodes = 0;
  for (;;) {
    int seed = 0;
    if (MODE == 2) {
      if (wid != 0) return;
      seed = seed_lo;
    } else {
      if (lane == 0) seed = atomicSub(&ctl->next_seed, 1);
      seed = __shfl(seed, 0);
      if (seed < seed_lo) return;
      if (__hip_atomic_load(&ctl->aborted, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    }
    int bound = om - 1;  // prune when size + |U| <= bound
    if (MODE == 0) bound = (int)(__hip_atomic_load(&ctl->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32);
    if (deg[seed] < bound) {  // pruning 1
      if (MODE == 2) return;
      continue;
    }
    // level 0: lower neighbours of the seed that pass the degree mask
    int cnt = 0;
    for (int w = lane; w < W; w += 64) {
      const unsigned long long v = adj[(size_t)seed * W + w] & below(seed, w) & dmask[w];
      stk[w] = v;
      cnt += __popcll(v);
    }
    cnt = wsum_i(cnt);
    bool done = false;
    int depth = 0;
    if (cnt == 0) {  // the seed alone
      if (MODE == 0) {
        if (1 > bound && lane == 0) atomicMax(&ctl->best, (1ull << 32) | (unsigned)seed);
      } else if (om <= 1) {
        done = true;
      }
      depth = -1;
    }
    while (depth >= 0 && !done) {
      unsigned long long* U = stk + (size_t)depth * W;
      const int size = depth + 1;
      int c = 0, top = -1;
      for (int w = lane; w < W; w += 64) {
        const unsigned long long v = U[w];
        c += __popcll(v);
        if (v) top = w * 64 + 63 - __clzll(v);
      }
      c = wsum_i(c);
      top = wmax_i(top);
      if (MODE == 0) {
        int b = 0;
        if (lane == 0)
          b = (int)(__hip_atomic_load(&ctl->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32);
        bound = __shfl(b, 0);
      }
      if (c == 0 || size + c <= bound) {
        --depth;
        continue;
      }
      {
        int steps = 0;
        const bool fits = colour_fits(U, W, adj, lane, bound - size, &steps);
        local_nodes += steps;
        if (fits) {
          --depth;
          continue;
        }
      }
      // pop the highest candidate and build the child's candidate set
      if (lane == ((top >> 6) & 63)) U[top >> 6] &= ~(1ull << (top & 63));
      __builtin_amdgcn_wave_barrier();
      unsigned long long* C = U + W;
      int cc = 0;
      for (int w = lane; w < W; w += 64) {
        const unsigned long long v = U[w] & adj[(size_t)top * W + w] & dmask[w];
        C[w] = v;
        cc += __popcll(v);
      }
      cc = wsum_i(cc);
      if (MODE == 2 && lane == 0) pth[depth] = top;
      if (cc == 0) {  // leaf: a clique of size + 1
        if (MODE == 0) {
          if (size + 1 > bound && lane == 0)
            atomicMax(&ctl->best, ((unsigned long long)(size + 1) << 32) | (unsigned)seed);
        } else if (size + 1 >= om) {
          done = true;
        }
      } else {
        ++depth;
      }
      if (++local_nodes >= 4096) {
        unsigned long long tot = 0;
        if (lane == 0) tot = atomicAdd(&ctl->nodes, (unsigned long long)local_nodes) + local_nodes;
        tot = __shfl(tot, 0);
        local_nodes = 0;
        const bool late = (long long)(wall_clock64() - t_start) > clock_limit;
        if ((long long)tot > node_budget || late) {
          if (lane == 0) atomicMax(&ctl->aborted, late ? 2 : 1);
          return;
        }
      }
    }
    if (MODE == 1 && done && lane == 0) atomicMax(&ctl->found, (unsigned long long)seed + 1ull);
    if (MODE == 2) {
      if (lane == 0) {
        if (done) {
          // deepest vertex first, the seed last (unwinding order of findClique.cpp:73,127-130)
          int k = 0;
          for (int d = depth; d >= 0; --d) out[k++] = pth[d];
          out[k++] = seed;
          out[n] = k;
        } else {
          out[n] = (om <= 1) ? 1 : 0;
          if (om <= 1) out[0] = seed;
        }
      }
      return;
    }
  }
}

__global__ void k_exact_dmask(int n, int W, int thr, const int* __restrict__ deg,
                              unsigned long long* __restrict__ dmask) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  unsigned long long b = 0ull;
  for (int k = 0; k < 64; ++k) {
    const int u = w * 64 + k;
    if (u < n && deg[u] >= thr) b |= 1ull << k;
  }
  dmask[w] = b;
}

// degrees without the diagonal and bits >= n (graphIO.cpp:188-257)
__global__ void k_exact_degrees(int n, int W, const unsigned long long* __restrict__ adj, int* __restrict__ deg) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  int s = 0;
  for (int w = 0; w < W; ++w) {
    unsigned long long b = adj[(size_t)v * W + w];
    if (w == W - 1 && (n & 63)) b &= (1ull << (n & 63)) - 1ull;
    if (w == (v >> 6)) b &= ~(1ull << (v & 63));
    s += __popcll(b);
  }
  deg[v] = s;
}

// symmetric copy of the adjacency with the diagonal and the bits >= n cleared
__global__ void k_exact_clean(int n, int W, const unsigned long long* __restrict__ adj,
                              unsigned long long* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * W) return;
  const int v = (int)(t / W), w = (int)(t % W);
  unsigned long long b = adj[t];
  if (w == W - 1 && (n & 63)) b &= (1ull << (n & 63)) - 1ull;
  if (w == (v >> 6)) b &= ~(1ull << (v & 63));
  out[t] = b;
}

int clique_exact_run(rpgo_ctx* ctx, const unsigned long long* adj_in, int n, int lower_bound, int32_t* idx_out,
                     int32_t* k_out, long long node_budget) {
  CliqueState& s = ctx->clq;
  hipStream_t st = ctx->stream;
  if (n <= 0) {
    *k_out = 0;  // findMaxClique: maxClique with lower bound 0 finds nothing
    return RPGO_OK;
  }
  const int W = (n + 63) / 64;
  int rc;
  if ((rc = s.ex_adj.reserve(sizeof(unsigned long long) * (size_t)n * W))) return rc;
  if ((rc = s.deg.reserve(sizeof(int) * n))) return rc;
  if ((rc = s.dmask.reserve(sizeof(unsigned long long) * W))) return rc;
  if ((rc = s.ex_ctl.reserve(sizeof(ExactCtl)))) return rc;
  if ((rc = s.out.reserve(sizeof(int) * (n + 1)))) return rc;
  unsigned long long* adj = s.ex_adj.as<unsigned long long>();
  k_exact_clean<<<(unsigned)(((long long)n * W + 255) / 256), 256, 0, st>>>(n, W, adj_in, adj);
  k_exact_degrees<<<(n + 255) / 256, 256, 0, st>>>(n, W, adj, s.deg.as<int>());
  RPGO_HIP_CHECK(hipGetLastError());
  std::vector<int> deg(n);
  RPGO_HIP_CHECK(hipMemcpyAsync(deg.data(), s.deg.ptr, sizeof(int) * n, hipMemcpyDeviceToHost, st));
  RPGO_HIP_CHECK(hipStreamSynchronize(st));
  const int maxdeg = *std::max_element(deg.begin(), deg.end());
  const int depth_cap = maxdeg + 2;
  const size_t stack_words = (size_t)kExactWaves * depth_cap * W;
  if ((rc = s.ex_stack.reserve(sizeof(unsigned long long) * stack_words))) return rc;
  if ((rc = s.ex_path.reserve(sizeof(int) * (size_t)kExactWaves * depth_cap))) return rc;
  ExactCtl* ctl = s.ex_ctl.as<ExactCtl>();
  // time limit of each phase in wall-clock ticks (rate in kHz)
  int clock_khz = 0;
  if (hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, ctx->device) != hipSuccess ||
      clock_khz <= 0)
    clock_khz = 100000;
  const long long clock_limit = (long long)clock_khz * 1000ll * kExactPhaseSeconds;
  // phase A: omega, seeded with a clique size known to exist
  const int b0 = std::max(0, lower_bound - 1);
  ExactCtl h{};
  h.best = (unsigned long long)b0 << 32;
  h.next_seed = n - 1;
  RPGO_HIP_CHECK(hipMemcpyAsync(ctl, &h, sizeof(h), hipMemcpyHostToDevice, st));
  k_exact_dmask<<<(W + 255) / 256, 256, 0, st>>>(n, W, b0, s.deg.as<int>(), s.dmask.as<unsigned long long>());
  k_exact<0><<<kExactWaves / 4, 256, 0, st>>>(n, W, adj, s.deg.as<int>(), s.dmask.as<unsigned long long>(), 0, 0,
                                              ctl, s.ex_stack.as<unsigned long long>(), depth_cap,
                                              s.ex_path.as<int>(), s.out.as<int>(), node_budget, clock_limit);
  RPGO_HIP_CHECK(hipGetLastError());
  RPGO_HIP_CHECK(hipMemcpyAsync(&h, ctl, sizeof(h), hipMemcpyDeviceToHost, st));
  RPGO_HIP_CHECK(hipStreamSynchronize(st));
  if (h.aborted) {
    if (h.aborted == 2)
      set_error("rpgo_max_clique: exact search exceeded its time limit (%d s)", kExactPhaseSeconds);
    else
      set_error("rpgo_max_clique: exact search exceeded the node budget (%lld)", node_budget);
    return RPGO_ERR_STATE;
  }
  const int om = (int)(h.best >> 32);
  const int seed_a = (int)(h.best & 0xffffffffu);
  if (om <= b0) {
    set_error("rpgo_max_clique: exact search found no clique above the lower bound %d", lower_bound);
    return RPGO_ERR_STATE;
  }
  // phase B: the highest seed >= seed_a with an omega-clique
  h.next_seed = n - 1;
  h.found = 0;
  h.nodes = 0;
  RPGO_HIP_CHECK(hipMemcpyAsync(ctl, &h, sizeof(h), hipMemcpyHostToDevice, st));
  k_exact_dmask<<<(W + 255) / 256, 256, 0, st>>>(n, W, om - 1, s.deg.as<int>(), s.dmask.as<unsigned long long>());
  k_exact<1><<<kExactWaves / 4, 256, 0, st>>>(n, W, adj, s.deg.as<int>(), s.dmask.as<unsigned long long>(), om,
                                              seed_a, ctl, s.ex_stack.as<unsigned long long>(), depth_cap,
                                              s.ex_path.as<int>(), s.out.as<int>(), node_budget, clock_limit);
  RPGO_HIP_CHECK(hipGetLastError());
  RPGO_HIP_CHECK(hipMemcpyAsync(&h, ctl, sizeof(h), hipMemcpyDeviceToHost, st));
  RPGO_HIP_CHECK(hipStreamSynchronize(st));
  if (h.aborted || h.found == 0) {
    set_error("rpgo_max_clique: exact search %s",
              h.aborted == 2 ? "exceeded its time limit"
                             : (h.aborted ? "exceeded the node budget" : "lost the maximum clique"));
    return RPGO_ERR_STATE;
  }
  const int seed = (int)(h.found - 1);
  h.nodes = 0;
  RPGO_HIP_CHECK(hipMemcpyAsync(ctl, &h, sizeof(h), hipMemcpyHostToDevice, st));
  // phase C: replay the winning seed
  k_exact<2><<<1, 64, 0, st>>>(n, W, adj, s.deg.as<int>(), s.dmask.as<unsigned long long>(), om, seed, ctl,
                               s.ex_stack.as<unsigned long long>(), depth_cap, s.ex_path.as<int>(), s.out.as<int>(),
                               node_budget, clock_limit);
  RPGO_HIP_CHECK(hipGetLastError());
  int k = 0;
  RPGO_HIP_CHECK(hipMemcpyAsync(&k, s.out.as<int>() + n, sizeof(int), hipMemcpyDeviceToHost, st));
  RPGO_HIP_CHECK(hipMemcpyAsync(&h, ctl, sizeof(h), hipMemcpyDeviceToHost, st));
  RPGO_HIP_CHECK(hipStreamSynchronize(st));
  if (h.aborted) {  // the replay stopped before writing its result
    set_error("rpgo_max_clique: exact search %s in its replay",
              h.aborted == 2 ? "exceeded its time limit" : "exceeded the node budget");
    return RPGO_ERR_STATE;
  }
  if (k != om) {
    set_error("rpgo_max_clique: replay found size %d, expected %d", k, om);
    return RPGO_ERR_STATE;
  }
  if (idx_out) RPGO_HIP_CHECK(hipMemcpy(idx_out, s.out.ptr, sizeof(int) * k, hipMemcpyDeviceToHost));
  *k_out = k;
  return RPGO_OK;
}

}  // namespace rpgo
