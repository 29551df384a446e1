// Maximum clique on the bit-packed PCM adjacency.
//
// HEU_REF mode reproduces FMC::maxCliqueHeu (findCliqueHeu.cpp:32-117, called
// through findMaxCliqueHeu src/utils/GraphUtils.cpp:19-27) bit for bit:
//   * candidates c = 0..n-1 in order, skipped when best > deg(c)   (pruning 1)
//   * S = [c] ++ { u in N(c) : deg(u) >= best } ascending          (pruning 3)
//   * repeat: icc++, m = last of S, S = S /\ N(m) /\ {deg >= best} (pruning 5)
//   * the candidate beating best keeps the working buffer it leaves behind:
//     position p holds the value last written there, not the greedy picks.
// The sequential scan is split into speculative waves: every remaining
// candidate is evaluated in parallel against the current best (one wavefront
// per candidate, S as a bitset in registers), the first candidate that beats
// it is taken, and the scan resumes after it.  Each wave advances best, so
// the number of waves is the number of improvements.  The winner is then
// replayed by one wavefront that materialises the reference's stale buffer.
#include "rpgo_internal.h"

#include <algorithm>
#include <climits>
#include <vector>

namespace rpgo {

__global__ void k_degrees(int n, int W, const unsigned long long* __restrict__ adj,
                          int* __restrict__ deg) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  // only bits u < n, u != v count (FMC's CSR skips the diagonal, graphIO.cpp:199)
  int s = 0;
  for (int w = 0; w < W; ++w) {
    unsigned long long b = adj[(size_t)v * W + w];
    if (w == W - 1 && (n & 63)) b &= (1ull << (n & 63)) - 1ull;
    if (w == (v >> 6)) b &= ~(1ull << (v & 63));
    s += __popcll(b);
  }
  deg[v] = s;
}

__global__ void k_dmask(int n, int W, int best, const int* __restrict__ deg,
                        unsigned long long* __restrict__ dmask) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  unsigned long long b = 0ull;
  for (int k = 0; k < 64; ++k) {
    const int u = w * 64 + k;
    if (u < n && deg[u] >= best) b |= 1ull << k;
  }
  dmask[w] = b;
}

// words per lane held in registers (n <= 64 * 64 * kMaxWPL)
constexpr int kMaxWPL = 8;
// search nodes the exact mode may visit before it reports failure
constexpr long long kExactNodeBudget = 400000000ll;

// Highest set bit over the wave's distributed bitset; -1 when empty.
__device__ __forceinline__ int wave_top_bit(const unsigned long long* S, int lane, int W) {
  int mine = -1;
#pragma unroll
  for (int q = kMaxWPL - 1; q >= 0; --q) {
    const int w = lane + 64 * q;
    if (mine < 0 && w < W && S[q]) mine = w * 64 + 63 - __clzll(S[q]);
  }
  // wave max
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(mine, off);
    mine = o > mine ? o : mine;
  }
  return mine;
}

// One wavefront per candidate c >= pos: greedy size icc(c, best).
__global__ void __launch_bounds__(256) k_icc(int n, int W, int pos, int best,
                                             const unsigned long long* __restrict__ adj,
                                             const int* __restrict__ deg,
                                             const unsigned long long* __restrict__ dmask,
                                             int* __restrict__ icc_out) {
  const int c = pos + blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= n) return;
  const int lane = threadIdx.x & 63;
  if (best > deg[c]) {  // pruning 1
    if (lane == 0) icc_out[c] = INT_MIN;
    return;
  }
  unsigned long long S[kMaxWPL];
#pragma unroll
  for (int q = 0; q < kMaxWPL; ++q) {
    const int w = lane + 64 * q;
    S[q] = (w < W) ? (adj[(size_t)c * W + w] & dmask[w] &
                      ((w == (c >> 6)) ? ~(1ull << (c & 63)) : ~0ull))
                   : 0ull;
  }
  int icc = 0;
  for (;;) {
    ++icc;
    const int m = wave_top_bit(S, lane, W);
    if (m < 0) break;  // S held only the seed c: removing it empties S
    // S /\ N(m); m itself always leaves S, so the loop ends within n steps
    const unsigned long long* row = adj + (size_t)m * W;
#pragma unroll
    for (int q = 0; q < kMaxWPL; ++q) {
      const int w = lane + 64 * q;
      if (w < W) S[q] &= row[w] & ((w == (m >> 6)) ? ~(1ull << (m & 63)) : ~0ull);
    }
  }
  if (lane == 0) icc_out[c] = icc;
}

__global__ void k_first_better(int n, int pos, int best, const int* __restrict__ icc,
                               int* __restrict__ first) {
  const int c = pos + blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  if (icc[c] > best) atomicMin(first, c);
}

// Replay of the winning candidate: writes the reference's buffer positions
// 0..s0-1 (position p keeps the value last written to it).  One wavefront.
__global__ void __launch_bounds__(64) k_replay(int n, int W, int c, int best,
                                               const unsigned long long* __restrict__ adj,
                                               const unsigned long long* __restrict__ dmask,
                                               int* __restrict__ out, int* __restrict__ icc_out) {
  const int lane = threadIdx.x;
  unsigned long long S[kMaxWPL];
#pragma unroll
  for (int q = 0; q < kMaxWPL; ++q) {
    const int w = lane + 64 * q;
    S[q] = (w < W) ? (adj[(size_t)c * W + w] & dmask[w] &
                      ((w == (c >> 6)) ? ~(1ull << (c & 63)) : ~0ull))
                   : 0ull;
  }
  auto wave_count = [&](const unsigned long long* B) {
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < kMaxWPL; ++q) cnt += __popcll(B[q]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
    return cnt;
  };
  if (lane == 0) out[0] = c;
  int s_t = 1 + wave_count(S);  // list size before the iteration
  int icc = 0;
  for (;;) {
    ++icc;
    const int m = wave_top_bit(S, lane, W);
    unsigned long long N[kMaxWPL];
    if (m >= 0) {
      const unsigned long long* row = adj + (size_t)m * W;
#pragma unroll
      for (int q = 0; q < kMaxWPL; ++q) {
        const int w = lane + 64 * q;
        N[q] = (w < W) ? (S[q] & row[w] & ((w == (m >> 6)) ? ~(1ull << (m & 63)) : ~0ull)) : 0ull;
      }
    } else {
#pragma unroll
      for (int q = 0; q < kMaxWPL; ++q) N[q] = 0ull;
    }
    // list after the iteration: the seed survives unless it was the pick
    const int s_next = (m >= 0) ? 1 + wave_count(N) : 0;
    // positions [max(s_next,1), s_t) keep S_t's entries: ranks p-1 of the set part
    const int lo = s_next > 1 ? s_next : 1;
    if (lo < s_t) {
      // rank of each of this lane's words: exclusive scan in word order
      // (word w = lane + 64 q; order is q-major across lanes)
      int base = 0;
#pragma unroll
      for (int q = 0; q < kMaxWPL; ++q) {
        const int cnt = __popcll(S[q]);
        int incl = cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          const int o = __shfl_up(incl, off);
          if (lane >= off) incl += o;
        }
        const int total = __shfl(incl, 63);
        int r = base + incl - cnt;  // rank of the first set bit of this word
        unsigned long long b = S[q];
        const int w = lane + 64 * q;
        while (b) {
          const int bit = __ffsll((long long)b) - 1;
          b &= b - 1ull;
          const int p = r + 1;
          if (p >= lo && p < s_t) out[p] = w * 64 + bit;
          ++r;
        }
        base += total;
      }
    }
#pragma unroll
    for (int q = 0; q < kMaxWPL; ++q) S[q] = N[q];
    s_t = s_next;
    if (m < 0) break;
  }
  if (lane == 0) *icc_out = icc;
  (void)best;
  (void)n;
}

// Greedy scan over candidates pos0 .. n-1 starting from the bound best0;
// mode EXACT continues with the exact search, mode HEU_REF replays the winner.
static int clique_scan(rpgo_ctx* ctx, const uint64_t* d_adj_u64, int n, int mode, int pos0, int best0,
                       int32_t* idx_out, int32_t* k_out);

int clique_run(rpgo_ctx* ctx, const uint64_t* d_adj_u64, int n, int mode, int32_t* idx_out,
               int32_t* k_out) {
  return clique_scan(ctx, d_adj_u64, n, mode, 0, -1, idx_out, k_out);
}

// FMC::maxCliqueHeuIncremental (findCliqueHeu.cpp:119-209) through
// findMaxCliqueHeuIncremental (GraphUtils.cpp:30-44): candidates are only the
// last num_new vertices, the bound starts at the previous clique size, and
// the size is reported only when it strictly grew (0 otherwise).
int clique_incremental(rpgo_ctx* ctx, const uint64_t* d_adj_u64, int n, int num_new, int prev_size,
                       int32_t* idx_out, int32_t* k_out) {
  if (num_new < 0 || num_new > n || prev_size < 0) {
    set_error("rpgo_max_clique_incremental: num_new %d / prev_size %d invalid for %d vertices", num_new,
              prev_size, n);
    return RPGO_ERR_ARG;
  }
  if (n <= 0 || num_new == 0) {
    *k_out = 0;
    return RPGO_OK;
  }
  int32_t k = 0;
  const int rc = clique_scan(ctx, d_adj_u64, n, RPGO_CLIQUE_HEU_REF, n - num_new, prev_size, idx_out, &k);
  if (rc) return rc;
  *k_out = k > prev_size ? k : 0;
  return RPGO_OK;
}

static int clique_scan(rpgo_ctx* ctx, const uint64_t* d_adj_u64, int n, int mode, int pos0, int best0,
                       int32_t* idx_out, int32_t* k_out) {
  CliqueState& s = ctx->clq;
  if (mode != RPGO_CLIQUE_HEU_REF && mode != RPGO_CLIQUE_EXACT) {
    set_error("rpgo_max_clique: unknown mode %d", mode);
    return RPGO_ERR_ARG;
  }
  if (n <= 0) {
    *k_out = mode == RPGO_CLIQUE_EXACT ? 0 : -1;  // FMC: maxClique finds nothing, maxCliqueHeu returns -1
    return RPGO_OK;
  }
  const int W = (n + 63) / 64;
  if (W > 64 * kMaxWPL) {
    set_error("rpgo_max_clique: %d vertices exceed the %d supported", n, 64 * 64 * kMaxWPL);
    return RPGO_ERR_ARG;
  }
  const unsigned long long* adj = reinterpret_cast<const unsigned long long*>(d_adj_u64);
  int rc;
  if ((rc = s.deg.reserve(sizeof(int) * n))) return rc;
  if ((rc = s.dmask.reserve(sizeof(unsigned long long) * W))) return rc;
  if ((rc = s.icc.reserve(sizeof(int) * n))) return rc;
  if ((rc = s.out.reserve(sizeof(int) * (n + 1)))) return rc;
  if ((rc = s.best.reserve(sizeof(int) * 4))) return rc;
  hipStream_t st = ctx->stream;
  k_degrees<<<(n + 255) / 256, 256, 0, st>>>(n, W, adj, s.deg.as<int>());
  RPGO_HIP_CHECK(hipGetLastError());
  int best = best0, pos = pos0, win = -1, win_best = -1;
  int* d_first = s.best.as<int>();
  while (pos < n) {
    k_dmask<<<(W + 255) / 256, 256, 0, st>>>(n, W, best, s.deg.as<int>(),
                                              s.dmask.as<unsigned long long>());
    const int rem = n - pos;
    k_icc<<<(rem + 3) / 4, 256, 0, st>>>(n, W, pos, best, adj, s.deg.as<int>(),
                                         s.dmask.as<unsigned long long>(), s.icc.as<int>());
    RPGO_HIP_CHECK(hipGetLastError());
    int first = n;
    RPGO_HIP_CHECK(hipMemcpyAsync(d_first, &first, sizeof(int), hipMemcpyHostToDevice, st));
    k_first_better<<<(rem + 255) / 256, 256, 0, st>>>(n, pos, best, s.icc.as<int>(), d_first);
    RPGO_HIP_CHECK(hipGetLastError());
    int got[1];
    RPGO_HIP_CHECK(hipMemcpyAsync(got, d_first, sizeof(int), hipMemcpyDeviceToHost, st));
    RPGO_HIP_CHECK(hipStreamSynchronize(st));
    if (got[0] >= n) break;
    int icc_c = 0;
    RPGO_HIP_CHECK(hipMemcpy(&icc_c, s.icc.as<int>() + got[0], sizeof(int), hipMemcpyDeviceToHost));
    win = got[0];
    win_best = best;
    best = icc_c;
    pos = win + 1;
  }
  if (mode == RPGO_CLIQUE_EXACT) {
    // the greedy size is a clique size that exists: the exact search's lower bound
    return clique_exact_run(ctx, adj, n, std::max(best, 1), idx_out, k_out, kExactNodeBudget);
  }
  if (win >= 0) {
    // dmask for the winner's threshold
    k_dmask<<<(W + 255) / 256, 256, 0, st>>>(n, W, win_best, s.deg.as<int>(),
                                              s.dmask.as<unsigned long long>());
    k_replay<<<1, 64, 0, st>>>(n, W, win, win_best, adj, s.dmask.as<unsigned long long>(),
                               s.out.as<int>(), s.best.as<int>() + 1);
    RPGO_HIP_CHECK(hipGetLastError());
    int icc_r = 0;
    RPGO_HIP_CHECK(hipMemcpyAsync(&icc_r, s.best.as<int>() + 1, sizeof(int), hipMemcpyDeviceToHost, st));
    RPGO_HIP_CHECK(hipStreamSynchronize(st));
    if (icc_r != best) {
      set_error("rpgo_max_clique: replay size %d differs from the scan's %d", icc_r, best);
      return RPGO_ERR_STATE;
    }
    if (idx_out && best > 0)
      RPGO_HIP_CHECK(hipMemcpy(idx_out, s.out.as<int>(), sizeof(int) * best, hipMemcpyDeviceToHost));
  }
  *k_out = best;
  return RPGO_OK;
}

}  // namespace rpgo
