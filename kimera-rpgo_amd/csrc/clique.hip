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
#include <mutex>
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

// first_reset (may be null): thread 0 also sets *first_reset = n, the "no
// winner" value k_first_better lowers, so the scan needs no host-to-device copy
__global__ void k_dmask(int n, int W, int best, const int* __restrict__ deg,
                        unsigned long long* __restrict__ dmask, int* __restrict__ first_reset) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (first_reset && w == 0) *first_reset = n;
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
// Word w = lane + 64 q, so the highest vertex lies in the highest q with a
// non-empty word, on the highest such lane: one ballot per q (scalar), then
// that lane's top bit by v_readlane, instead of a six-step shuffle reduction.
// Called by whole wavefronts (wave-uniform control flow).
__device__ __forceinline__ int wave_top_bit(const unsigned long long* S, int lane, int W) {
#pragma unroll
  for (int q = kMaxWPL - 1; q >= 0; --q) {
    if (64 * q >= W) continue;  // no words here (wave-uniform)
    const int w = lane + 64 * q;
    const unsigned long long lanes = __ballot(w < W && S[q] != 0ull);
    if (lanes) {
      const int L = 63 - __clzll(lanes);
      return __builtin_amdgcn_readlane(w * 64 + 63 - __clzll(S[q]), L);
    }
  }
  return -1;
}

__device__ __forceinline__ int wave_popcount(const unsigned long long* S) {
  int cnt = 0;
#pragma unroll
  for (int q = 0; q < kMaxWPL; ++q) cnt += __popcll(S[q]);
  return __ockl_wfred_add_i32(cnt);  // DPP wavefront sum
}

__device__ __forceinline__ unsigned long long readlane_u64(unsigned long long v, int l) {
  const unsigned int lo = __builtin_amdgcn_readlane((int)(unsigned int)v, l);
  const unsigned int hi = __builtin_amdgcn_readlane((int)(unsigned int)(v >> 32), l);
  return ((unsigned long long)hi << 32) | lo;
}

// Most picks of one block of the walk whose rows are loaded together.
constexpr int kWalkChunk = 16;

// S /\ N(p) for the next K picks p of P (lowest first; P loses them) in
// rows blk[p * W]: the rows' words are loaded together per word slot q.  A
// pick past the end of P repeats the first (a row ANDed twice changes
// nothing).  Lanes whose word of S is empty read word wt of the same rows,
// one shared line, and keep their zero; a slot empty on every lane is skipped.
template <int K>
__device__ __forceinline__ void and_pick_rows(unsigned long long* S, unsigned long long& P,
                                              const unsigned long long* __restrict__ blk, int W, int wt,
                                              int lane) {
  int pi[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    pi[k] = P ? __builtin_ctzll(P) : pi[0];
    P &= P - 1ull;
  }
#pragma unroll
  for (int q = 0; q < kMaxWPL; ++q) {
    if (64 * q >= W || !__ballot(lane + 64 * q < W && S[q] != 0ull)) continue;  // slot q empty on every lane
    const int w = lane + 64 * q;
    const int wa = (w < W && S[q] != 0ull) ? w : wt;
    unsigned long long r[K];
#pragma unroll
    for (int k = 0; k < K; ++k) r[k] = blk[(size_t)pi[k] * W + wa];
    unsigned long long a = S[q];
#pragma unroll
    for (int k = 0; k < K; ++k) a &= r[k];
    S[q] = a;
  }
}

// Greedy size icc(c, best) of candidate c, one wavefront (S as a bitset in
// registers, word w = lane + 64 q; dmask holds the deg >= best filter, LDS or
// global).  The walk picks the highest vertex m of S and keeps S /\ N(m)
// until S is empty; it is taken a word of S at a time.  While S
// keeps a vertex of its highest non-empty word wt, every pick lies in wt: the
// vertices of S outside wt are all lower.  So the picks inside wt follow from
// the 64 x 64 diagonal block of the adjacency alone (one row word per vertex,
// loaded at once, the walk then runs on scalars by v_readlane), and the word
// leaves S when its last vertex is picked or removed.  The other words of S
// keep the common neighbours of those picks, whose rows are loaded
// kWalkChunk at a time with no load waiting on another.  A dense block then
// costs two load round trips per 64 vertices instead of one per pick.  Only
// icc > best matters to the scan, so the walk stops early (INT_MIN) after a
// chunk of rows once it cannot get there (the word's picks are all counted,
// and every later pick removes at least itself from S, so at most |S| + 1
// increments remain); a candidate that beats best never stops early and
// returns its exact size.
__device__ __forceinline__ int icc_walk(int W, int c, const unsigned long long* __restrict__ adj,
                                        const unsigned long long* dmask, int lane, int best) {
  unsigned long long S[kMaxWPL];
#pragma unroll
  for (int q = 0; q < kMaxWPL; ++q) {
    const int w = lane + 64 * q;
    S[q] = (w < W) ? (adj[(size_t)c * W + w] & dmask[w] &
                      ((w == (c >> 6)) ? ~(1ull << (c & 63)) : ~0ull))
                   : 0ull;
  }
  if (wave_popcount(S) + 1 <= best) return INT_MIN;
  int picks = 0;
  for (;;) {
    // highest non-empty word wt of S and its vertices U (wave-uniform)
    int wt = -1;
    unsigned long long U = 0ull;
#pragma unroll
    for (int q = kMaxWPL - 1; q >= 0; --q) {
      if (64 * q >= W) continue;  // no words here (wave-uniform)
      const unsigned long long lanes = __ballot(lane + 64 * q < W && S[q] != 0ull);
      if (lanes) {
        const int L = 63 - __clzll(lanes);
        wt = 64 * q + L;
        U = readlane_u64(S[q], L);
        break;
      }
    }
    if (wt < 0) return picks + 1;  // S held only the seed c
    const unsigned long long* blk = adj + (size_t)64 * wt * W;  // rows 64 wt .. 64 wt + 63
    // lane k: word wt of row 64 wt + k for k in U (a lone vertex is its own pick);
    // the other lanes read the top vertex's word and drop it
    unsigned long long dw = 0ull;
    if (U & (U - 1)) {
      const bool in = (U >> lane) & 1ull;
      const unsigned long long v = blk[(size_t)(in ? lane : 63 - __clzll(U)) * W + wt];
      dw = in ? v : 0ull;
    }
    unsigned long long P = 0ull;  // picks in word wt, highest first
    while (U) {
      const int p = 63 - __clzll(U);
      P |= 1ull << p;
      U &= readlane_u64(dw, p) & ~(1ull << p);
    }
    picks += __popcll(P);
#pragma unroll
    for (int q = 0; q < kMaxWPL; ++q)
      if (lane + 64 * q == wt) S[q] = 0ull;
    // S /\ N(p) over the picks p, up to kWalkChunk rows at a time; a short
    // remainder takes the next power of two, so few loads are wasted (the
    // walk is bound by the loads it issues)
    while (P) {
      const int np = __popcll(P);
      if (np > 8) and_pick_rows<kWalkChunk>(S, P, blk, W, wt, lane);
      else if (np > 4) and_pick_rows<8>(S, P, blk, W, wt, lane);
      else if (np > 2) and_pick_rows<4>(S, P, blk, W, wt, lane);
      else if (np == 2) and_pick_rows<2>(S, P, blk, W, wt, lane);
      else and_pick_rows<1>(S, P, blk, W, wt, lane);
      // every pick of word wt is already counted and S only shrinks
      if (picks + wave_popcount(S) + 1 <= best) return INT_MIN;
    }
  }
}

// Picks whose rows a replay stages in LDS at once: kReplayChunk x kMaxWPL x
// 64 words of 8 B (32 KB) per replaying wavefront.
constexpr int kReplayChunk = 8;
constexpr int kReplayLds = kReplayChunk * kMaxWPL * 64;

// Replay of the winning candidate c: the walk of icc_walk, one word of S at a
// time, that also writes the reference's buffer positions 0..s0-1 (position p
// keeps the value last written to it) and returns the greedy size.  One
// wavefront; rl is its LDS staging area (kReplayLds words).
// Step t (pick m_t, S_t -> S_{t+1}) leaves the list at s_{t+1} = 1 + |S_{t+1}|
// entries (the seed stays in front) and positions [s_{t+1}, s_t) keep S_t's
// entries: position p holds S_t's element of rank p - 1 in ascending order.
// The writes need |S_{t+1}| after every pick, so within word wt the picks'
// rows are staged in LDS kReplayChunk at a time (loads in flight together);
// the steps then run on registers and LDS: S_t is U_t (the rest of word wt,
// above every other vertex of S_t) and the lower words L_t.
__device__ __forceinline__ int replay_walk(int W, int c, const unsigned long long* __restrict__ adj,
                                           const unsigned long long* dmask, int* __restrict__ out, int lane,
                                           unsigned long long* rl, int* s0 = nullptr) {
  unsigned long long S[kMaxWPL];
#pragma unroll
  for (int q = 0; q < kMaxWPL; ++q) {
    const int w = lane + 64 * q;
    S[q] = (w < W) ? (adj[(size_t)c * W + w] & dmask[w] &
                      ((w == (c >> 6)) ? ~(1ull << (c & 63)) : ~0ull))
                   : 0ull;
  }
  if (lane == 0) out[0] = c;
  int s_t = 1 + wave_popcount(S);  // list size before the step
  if (s0) *s0 = s_t;                // positions 0 .. s_t - 1 are written
  int picks = 0;
  for (;;) {
    int wt = -1;
    unsigned long long U = 0ull;
#pragma unroll
    for (int q = kMaxWPL - 1; q >= 0; --q) {
      if (64 * q >= W) continue;
      const unsigned long long lanes = __ballot(lane + 64 * q < W && S[q] != 0ull);
      if (lanes) {
        const int L = 63 - __clzll(lanes);
        wt = 64 * q + L;
        U = readlane_u64(S[q], L);
        break;
      }
    }
    // S held only the seed: the last step leaves the list empty (s_t == 1,
    // nothing is written)
    if (wt < 0) return picks + 1;
#pragma unroll
    for (int q = 0; q < kMaxWPL; ++q)
      if (lane + 64 * q == wt) S[q] = 0ull;  // S is L_t from here
    const unsigned long long* blk = adj + (size_t)64 * wt * W;
    unsigned long long dw = 0ull;
    if (U & (U - 1)) {
      const bool in = (U >> lane) & 1ull;
      const unsigned long long v = blk[(size_t)(in ? lane : 63 - __clzll(U)) * W + wt];
      dw = in ? v : 0ull;
    }
    unsigned long long P = 0ull, Ug = U;
    while (Ug) {
      const int p = 63 - __clzll(Ug);
      P |= 1ull << p;
      Ug &= readlane_u64(dw, p) & ~(1ull << p);
    }
    while (P) {
      // the next kReplayChunk picks, highest first (a short chunk repeats
      // its first pick: ANDing a row twice changes nothing)
      int pk[kReplayChunk];
      int nk = 0;
#pragma unroll
      for (int k = 0; k < kReplayChunk; ++k) {
        pk[k] = P ? 63 - __clzll(P) : pk[0];
        if (P) {
          P &= ~(1ull << pk[k]);
          ++nk;
        }
      }
      // per word q: the chunk's row words loaded together, kept in LDS for
      // the writes (word lane + 64 q of pick k at rl[(k kMaxWPL + q) 64 +
      // lane]; each lane reads back only its own), and L after each pick with
      // its size; the kReplayChunk wave sums are then taken together
      unsigned long long L0[kMaxWPL];
      int cnt[kReplayChunk];
#pragma unroll
      for (int k = 0; k < kReplayChunk; ++k) cnt[k] = 0;
#pragma unroll
      for (int q = 0; q < kMaxWPL; ++q) {
        L0[q] = S[q];
        if (64 * q >= W) continue;
        const int w = lane + 64 * q;
        const int wa = (w < W && S[q] != 0ull) ? w : wt;  // dead lanes: one shared line
        unsigned long long r[kReplayChunk];
#pragma unroll
        for (int k = 0; k < kReplayChunk; ++k) r[k] = blk[(size_t)pk[k] * W + wa];
        unsigned long long a = S[q];
#pragma unroll
        for (int k = 0; k < kReplayChunk; ++k) {
          rl[(k * kMaxWPL + q) * 64 + lane] = r[k];
          a &= r[k];
          cnt[k] += __popcll(a);
        }
        S[q] = a;
      }
#pragma unroll
      for (int k = 0; k < kReplayChunk; ++k) cnt[k] = __ockl_wfred_add_i32(cnt[k]);
#pragma unroll
      for (int k = 0; k < kReplayChunk; ++k) {
        if (k >= nk) break;
        const int m = 63 - __clzll(U);  // the k-th staged pick
        const unsigned long long Un = U & readlane_u64(dw, m) & ~(1ull << m);
        const int s_next = 1 + __popcll(Un) + cnt[k];
        // positions [s_next, s_t): the top of S_t, first U_t (highest bit
        // first), then L_t, rebuilt from the staged rows when reached
        int p = s_t - 1;
        unsigned long long Ut = U;
        for (; p >= s_next && Ut; --p) {
          const int t = 63 - __clzll(Ut);
          if (lane == 0) out[p] = 64 * wt + t;
          Ut &= ~(1ull << t);
        }
        if (p >= s_next) {
          unsigned long long Lp[kMaxWPL];
#pragma unroll
          for (int q = 0; q < kMaxWPL; ++q) {
            Lp[q] = L0[q];
#pragma unroll
            for (int j = 0; j < k; ++j)
              if (64 * q < W) Lp[q] &= rl[(j * kMaxWPL + q) * 64 + lane];
          }
          for (; p >= s_next; --p) {
            const int t = wave_top_bit(Lp, lane, W);
            if (lane == 0) out[p] = t;
            const int w = t >> 6;
#pragma unroll
            for (int q = 0; q < kMaxWPL; ++q)
              if (lane + 64 * q == w) Lp[q] &= ~(1ull << (t & 63));
          }
        }
        U = Un;
        s_t = s_next;
        ++picks;
      }
    }
  }
}

// A walk against best stays inside the vertices with deg >= best (dmask): S
// starts within it and only shrinks, and the candidates that walk pass the
// same filter.  The rounds of the single-graph scan therefore walk the
// subgraph of those vertices, renumbered in the same order (so "highest"
// picks the same vertex), with rows of ceil(n'/64) words instead of W: the
// same sizes for fewer row bytes, which bound the rounds.
// One workgroup of 512: new indices by a scan of the dmask words' popcounts.
__global__ void __launch_bounds__(512) k_compact_index(int n, int W, const unsigned long long* __restrict__ dmask,
                                                       int* __restrict__ cidx, int* __restrict__ ccol,
                                                       unsigned long long* __restrict__ cdmask,
                                                       int* __restrict__ cn) {
  __shared__ int pre[512];
  const int t = threadIdx.x;
  const unsigned long long b = t < W ? dmask[t] : 0ull;
  pre[t] = __popcll(b);
  __syncthreads();
  for (int d = 1; d < 512; d <<= 1) {  // inclusive scan
    const int v = t >= d ? pre[t - d] : 0;
    __syncthreads();
    pre[t] += v;
    __syncthreads();
  }
  const int total = pre[511];
  int r = pre[t] - __popcll(b);
  if (t < W) {
    for (int k = 0; k < 64; ++k) {
      const int v = 64 * t + k;
      if (v >= n) break;
      if ((b >> k) & 1ull) {
        cidx[v] = r;
        ccol[r++] = v;
      } else {
        cidx[v] = -1;
      }
    }
  }
  if (t < W) {
    const int lo = 64 * t;
    cdmask[t] = total >= lo + 64 ? ~0ull : (total > lo ? (1ull << (total - lo)) - 1ull : 0ull);
  }
  if (t == 0) *cn = total;
}

// word w' of new row v': bits of old row ccol[v'] at old columns
// ccol[64 w' .. 64 w' + 63] (ascending, so the old words come in order)
__global__ void k_compact_adj(int n, int W, const unsigned long long* __restrict__ adj,
                              const int* __restrict__ ccol, const int* __restrict__ cn,
                              unsigned long long* __restrict__ cadj) {
  const int nc = *cn, Wc = (nc + 63) / 64;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nc * Wc) return;
  const int vr = (int)(t / Wc), wr = (int)(t % Wc);
  const unsigned long long* row = adj + (size_t)ccol[vr] * W;
  unsigned long long out = 0ull, cur = 0ull;
  int cw = -1;
  for (int k = 0; k < 64; ++k) {
    const int u2 = 64 * wr + k;
    if (u2 >= nc) break;
    const int u = ccol[u2];
    if ((u >> 6) != cw) {
      cw = u >> 6;
      cur = row[cw];
    }
    out |= ((cur >> (u & 63)) & 1ull) << k;
  }
  cadj[(size_t)vr * Wc + wr] = out;
}

// k_icc on the compacted graph; results by old index.
__global__ void __launch_bounds__(256) k_icc_compact(int n, int pos, int best, const int* __restrict__ deg,
                                                     const int* __restrict__ cidx, const int* __restrict__ cn,
                                                     const unsigned long long* __restrict__ cadj,
                                                     const unsigned long long* __restrict__ cdmask,
                                                     int* __restrict__ icc_out) {
  const int c = pos + blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= n) return;
  const int lane = threadIdx.x & 63;
  if (best > deg[c]) {  // pruning 1
    if (lane == 0) icc_out[c] = INT_MIN;
    return;
  }
  const int icc = icc_walk((*cn + 63) / 64, cidx[c], cadj, cdmask, lane, best);
  if (lane == 0) icc_out[c] = icc;
}

// The winner's replay on the compacted graph (its dmask is deg >= the best
// the winner beat, which the replay walk never leaves); the buffer holds new
// indices until k_remap_out.  res[0] = greedy size, res[1] = positions written.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2)))
k_replay_compact(int c, const int* __restrict__ cidx, const int* __restrict__ cn,
                 const unsigned long long* __restrict__ cadj, const unsigned long long* __restrict__ cdmask,
                 int* __restrict__ out, int* __restrict__ res) {
  __shared__ unsigned long long rl[kReplayLds];
  int s0 = 0;
  const int icc = replay_walk((*cn + 63) / 64, cidx[c], cadj, cdmask, out, threadIdx.x, rl, &s0);
  if (threadIdx.x == 0) {
    res[0] = icc;
    res[1] = s0;
  }
}

// buffer positions 0 .. res[1] - 1 back to old indices
__global__ void k_remap_out(int* __restrict__ out, const int* __restrict__ ccol, const int* __restrict__ res) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < res[1]) out[p] = ccol[out[p]];
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
  const int icc = icc_walk(W, c, adj, dmask, lane, best);
  if (lane == 0) icc_out[c] = icc;
}

__global__ void k_first_better(int n, int pos, int best, const int* __restrict__ icc,
                               int* __restrict__ first) {
  const int c = pos + blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  if (icc[c] > best) atomicMin(first, c);
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2))) k_replay(int n, int W, int c, int best,
                                               const unsigned long long* __restrict__ adj,
                                               const unsigned long long* __restrict__ dmask,
                                               int* __restrict__ out, int* __restrict__ icc_out) {
  __shared__ unsigned long long rl[kReplayLds];
  const int icc = replay_walk(W, c, adj, dmask, out, threadIdx.x, rl);
  if (threadIdx.x == 0) *icc_out = icc;
  (void)best;
  (void)n;
}

// ---------------------------------------------------------------------------
// Batched heuristic: one workgroup of NW wavefronts per graph runs the whole
// sequential scan of FMC::maxCliqueHeu on the device.  Candidates are taken
// in rounds of NW consecutive ones (one per wavefront) against the current
// best; the first candidate of the first round that beats best is the one
// the sequential scan takes, and the scan resumes after it.  The winner is
// replayed by wavefront 0.  Graph g's adjacency starts at adj_off[g] with m[g]
// rows of ceil(m[g]/64) words; its degrees and clique go to lc_off[g].
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Chip-wide rounds of the batched scan.  One workgroup per graph leaves the
// largest graphs' scans on one CU each.  A round instead evaluates, for every
// graph still scanning, a window of kCbWindow candidates from its position
// with its current best, one wavefront each over the whole chip.  The first
// candidate of the window whose greedy size beats best is the scan's next
// winner: every candidate before it saw the same best in the sequential scan
// and did not beat it.  The position moves past the winner (or past the
// window), and candidates after a winner are evaluated again next round with
// the new best.
// ---------------------------------------------------------------------------
// 256 candidates per graph per round: on the 256-graph batch (up to ~6,000
// accepted loop closures per graph) the clique stage took 22.4 / 19.5 / 21.4 /
// 27.0 ms with windows of 128 / 256 / 512 / 1,024 on MI355X (wider windows
// re-evaluate more candidates after each winner, narrower ones need more rounds)
constexpr int kCbWindow = 256;   // candidates per graph per round
constexpr int kCbExtraRounds = 8;  // rounds for winners, beyond one per window of the largest
                                   // graph; the per-graph scan finishes what is left

struct CbState {
  int best, pos, win, win_best;
};

__global__ void k_cb_init(int n_graphs, CbState* __restrict__ st) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n_graphs) st[g] = CbState{-1, 0, -1, -1};
}

// deg[v] of every vertex of every graph (vertex slots indexed like the loop
// closures; slots beyond a graph's accepted count are left alone)
__global__ void k_cb_degrees(long long n_slots, const int* __restrict__ slot_graph,
                             const long long* __restrict__ adj_off, const long long* __restrict__ lc_off,
                             const int* __restrict__ m_arr, const unsigned long long* __restrict__ adj_all,
                             int* __restrict__ deg_all) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_slots) return;
  const int g = slot_graph[t];
  const int v = (int)(t - lc_off[g]), n = m_arr[g];
  if (v >= n) return;
  const int W = (n + 63) / 64;
  const unsigned long long* A = adj_all + adj_off[g];
  int d = 0;
  for (int w = 0; w < W; ++w) {
    unsigned long long b = A[(size_t)v * W + w];
    if (w == W - 1 && (n & 63)) b &= (1ull << (n & 63)) - 1ull;
    if (w == (v >> 6)) b &= ~(1ull << (v & 63));
    d += __popcll(b);
  }
  deg_all[t] = d;
}

// degree filter deg >= best of every graph still scanning; word w of graph g
// at dm[lc_off[g] + w]
__global__ void k_cb_mask(long long n_slots, const int* __restrict__ slot_graph,
                          const long long* __restrict__ lc_off, const int* __restrict__ m_arr,
                          const int* __restrict__ deg_all, const CbState* __restrict__ st,
                          unsigned long long* __restrict__ dm) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_slots) return;
  const int g = slot_graph[t];
  const int w = (int)(t - lc_off[g]), n = m_arr[g];
  if (w >= (n + 63) / 64 || st[g].pos >= n) return;
  const int best = st[g].best;
  const int* deg = deg_all + lc_off[g];
  unsigned long long b = 0ull;
  for (int k = 0; k < 64; ++k) {
    const int u = w * 64 + k;
    if (u < n && deg[u] >= best) b |= 1ull << k;
  }
  dm[t] = b;
}

// one wavefront per (graph, window slot): greedy size of candidate pos + k.
// Workgroup b runs on XCD x = b % 8; the XCD's i-th workgroup takes window
// quarter i % (kCbWindow / 4) of the (i / (kCbWindow / 4))-th graph of XCD x's
// group in xorder, so each graph's walks share one L2.
__global__ void __launch_bounds__(256) k_cb_icc(int n_graphs, const long long* __restrict__ adj_off,
                                                const long long* __restrict__ lc_off,
                                                const int* __restrict__ m_arr,
                                                const unsigned long long* __restrict__ adj_all,
                                                const int* __restrict__ deg_all,
                                                const unsigned long long* __restrict__ dm,
                                                const CbState* __restrict__ st, const int* __restrict__ xorder,
                                                int* __restrict__ icc_out) {
  constexpr int kBlkPerGraph = kCbWindow / 4;
  const int x = blockIdx.x & 7;
  const int i = (int)(blockIdx.x >> 3);
  const int gi = xorder[n_graphs + x] + i / kBlkPerGraph;
  if (gi >= xorder[n_graphs + x + 1]) return;  // beyond this XCD's group
  const int g = xorder[gi];
  const int k = (i % kBlkPerGraph) * 4 + (threadIdx.x >> 6);
  const long long t = (long long)g * kCbWindow + k;
  const int n = m_arr[g], pos = st[g].pos, best = st[g].best;
  const int c = pos + k;
  if (pos >= n || c >= n) return;  // the select step reads only slots below n - pos
  const int lane = threadIdx.x & 63;
  int icc = INT_MIN;
  // best < 1: the window's first candidate wins (every greedy size is >= 1),
  // so the others need no walk
  if (!(best > deg_all[lc_off[g] + c]) && (best >= 1 || k == 0))  // pruning 1
    icc = icc_walk((n + 63) / 64, c, adj_all + adj_off[g], dm + lc_off[g], lane, best);
  if (lane == 0) icc_out[t] = icc;
}

// one wavefront per graph: the window's first winner (by ballot over 64
// slots at a time), or the next window
__global__ void __launch_bounds__(256) k_cb_select(int n_graphs, const int* __restrict__ m_arr,
                                                   const int* __restrict__ icc, CbState* __restrict__ st) {
  const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (g >= n_graphs) return;
  CbState s = st[g];
  const int n = m_arr[g];
  if (s.pos >= n) return;
  const int span = min(kCbWindow, n - s.pos);
  const int* row = icc + (size_t)g * kCbWindow;
  int k = span;
  for (int base = 0; base < span; base += 64) {
    const int kk = base + lane;
    const unsigned long long b = __ballot(kk < span && row[kk] > s.best);
    if (b) {
      k = base + __builtin_ctzll(b);
      break;
    }
  }
  if (lane != 0) return;
  if (k < span) {
    s.win = s.pos + k;
    s.win_best = s.best;
    s.best = row[k];
    s.pos = s.pos + k + 1;
  } else {
    s.pos += span;
  }
  st[g] = s;
}

template <int NW>
__global__ void __launch_bounds__(64 * NW) k_clique_batch(int n_graphs, const long long* __restrict__ adj_off,
                                                          const long long* __restrict__ lc_off,
                                                          const int* __restrict__ m_arr,
                                                          const unsigned long long* __restrict__ adj_all,
                                                          int* __restrict__ deg_all, int* __restrict__ out_all,
                                                          int* __restrict__ k_out,
                                                          const CbState* __restrict__ start) {
  __shared__ unsigned long long dm[64 * kMaxWPL];
  __shared__ unsigned long long rl[kReplayLds];  // wavefront 0's replay
  __shared__ int s_icc[NW];
  __shared__ int s_win, s_win_icc;
  const int g = blockIdx.x;
  if (g >= n_graphs) return;
  const int n = m_arr[g];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  if (n <= 0) {
    if (tid == 0) k_out[g] = -1;  // maxCliqueHeu on an empty graph
    return;
  }
  const int W = (n + 63) / 64;
  if (W > 64 * kMaxWPL) {
    if (tid == 0) k_out[g] = INT_MIN;  // rejected on the host before launch
    return;
  }
  const unsigned long long* A = adj_all + adj_off[g];
  int* deg = deg_all + lc_off[g];
  int* out = out_all + lc_off[g];
  // after the chip-wide rounds the degrees are already there (k_cb_degrees)
  for (int v = start ? n : tid; v < n; v += 64 * NW) {
    int d = 0;
    for (int w = 0; w < W; ++w) {
      unsigned long long b = A[(size_t)v * W + w];
      if (w == W - 1 && (n & 63)) b &= (1ull << (n & 63)) - 1ull;
      if (w == (v >> 6)) b &= ~(1ull << (v & 63));
      d += __popcll(b);
    }
    deg[v] = d;
  }
  __syncthreads();
  // the scan so far (chip-wide rounds), or its start
  int best = -1, pos = 0, win = -1, win_best = -1;
  if (start) {
    best = start[g].best;
    pos = start[g].pos;
    win = start[g].win;
    win_best = start[g].win_best;
  }
  for (;;) {
    for (int w = tid; w < W; w += 64 * NW) {
      unsigned long long b = 0ull;
      for (int k = 0; k < 64; ++k) {
        const int u = w * 64 + k;
        if (u < n && deg[u] >= best) b |= 1ull << k;
      }
      dm[w] = b;
    }
    __syncthreads();
    bool found = false;
    for (int base = pos; base < n; base += NW) {
      const int c = base + wid;
      int icc = INT_MIN;
      // pruning 1; with best < 1 the first candidate wins (greedy sizes are >= 1)
      if (c < n && !(best > deg[c]) && (best >= 1 || wid == 0)) icc = icc_walk(W, c, A, dm, lane, best);
      if (lane == 0) s_icc[wid] = icc;
      __syncthreads();
      if (tid == 0) {
        int first = -1, fi = 0;
        for (int k = 0; k < NW; ++k)
          if (s_icc[k] > best) {
            first = base + k;
            fi = s_icc[k];
            break;
          }
        s_win = first;
        s_win_icc = fi;
      }
      __syncthreads();
      const int first = s_win;
      if (first >= 0) {
        win = first;
        win_best = best;
        best = s_win_icc;
        pos = first + 1;
        found = true;
        break;
      }
    }
    __syncthreads();
    if (!found) break;
  }
  if (win >= 0) {
    for (int w = tid; w < W; w += 64 * NW) {
      unsigned long long b = 0ull;
      for (int k = 0; k < 64; ++k) {
        const int u = w * 64 + k;
        if (u < n && deg[u] >= win_best) b |= 1ull << k;
      }
      dm[w] = b;
    }
    __syncthreads();
    if (wid == 0) {
      const int icc = replay_walk(W, win, A, dm, out, lane, rl);
      if (lane == 0) k_out[g] = icc == best ? best : INT_MIN;
    }
  } else if (tid == 0) {
    k_out[g] = best;
  }
}

size_t clique_spread_state_bytes(int n_graphs) { return sizeof(CbState) * (size_t)(n_graphs > 0 ? n_graphs : 1); }
size_t clique_spread_icc_bytes(int n_graphs) {
  return sizeof(int) * (size_t)kCbWindow * (size_t)(n_graphs > 0 ? n_graphs : 1);
}

int clique_batch_launch(rpgo_ctx* ctx, int n_graphs, int waves, const long long* d_adj_off,
                        const long long* d_lc_off, const int* d_m, const unsigned long long* adj, int* deg,
                        int* out_idx, int* out_k, const CliqueSpread* spread) {
  hipStream_t st = ctx->stream;
  if (waves != 4 && waves != 8 && waves != 16) {
    set_error("clique_batch_launch: unsupported shape (%d wavefronts)", waves);
    return RPGO_ERR_ARG;
  }
  CbState* cs = nullptr;
  if (spread && spread->n_slots > 0) {
    cs = static_cast<CbState*>(spread->state);
    const long long ns = spread->n_slots;
    const unsigned sg = (unsigned)((ns + 255) / 256);
    k_cb_init<<<(n_graphs + 255) / 256, 256, 0, st>>>(n_graphs, cs);
    k_cb_degrees<<<sg, 256, 0, st>>>(ns, spread->slot_graph, d_adj_off, d_lc_off, d_m, adj, deg);
    const unsigned gi = (unsigned)(8ll * spread->xmax * (kCbWindow / 4));
    const int rounds = (int)((spread->max_slots + kCbWindow - 1) / kCbWindow) + kCbExtraRounds;
    for (int r = 0; r < rounds; ++r) {
      k_cb_mask<<<sg, 256, 0, st>>>(ns, spread->slot_graph, d_lc_off, d_m, deg, cs, spread->dm);
      k_cb_icc<<<gi, 256, 0, st>>>(n_graphs, d_adj_off, d_lc_off, d_m, adj, deg, spread->dm, cs, spread->xorder,
                                   spread->icc);
      k_cb_select<<<(n_graphs + 3) / 4, 256, 0, st>>>(n_graphs, d_m, spread->icc, cs);
    }
    RPGO_HIP_CHECK(hipGetLastError());
  }
#define RPGO_CB(NW) \
  k_clique_batch<NW><<<n_graphs, 64 * NW, 0, st>>>(n_graphs, d_adj_off, d_lc_off, d_m, adj, deg, out_idx, out_k, cs)
  if (waves == 4) RPGO_CB(4);
  else if (waves == 8) RPGO_CB(8);
  else RPGO_CB(16);
#undef RPGO_CB
  RPGO_HIP_CHECK(hipGetLastError());
  return RPGO_OK;
}

int clique_walk_chunk() { return kWalkChunk; }

int clique_max_vertices() { return 64 * 64 * kMaxWPL; }

// Working memory of one heuristic scan over n vertices (W = ceil(n/64)):
// deg n ints, dmask W words, icc n ints, out n + 1 ints, best 4 ints.
struct ScanBufs {
  int* deg;
  unsigned long long* dmask;
  int* icc;
  int* out;
  int* best;
  // compacted graph of the rounds (all null: the rounds walk the full rows)
  int* cidx = nullptr;                  // n: new index of a vertex with deg >= best
  int* ccol = nullptr;                  // n: old index of new vertex v'
  unsigned long long* cdmask = nullptr;  // W: every new vertex
  unsigned long long* cadj = nullptr;    // n x W words (rows of ceil(n'/64) words)
  int* cn = nullptr;                     // n'
};

// Greedy scan over candidates pos0 .. n-1 starting from the bound best0;
// mode EXACT continues with the exact search, mode HEU_REF replays the winner.
// bufs == nullptr: the context's grow-only buffers (allocated on demand).
static int clique_scan(rpgo_ctx* ctx, const uint64_t* d_adj_u64, int n, int mode, int pos0, int best0,
                       int32_t* idx_out, int32_t* k_out, const ScanBufs* bufs = nullptr);

int clique_run(rpgo_ctx* ctx, const uint64_t* d_adj_u64, int n, int mode, int32_t* idx_out,
               int32_t* k_out) {
  return clique_scan(ctx, d_adj_u64, n, mode, 0, -1, idx_out, k_out);
}

// ---------------------------------------------------------------------------
// Allocation-free heuristic scan (rpgo_max_clique_bits_noalloc).  Its working
// memory is this file-scope device array, one copy per GPU, held until the
// process ends.  Host code reaches it only through the address
// hipGetSymbolAddress returns on each device (the array's name is a host-side
// stand-in, not a device address); that lookup runs once per device, from
// rpgo_ctx_create, with the device current.
// ---------------------------------------------------------------------------
constexpr int kNoallocMaxVertices = RPGO_NOALLOC_MAX_VERTICES;
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kNaDeg = 0;
constexpr size_t kNaDmask = kNaDeg + align256(sizeof(int) * kNoallocMaxVertices);
constexpr size_t kNaIcc = kNaDmask + align256(sizeof(unsigned long long) * ((kNoallocMaxVertices + 63) / 64));
constexpr size_t kNaOut = kNaIcc + align256(sizeof(int) * kNoallocMaxVertices);
constexpr size_t kNaBest = kNaOut + align256(sizeof(int) * (kNoallocMaxVertices + 1));
constexpr size_t kNaBytes = kNaBest + 256;
static_assert(kNoallocMaxVertices <= 64 * 64 * kMaxWPL, "no-alloc maximum beyond the register bitset");

static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int clique_noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("clique_noalloc_init: device %d beyond the %d supported", device, kMaxDevices);
    return RPGO_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return RPGO_OK;
  void* ptr = nullptr;
  // the caller (rpgo_ctx_create) has made `device` current
  RPGO_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = ptr;
  return RPGO_OK;
}

size_t clique_noalloc_bytes() { return kNaBytes; }

int clique_run_noalloc(rpgo_ctx* ctx, const uint64_t* d_adj_u64, int n, int32_t* idx_out, int32_t* k_out) {
  if (n > kNoallocMaxVertices) {
    set_error("rpgo_max_clique_bits_noalloc: %d vertices exceed the maximum %d", n, kNoallocMaxVertices);
    return RPGO_ERR_ARG;
  }
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (ctx->device >= 0 && ctx->device < kMaxDevices) base = g_scratch_ptr[ctx->device];
  }
  if (!base) {
    set_error("rpgo_max_clique_bits_noalloc: scratch of device %d not set up", ctx->device);
    return RPGO_ERR_STATE;
  }
  unsigned char* b = static_cast<unsigned char*>(base);
  const ScanBufs bufs{reinterpret_cast<int*>(b + kNaDeg), reinterpret_cast<unsigned long long*>(b + kNaDmask),
                      reinterpret_cast<int*>(b + kNaIcc), reinterpret_cast<int*>(b + kNaOut),
                      reinterpret_cast<int*>(b + kNaBest)};
  return clique_scan(ctx, d_adj_u64, n, RPGO_CLIQUE_HEU_REF, 0, -1, idx_out, k_out, &bufs);
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
                       int32_t* idx_out, int32_t* k_out, const ScanBufs* bufs) {
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
  ScanBufs b;
  if (bufs) {
    b = *bufs;
  } else {
    int rc;
    if ((rc = s.deg.reserve(sizeof(int) * n))) return rc;
    if ((rc = s.dmask.reserve(sizeof(unsigned long long) * W))) return rc;
    if ((rc = s.icc.reserve(sizeof(int) * n))) return rc;
    if ((rc = s.out.reserve(sizeof(int) * (n + 1)))) return rc;
    if ((rc = s.best.reserve(sizeof(int) * 4))) return rc;
    if ((rc = s.cidx.reserve(sizeof(int) * n))) return rc;
    if ((rc = s.ccol.reserve(sizeof(int) * n))) return rc;
    if ((rc = s.cdmask.reserve(sizeof(unsigned long long) * W))) return rc;
    if ((rc = s.cadj.reserve(sizeof(unsigned long long) * (size_t)n * W))) return rc;
    if ((rc = s.cn.reserve(sizeof(int)))) return rc;
    b = ScanBufs{s.deg.as<int>(), s.dmask.as<unsigned long long>(), s.icc.as<int>(), s.out.as<int>(),
                 s.best.as<int>(), s.cidx.as<int>(), s.ccol.as<int>(), s.cdmask.as<unsigned long long>(),
                 s.cadj.as<unsigned long long>(), s.cn.as<int>()};
  }
  hipStream_t st = ctx->stream;
  k_degrees<<<(n + 255) / 256, 256, 0, st>>>(n, W, adj, b.deg);
  RPGO_HIP_CHECK(hipGetLastError());
  int best = best0, pos = pos0, win = -1, win_best = -1;
  int* d_first = b.best;
  while (pos < n) {
    k_dmask<<<(W + 255) / 256, 256, 0, st>>>(n, W, best, b.deg, b.dmask, d_first);
    // best < 1: candidate pos wins (every greedy size is >= 1), so it alone is walked
    const int rem = best < 1 ? 1 : n - pos;
    if (best >= 1 && b.cadj && W > 64) {
      // the round's subgraph of deg >= best, renumbered in order (rows of at
      // most 64 words take one load per lane either way: not compacted)
      k_compact_index<<<1, 512, 0, st>>>(n, W, b.dmask, b.cidx, b.ccol, b.cdmask, b.cn);
      k_compact_adj<<<(unsigned)(((long long)n * W + 255) / 256), 256, 0, st>>>(n, W, adj, b.ccol, b.cn, b.cadj);
      k_icc_compact<<<(rem + 3) / 4, 256, 0, st>>>(n, pos, best, b.deg, b.cidx, b.cn, b.cadj, b.cdmask, b.icc);
    } else {
      k_icc<<<(rem + 3) / 4, 256, 0, st>>>(n, W, pos, best, adj, b.deg, b.dmask, b.icc);
    }
    RPGO_HIP_CHECK(hipGetLastError());
    k_first_better<<<(rem + 255) / 256, 256, 0, st>>>(n, pos, best, b.icc, d_first);
    RPGO_HIP_CHECK(hipGetLastError());
    int got[1];
    RPGO_HIP_CHECK(hipMemcpyAsync(got, d_first, sizeof(int), hipMemcpyDeviceToHost, st));
    RPGO_HIP_CHECK(hipStreamSynchronize(st));
    if (got[0] >= n) break;
    int icc_c = 0;
    RPGO_HIP_CHECK(hipMemcpyAsync(&icc_c, b.icc + got[0], sizeof(int), hipMemcpyDeviceToHost, st));
    RPGO_HIP_CHECK(hipStreamSynchronize(st));
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
    k_dmask<<<(W + 255) / 256, 256, 0, st>>>(n, W, win_best, b.deg, b.dmask, nullptr);
    if (win_best >= 1 && b.cadj && W > 64) {
      // on the subgraph of deg >= win_best, as the rounds
      k_compact_index<<<1, 512, 0, st>>>(n, W, b.dmask, b.cidx, b.ccol, b.cdmask, b.cn);
      k_compact_adj<<<(unsigned)(((long long)n * W + 255) / 256), 256, 0, st>>>(n, W, adj, b.ccol, b.cn, b.cadj);
      k_replay_compact<<<1, 64, 0, st>>>(win, b.cidx, b.cn, b.cadj, b.cdmask, b.out, b.best + 1);
      k_remap_out<<<(n + 255) / 256, 256, 0, st>>>(b.out, b.ccol, b.best + 1);
    } else {
      k_replay<<<1, 64, 0, st>>>(n, W, win, win_best, adj, b.dmask, b.out, b.best + 1);
    }
    RPGO_HIP_CHECK(hipGetLastError());
    int icc_r = 0;
    RPGO_HIP_CHECK(hipMemcpyAsync(&icc_r, b.best + 1, sizeof(int), hipMemcpyDeviceToHost, st));
    RPGO_HIP_CHECK(hipStreamSynchronize(st));
    if (icc_r != best) {
      set_error("rpgo_max_clique: replay size %d differs from the scan's %d", icc_r, best);
      return RPGO_ERR_STATE;
    }
    if (idx_out && best > 0) {
      RPGO_HIP_CHECK(hipMemcpyAsync(idx_out, b.out, sizeof(int) * best, hipMemcpyDeviceToHost, st));
      RPGO_HIP_CHECK(hipStreamSynchronize(st));
    }
  }
  *k_out = best;
  return RPGO_OK;
}

}  // namespace rpgo
