// FP32 PCM pairwise kernel (SE(3)) with an FP64 replay of near-threshold pairs.
//
// The reference forms the covariance of Trajectory::getBetween(a, c) as a
// difference of absolute chain covariances, Sigma_c - Ad Sigma_a Ad^T
// (GeometryUtils.h:148-149; LLT fallback :152-167).  Both terms grow with the
// trajectory, so the difference is meaningless in FP32.  In exact arithmetic
// the result is the covariance of the odometry segment between min(a,c) and
// max(a,c), expressed in the frame of max(a,c) (the LLT fallback covers a > c).
// This path builds that segment covariance from cancellation-free parts:
//
//   anchors every kB poses (A_p = p*kB);
//   Lcov_k = Seg(A_blk(k), k)      in frame k          (forward, within block)
//   Rcov_k = Seg(k, A_blk(k)+1)    in frame A_blk(k)+1 (suffix sum, within block)
//   Sanc(p, q) = Seg(A_p, A_q)     in frame A_q        (anchor table, FP64 sums)
//
//   Seg(lo, hi) in frame hi = Ad(Pl_hi^-1) M Ad(Pl_hi^-1)^T + Lcov_hi, with
//     same block : M = -Ad(Pl_lo) Lcov_lo Ad(Pl_lo)^T          (bounded cancellation)
//     cross block: M =  Ad(G) Rcov_lo Ad(G)^T + Sanc(blk(lo)+1, blk(hi)),
//                  G = T_{A_hi}^-1 T_{A_blk(lo)+1}
//
// Poses are kept relative to their block anchor (Pl_k = T_A^-1 T_k, FP32) and
// anchor poses in FP64, so relative poses never subtract large coordinates in
// FP32.  The anchor-to-anchor poses a segment needs (T_{A_p}^-1 T_{A_q} and G)
// are composed once per anchor pair in FP64, rounded to FP32 and stored beside
// Sanc, so the pair loop does no FP64 work.  The rest of the pair check
// (Pcm.h:703-717: compose with the newer loop closure, inverse, compose with
// the older one, compose with the second odometry segment, Mahalanobis norm
// of the Log) runs in FP32.
//
// A pair whose squared distance lies within a relative band of the squared
// threshold, whose Log is near pi, or whose covariance is not numerically PD
// is queued and replayed in FP64 with the reference's own formula and
// operation order (replay64.hip), so the adjacency matches the FP64 path bit
// for bit.  The band is recheck_margin plus a term for the reference's own
// FP64 cancellation error, which grows with the chain covariance.
#include <climits>

#include "geom32.h"
#include "replay64.h"
#include "rpgo_internal.h"

namespace rpgo {

constexpr int kB = 64;          // poses per anchor block
constexpr int kPoseRec = 60;    // floats per pose record
constexpr int kLcRec = 48;      // floats per loop-closure record
constexpr int kQueueCap = kRecheckQueueCap;

// pose record layout (floats): Pl[12] Lcov[21] Rcov[21] trc ti[3] pad[2]
// (ti: translation of Pl^-1)
struct PoseRec {
  P32 Pl;
  S6 L;
  S6 R;
  float trc;
  float ti[3];
  float pad[2];
};
static_assert(sizeof(PoseRec) == kPoseRec * 4, "pose record size");

// loop-closure record: Z[12] Zinv[12] S[21] a b ri
struct LcRec {
  P32 Z;
  P32 Zi;
  S6 S;
  int a;
  int b;
  int pad;
};
static_assert(sizeof(LcRec) == kLcRec * 4, "lc record size");

// Entry (r, q) of the nb x nb anchor-pair table, r <= q:
//   S   = Sanc(r, q): covariance of the odometry from anchor r to anchor q, in frame q;
//   rel = T_{A_{r-1}}^-1 T_{A_q} and G = T_{A_q}^-1 T_{A_r} (r >= 1), composed in
//         FP64 from the anchor poses and rounded once.
// A segment from block p to a later block q reads entry (p + 1, q): S, rel
// and G of one gather (192 B, two 128-byte lines).
struct alignas(64) AncRec {
  S6 S;
  P32 rel;
  P32 G;
  float pad[3];
};
static_assert(sizeof(AncRec) == 192, "anchor entry size");

// Pl^-1 of a pose record: the transpose of its rotation and the stored ti
__device__ __forceinline__ void rec_inv(const PoseRec& r, P32& o) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) o.R[i * 3 + j] = r.Pl.R[j * 3 + i];
  o.t[0] = r.ti[0], o.t[1] = r.ti[1], o.t[2] = r.ti[2];
}

__device__ __forceinline__ void s6_from36(const double* c, S6& s) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = i; j < 3; ++j) {
      s.A[s3(i, j)] = (float)c[i * 6 + j];
      s.C[s3(i, j)] = (float)c[(i + 3) * 6 + j + 3];
    }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) s.B[i * 3 + j] = (float)c[i * 6 + j + 3];
}

__device__ __forceinline__ void p32_from64(const double* p, P32& o) {
#pragma unroll
  for (int e = 0; e < 9; ++e) o.R[e] = (float)p[e];
  o.t[0] = (float)p[9]; o.t[1] = (float)p[10]; o.t[2] = (float)p[11];
}

__device__ __forceinline__ void add36(double* acc, const double* x) {
#pragma unroll
  for (int e = 0; e < 36; ++e) acc[e] += x[e];
}

// ---------------------------------------------------------------------------
// precompute
// ---------------------------------------------------------------------------

// NaN rotation covariances switch the whole run to the FP64 replay.
__global__ void k_any_no_rotation(int n, const double* __restrict__ rec, int stride, int flag_off,
                                  int* __restrict__ any) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n && rec[(size_t)k * stride + flag_off] == 0.0) atomicOr(any, 1);
}

// Anchor block p of one graph: local poses, forward Lcov, suffix Rcov, and
// the block's own segment (Bend) and span (Pblock), all accumulated in FP64.
__device__ __forceinline__ void block_local_body(int p, int n_poses, const double* __restrict__ delta,
                                                 const double* __restrict__ hinv, PoseRec* __restrict__ prec,
                                                 double* __restrict__ pl64, double* __restrict__ pblock64,
                                                 double* __restrict__ bend64) {
  using R = R64<3>;
  const int k0 = p * kB;
  const int k1 = min(k0 + kB, n_poses);  // poses of this block: [k0, k1)
  const bool has_next = k1 < n_poses;    // anchor A_{p+1} = k1 exists
  double P[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
  double L[36];
#pragma unroll
  for (int e = 0; e < 36; ++e) L[e] = 0.0;
  const int kend = has_next ? k1 : k1 - 1;  // last step index (exclusive) of the walk
  for (int k = k0;; ++k) {
    // record pose k (k < k1) or the continuation at the next anchor
    if (k < k1) {
      PoseRec r;
      p32_from64(P, r.Pl);
      s6_from36(L, r.L);
      s6_zero(r.R);
      r.trc = 0.f;
      {
        P32 Pi;
        p32_inv(r.Pl, Pi);
        r.ti[0] = Pi.t[0], r.ti[1] = Pi.t[1], r.ti[2] = Pi.t[2];
      }
      r.pad[0] = r.pad[1] = 0.f;
      prec[k] = r;
#pragma unroll
      for (int e = 0; e < 12; ++e) pl64[(size_t)k * 12 + e] = P[e];
    } else {
#pragma unroll
      for (int e = 0; e < 12; ++e) pblock64[(size_t)p * 12 + e] = P[e];
#pragma unroll
      for (int e = 0; e < 36; ++e) bend64[(size_t)p * 36 + e] = L[e];
      break;
    }
    if (k >= kend) break;
    // advance: P <- P * delta_k ; L <- H L H^T + Q_k
    const double* dk = delta + (size_t)k * R::PWC;
    double Pn[12], HLH[36];
    R::pose_compose(P, dk, Pn);
    congruence64<6>(hinv + (size_t)k * 36, L, HLH);
#pragma unroll
    for (int e = 0; e < 12; ++e) P[e] = Pn[e];
#pragma unroll
    for (int e = 0; e < 36; ++e) L[e] = HLH[e] + dk[12 + e];
  }
  if (!has_next) return;
  // suffix sums Rcov_k = sum_{m >= k} Ad(Pblock^-1 Pl_{m+1}) Q_m Ad(.)^T
  double Pbi[12];
  R::pose_inverse(pblock64 + (size_t)p * 12, Pbi);
  double acc[36];
#pragma unroll
  for (int e = 0; e < 36; ++e) acc[e] = 0.0;
  for (int m = k1 - 1; m >= k0; --m) {
    const double* Pm1 = (m + 1 < k1) ? pl64 + (size_t)(m + 1) * 12 : pblock64 + (size_t)p * 12;
    double G[12], Ad[36], term[36];
    R::pose_compose(Pbi, Pm1, G);
    R::adjoint(G, Ad);
    congruence64<6>(Ad, delta + (size_t)m * R::PWC + 12, term);
    add36(acc, term);
    S6 s;
    s6_from36(acc, s);
    prec[m].R = s;
  }
}

// one thread per anchor block
__global__ void __launch_bounds__(64) k_block_local(int n_poses, int n_blocks, const double* __restrict__ delta,
                              const double* __restrict__ hinv, PoseRec* __restrict__ prec,
                              double* __restrict__ pl64, double* __restrict__ pblock64,
                              double* __restrict__ bend64) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_blocks) return;
  block_local_body(p, n_poses, delta, hinv, prec, pl64, pblock64, bend64);
}

// Anchor poses T_{A_p} in FP64 (sequential over blocks).
__device__ __forceinline__ void anchor_poses_body(int n_blocks, const double* __restrict__ pose0,
                                                  const double* __restrict__ pblock64, double* __restrict__ ta) {
  double P[12];
#pragma unroll
  for (int e = 0; e < 12; ++e) P[e] = pose0[e];
  for (int p = 0; p < n_blocks; ++p) {
#pragma unroll
    for (int e = 0; e < 12; ++e) ta[(size_t)p * 12 + e] = P[e];
    if (p + 1 < n_blocks) {
      double Pn[12];
      R64<3>::pose_compose(P, pblock64 + (size_t)p * 12, Pn);
#pragma unroll
      for (int e = 0; e < 12; ++e) P[e] = Pn[e];
    }
  }
}

__global__ void k_anchor_poses(int n_blocks, const double* __restrict__ pose0,
                               const double* __restrict__ pblock64, double* __restrict__ ta) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  anchor_poses_body(n_blocks, pose0, pblock64, ta);
}

// term(r, q) = Ad(T_{A_q}^-1 T_{A_{r+1}}) Bend_r Ad(.)^T for r < q (FP64),
// entry t = q nb + r; also rel and G of anchor-pair entry (r + 1, q)
__device__ __forceinline__ void anchor_terms_body(long long t, int nb, const double* __restrict__ ta,
                                                  const double* __restrict__ bend64, double* __restrict__ terms,
                                                  AncRec* __restrict__ anc) {
  using R = R64<3>;
  const int q = (int)(t / nb), r = (int)(t % nb);
  if (r >= q) return;
  double qi[12], G[12], Ad[36], term[36];
  R::pose_inverse(ta + (size_t)q * 12, qi);
  R::pose_compose(qi, ta + (size_t)(r + 1) * 12, G);
  R::adjoint(G, Ad);
  congruence64<6>(Ad, bend64 + (size_t)r * 36, term);
#pragma unroll
  for (int e = 0; e < 36; ++e) terms[(size_t)t * 36 + e] = term[e];
  AncRec& e = anc[(size_t)(r + 1) * nb + q];
  p32_from64(G, e.G);
  double ri[12], rel[12];
  R::pose_inverse(ta + (size_t)r * 12, ri);
  R::pose_compose(ri, ta + (size_t)q * 12, rel);
  p32_from64(rel, e.rel);
  e.pad[0] = e.pad[1] = e.pad[2] = 0.f;
}

__global__ void __launch_bounds__(256) k_anchor_terms(int nb, const double* __restrict__ ta,
                               const double* __restrict__ bend64, double* __restrict__ terms,
                               AncRec* __restrict__ anc) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nb * nb) return;
  anchor_terms_body(t, nb, ta, bend64, terms, anc);
}

// Sanc(r, q) = sum_{r' = r}^{q-1} term(r', q)  (suffix sum over r for each q)
__device__ __forceinline__ void anchor_suffix_body(int q, int nb, const double* __restrict__ terms,
                                                   AncRec* __restrict__ anc) {
  double acc[36];
#pragma unroll
  for (int e = 0; e < 36; ++e) acc[e] = 0.0;
  S6 z;
  s6_zero(z);
  anc[(size_t)q * nb + q].S = z;
  for (int r = q - 1; r >= 0; --r) {
    add36(acc, terms + ((size_t)q * nb + r) * 36);
    S6 s;
    s6_from36(acc, s);
    anc[(size_t)r * nb + q].S = s;
  }
}

__global__ void k_anchor_suffix(int nb, const double* __restrict__ terms, AncRec* __restrict__ anc) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nb) return;
  anchor_suffix_body(q, nb, terms, anc);
}

// Chain covariance trace estimate of pose k (margin of the FP64 band).
__device__ __forceinline__ void pose_trace_body(int k, const AncRec* __restrict__ anc, PoseRec* __restrict__ prec) {
  const int p = k / kB;
  PoseRec& r = prec[k];
  P32 Pi;
  rec_inv(r, Pi);
  S6 T;
  s6_congruence(Pi, anc[p].S, T);  // Sanc(0, p)
  r.trc = s6_trace(T) + s6_trace(r.L);
}

__global__ void k_pose_trace(int n_poses, const AncRec* __restrict__ anc, PoseRec* __restrict__ prec) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_poses) return;
  pose_trace_body(k, anc, prec);
}

__device__ __forceinline__ void lc_record_body(int t, const int* __restrict__ cf, const int* __restrict__ ct,
                                               const double* __restrict__ clcm, LcRec* __restrict__ out) {
  const double* r = clcm + (size_t)t * R64<3>::PWC;
  LcRec o;
  p32_from64(r, o.Z);
  double zi[12];
  R64<3>::pose_inverse(r, zi);
  p32_from64(zi, o.Zi);
  s6_from36(r + 12, o.S);
  o.a = cf[t];
  o.b = ct[t];
  o.pad = 0;
  out[t] = o;
}

__global__ void k_lc_records(int m, const int* __restrict__ cf, const int* __restrict__ ct,
                             const double* __restrict__ clcm, LcRec* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  lc_record_body(t, cf, ct, clcm, out);
}

// ---------------------------------------------------------------------------
// pair check
// ---------------------------------------------------------------------------

struct Anchors {
  const AncRec* anc;  // nb x nb
  int nb;
};

// between(x, y) of the reference, x older-side key, y newer-side key:
// pose T_x^-1 T_y and covariance Seg(min, max) in frame max.  Returns the
// translation norm of the pose (for the error band).
__device__ __forceinline__ float between32(const Anchors& an, int x, const PoseRec& rx, int y,
                                           const PoseRec& ry, P32& pose, S6& cov) {
  const bool xlo = x <= y;
  const int lo = xlo ? x : y, hi = xlo ? y : x;
  const PoseRec& rl = xlo ? rx : ry;
  const PoseRec& rh = xlo ? ry : rx;
  const int plo = lo / kB, phi = hi / kB;
  const bool same = plo == phi;
  // T_{A_plo}^-1 T_{A_phi}: the identity within one block
  P32 rel;
  const AncRec* e = an.anc + (size_t)(same ? phi : plo + 1) * an.nb + phi;
  if (same) {
#pragma unroll
    for (int k = 0; k < 9; ++k) rel.R[k] = (k % 4 == 0) ? 1.f : 0.f;
    rel.t[0] = rel.t[1] = rel.t[2] = 0.f;
  } else {
    rel = e->rel;
  }
  P32 reli;
  p32_inv(rel, reli);
  // pose
  P32 xi_, tmp;
  rec_inv(rx, xi_);
  p32_mul(xi_, xlo ? rel : reli, tmp);
  p32_mul(tmp, ry.Pl, pose);
  // covariance
  S6 M;
  if (same) {
    s6_congruence(rl.Pl, rl.L, M);
#pragma unroll
    for (int k = 0; k < 6; ++k) M.A[k] = -M.A[k], M.C[k] = -M.C[k];
#pragma unroll
    for (int k = 0; k < 9; ++k) M.B[k] = -M.B[k];
  } else {
    s6_congruence_add(e->G, rl.R, e->S, M);
  }
  P32 hinv;
  rec_inv(rh, hinv);
  s6_congruence_add(hinv, M, rh.L, cov);
  return sqrtf(pose.t[0] * pose.t[0] + pose.t[1] * pose.t[1] + pose.t[2] * pose.t[2]);
}

// compose of PoseWithCovariance with a measurement (Z, Zinv, S):
// pose T Z, covariance Ad(Z^-1) C Ad(Z^-1)^T + S
__device__ __forceinline__ void compose_meas(const P32& T, const S6& C, const LcRec& z, P32& To,
                                             S6& Co) {
  p32_mul(T, z.Z, To);
  s6_congruence_add(z.Zi, C, z.S, Co);
}

// Pair work is split by row tile: tile jt holds the 64 newer loop closures
// j = 64 jt + lane, paired with the older ones i < 64 (jt + 1), which come in
// words of 64 (i-word w: i in [64 w, 64 w + 64)), each split into S runs of
// 64 / S (S = 4, 8 or 16, chosen per launch from the stage's size).  One
// wavefront takes one run of older loop closures, uniform across the
// wavefront, against its 64 lanes; a block of four wavefronts takes four
// consecutive runs of one tile.  The tile's j-side records (the loop closure
// and its two endpoint poses, 676 B per lane) are staged once per block in
// LDS and read at their points of use: held in registers for the whole walk
// they took 157 VGPRs and pushed the pair chain into scratch spills (1.26x
// slower at 5e7 pairs on MI355X).
constexpr int kTileWaves = 4;

// 169 floats per lane: an odd stride, so 32-bit LDS reads of the 64 lanes
// fall in distinct banks (a stride of 160 put all 64 lanes in one bank).
struct JSide {
  LcRec lj;
  PoseRec rc;  // pose lj.a
  PoseRec rd;  // pose lj.b
  float pad;
};
static_assert(sizeof(JSide) == 169 * 4, "j-side record stride");

// blocks of the tiles before jt: tile t has (t + 1) S runs, (t + 1) S / 4 blocks
__host__ __device__ __forceinline__ long long tile_blocks_before(long long jt, int S) {
  return (long long)(S / kTileWaves) * jt * (jt + 1) / 2;
}

// Workgroup b of a launch is dispatched to XCD b % 8.  The logical block
// index gives each XCD one contiguous eighth of the blocks, so consecutive
// blocks of a row tile (same j-side loop closures, neighbouring older ones)
// share one XCD's L2 for their anchor-table gathers (10 % fewer bytes fetched
// beyond L2 at 5e7 pairs on MI355X).
__device__ __forceinline__ long long xcd_block() {
  const long long n = gridDim.x, q = n >> 3, r = n & 7;
  const int x = blockIdx.x & 7;
  return x * q + min((long long)x, r) + (blockIdx.x >> 3);
}

// Two wavefronts per SIMD (256 VGPRs each).
__global__ void __launch_bounds__(64 * kTileWaves) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_pairs32(int m, int n_tiles, int S, const LcRec* __restrict__ lcs, const PoseRec* __restrict__ prec,
          Anchors an, float thr2, float tau, float kdyn, int W, unsigned long long* __restrict__ adj,
          double* __restrict__ dist, int2* __restrict__ queue, unsigned int* __restrict__ qcount,
          unsigned long long* __restrict__ n_consistent) {
  __shared__ JSide js[64];
  const long long blk = xcd_block();
  int lo = 0, hi = n_tiles;  // tile_blocks_before(lo) <= blk < tile_blocks_before(hi)
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (tile_blocks_before(mid, S) <= blk) lo = mid;
    else hi = mid;
  }
  const int jt = lo;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const int run = (int)(blk - tile_blocks_before(jt, S)) * kTileWaves + wave;  // run of tile jt
  const int j = jt * 64 + lane;
  const bool jvalid = j < m;
  const int jj = jvalid ? j : m - 1;
  if (wave == 0) js[lane].lj = lcs[jj];
  else if (wave == 1) js[lane].rc = prec[lcs[jj].a];
  else if (wave == 2) js[lane].rd = prec[lcs[jj].b];
  __syncthreads();
  const int iw = run / S, len = 64 / S;
  if (iw > jt) return;  // not reached: a tile's runs fill its blocks
  const int i0 = iw * 64 + (run % S) * len, i1 = min(i0 + len, m);
  unsigned long long rowbits = 0ull;
  unsigned int n_cons = 0;
  for (int i = i0; i < i1; ++i) {
    // opaque per iteration: the LDS reads below stay at their uses
    int off = lane * (int)sizeof(JSide);
    asm volatile("" : "+v"(off));
    const JSide& J = *reinterpret_cast<const JSide*>(reinterpret_cast<const char*>(js) + off);
    const bool active = jvalid && i < j;
    const LcRec& li = lcs[i];
    const PoseRec& ra = prec[li.a];
    const PoseRec& rb = prec[li.b];
    bool cons = false;
    if (active) {
      P32 TA, TX, TY, TW, TB, TL, TBi;
      S6 CA, CX, CW, CB, CL;
      between32(an, li.a, ra, J.lj.a, J.rc, TA, CA);
      compose_meas(TA, CA, J.lj, TX, CX);
      p32_inv(TX, TY);
      compose_meas(TY, CX, li, TW, CW);
      between32(an, li.b, rb, J.lj.b, J.rd, TB, CB);
      p32_mul(TW, TB, TL);
      p32_inv(TB, TBi);
      s6_congruence_add(TBi, CW, CB, CL);
      float xi[6];
      const bool log_ok = se3_log32(TL, xi);
      float d2, mp;
      const bool pd = mahal2_32(CL, xi, &d2, &mp);
      cons = d2 < thr2;
      // The reference subtracts chain covariances bounded by |Sigma_c| in PSD
      // order, so its FP64 error is ~eps64 * tr(chain), measured against the
      // smallest pivot of Sigma_L (lever arms scale error and covariance alike).
      const float chain = ra.trc + J.rc.trc + rb.trc + J.rd.trc;
      const float band = tau + kdyn * chain / fmaxf(mp, 1e-30f);
      const bool near = fabsf(d2 - thr2) <= band * thr2;
      if (!log_ok || !pd || near || !(d2 == d2) || d2 > 3.0e38f) {
        const unsigned int q = atomicAdd(qcount, 1u);
        if (q < (unsigned)kQueueCap) queue[q] = make_int2(i, j);
      }
      if (dist) {
        const double dv = sqrt((double)d2);
        dist[(size_t)j * m + i] = dv;
        dist[(size_t)i * m + j] = dv;
      }
    }
    if (cons) rowbits |= 1ull << (i & 63);
    const unsigned long long colbits = __ballot(cons);
    if (lane == 0 && colbits) {
      atomicOr(&adj[(size_t)i * W + jt], colbits);
      n_cons += (unsigned)__popcll(colbits);
    }
  }
  if (jvalid && rowbits) atomicOr(&adj[(size_t)j * W + iw], rowbits);
  if (lane == 0 && n_cons) atomicAdd(n_consistent, (unsigned long long)n_cons);
}

// consistent pair count: popcount of the full symmetric matrix / 2
__global__ void k_popcount(long long nwords, const unsigned long long* __restrict__ adj,
                           unsigned long long* __restrict__ out) {
  unsigned long long s = 0;
  for (long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x; w < nwords;
       w += (long long)gridDim.x * blockDim.x)
    s += __popcll(adj[w]);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0 && s) atomicAdd(out, s);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

// layout of fp32_buf: prec | lcrec | pl64 | pblock64 | bend64 | ta | anc | terms
struct Fp32Layout {
  size_t prec, lc, pl64, pb64, bend, ta, anc, terms, total;
};

static Fp32Layout fp32_layout(int N, int m, int nb) {
  Fp32Layout L;
  L.prec = 0;
  L.lc = L.prec + sizeof(PoseRec) * (size_t)N;
  L.pl64 = L.lc + sizeof(LcRec) * (size_t)(m > 0 ? m : 1);
  L.pb64 = L.pl64 + sizeof(double) * 12 * (size_t)N;
  L.bend = L.pb64 + sizeof(double) * 12 * (size_t)nb;
  L.ta = L.bend + sizeof(double) * 36 * (size_t)nb;
  L.anc = (L.ta + sizeof(double) * 12 * (size_t)nb + 255) & ~(size_t)255;  // AncRec is 64-byte aligned
  L.terms = L.anc + sizeof(AncRec) * (size_t)nb * nb;
  L.total = L.terms + sizeof(double) * 36 * (size_t)nb * nb;
  return L;
}

int pcm_fp32_prepare(rpgo_ctx* ctx, bool* use_fp64) {
  PcmState& s = ctx->pcm;
  hipStream_t st = ctx->stream;
  *use_fp64 = false;
  if (s.dim != 3 || s.n_chains > 1 || s.landmark) {  // these run the FP64 replay
    *use_fp64 = true;
    return RPGO_OK;
  }
  const int N = s.n_poses, nb = (N + kB - 1) / kB, m = s.m;
  // NaN rotation covariance anywhere: FP64 replay for every pair
  int* any = s.counters.as<int>() + 32;
  RPGO_HIP_CHECK(hipMemsetAsync(any, 0, sizeof(int), st));
  if (N > 1)
    k_any_no_rotation<<<(N - 1 + 255) / 256, 256, 0, st>>>(N - 1, s.delta.as<double>(), R64<3>::PWC,
                                                          48, any);
  if (s.n_lc > 0)
    k_any_no_rotation<<<(s.n_lc + 255) / 256, 256, 0, st>>>(s.n_lc, s.lcm.as<double>(), R64<3>::PWC,
                                                            48, any);
  RPGO_HIP_CHECK(hipGetLastError());
  int h_any = 0;
  RPGO_HIP_CHECK(hipMemcpyAsync(&h_any, any, sizeof(int), hipMemcpyDeviceToHost, st));
  RPGO_HIP_CHECK(hipStreamSynchronize(st));
  if (h_any) {
    *use_fp64 = true;
    return RPGO_OK;
  }
  const Fp32Layout L = fp32_layout(N, m, nb);
  int rc;
  if ((rc = s.fp32_buf.reserve(L.total))) return rc;
  char* base = s.fp32_buf.as<char>();
  PoseRec* prec = reinterpret_cast<PoseRec*>(base + L.prec);
  LcRec* lcr = reinterpret_cast<LcRec*>(base + L.lc);
  double* pl64 = reinterpret_cast<double*>(base + L.pl64);
  double* pb64 = reinterpret_cast<double*>(base + L.pb64);
  double* bend = reinterpret_cast<double*>(base + L.bend);
  double* ta = reinterpret_cast<double*>(base + L.ta);
  AncRec* anc = reinterpret_cast<AncRec*>(base + L.anc);
  double* terms = reinterpret_cast<double*>(base + L.terms);
  RPGO_HIP_CHECK(hipMemsetAsync(pb64, 0, sizeof(double) * 12 * (size_t)nb, st));
  RPGO_HIP_CHECK(hipMemsetAsync(bend, 0, sizeof(double) * 36 * (size_t)nb, st));
  k_block_local<<<(nb + 63) / 64, 64, 0, st>>>(N, nb, s.delta.as<double>(), s.hinv.as<double>(),
                                               prec, pl64, pb64, bend);
  RPGO_HIP_CHECK(hipGetLastError());
  k_anchor_poses<<<1, 1, 0, st>>>(nb, s.pose0.as<double>(), pb64, ta);
  RPGO_HIP_CHECK(hipGetLastError());
  const long long nt = (long long)nb * nb;
  k_anchor_terms<<<(unsigned)((nt + 255) / 256), 256, 0, st>>>(nb, ta, bend, terms, anc);
  k_anchor_suffix<<<(nb + 63) / 64, 64, 0, st>>>(nb, terms, anc);
  RPGO_HIP_CHECK(hipGetLastError());
  k_pose_trace<<<(N + 255) / 256, 256, 0, st>>>(N, anc, prec);
  if (m > 0)
    k_lc_records<<<(m + 255) / 256, 256, 0, st>>>(m, s.c_from.as<int>(), s.c_to.as<int>(),
                                                  s.c_lcm.as<double>(), lcr);
  RPGO_HIP_CHECK(hipGetLastError());
  s.fp32_nb = nb;
  return RPGO_OK;
}

int pcm_pairs_fp32_launch(rpgo_ctx* ctx, double lc_thr, double margin, double chain_coef,
                          bool want_dist) {
  PcmState& s = ctx->pcm;
  const int N = s.n_poses, nb = s.fp32_nb, m = s.m;
  const Fp32Layout L = fp32_layout(N, m, nb);
  char* base = s.fp32_buf.as<char>();
  Anchors an;
  an.anc = reinterpret_cast<const AncRec*>(base + L.anc);
  an.nb = nb;
  const PoseRec* prec = reinterpret_cast<const PoseRec*>(base + L.prec);
  const LcRec* lcr = reinterpret_cast<const LcRec*>(base + L.lc);
  int rc;
  if ((rc = s.queue.reserve(sizeof(int2) * (size_t)kQueueCap))) return rc;
  const int W = (m + 63) / 64;
  const long long JT = W;
  const float thr2 = (float)(lc_thr * lc_thr);
  const float tau = (float)(margin > 0.0 ? margin : 2e-3);
  const float kdyn = (float)(chain_coef > 0.0 ? chain_coef : 16.0 * 2.220446049250313e-16);
  unsigned int* qcount = s.counters.as<unsigned int>() + 16;
  unsigned long long* n_cons = s.counters.as<unsigned long long>();
  hipStream_t st = ctx->stream;
  double* dist = want_dist ? s.dist.as<double>() : nullptr;
  unsigned long long* adj = s.adj.as<unsigned long long>();
  int2* q = s.queue.as<int2>();
  // runs per i-word: S grows from 4 until the stage holds about 16 rounds of
  // the chip's resident wavefronts (8 per CU), so the partly filled last
  // round costs little, and stops at 16 so that a wavefront keeps 4 or more
  // older loop closures per staged tile.  On MI355X: 10,000 loop closures
  // take 4.60 / 4.42 / 4.50 / 4.77 ms at S = 1 / 4 / 8 / 16 (S = 1: whole
  // words, one tile per block of four); 1,000 take 105 / 93 / 93 / 115 us at
  // S = 8 / 16 / 32 / 64.
  const long long words = JT * (JT + 1) / 2;
  int S = 4;
  while (S < 16 && words * S < 128ll * ctx->n_cu) S *= 2;
  const long long blocks = tile_blocks_before(JT, S);
  if (blocks > 0)
    k_pairs32<<<(unsigned)blocks, 64 * kTileWaves, 0, st>>>(m, (int)JT, S, lcr, prec, an, thr2, tau, kdyn,
                                                            W, adj, dist, q, qcount, n_cons);
  RPGO_HIP_CHECK(hipGetLastError());
  return RPGO_OK;
}

// ---------------------------------------------------------------------------
// batched FP32 path: the single-graph tables and pair check per graph of a
// packed batch (graph g's rows at the offsets of BatchFp32Args)
// ---------------------------------------------------------------------------

size_t pcm_fp32_pose_rec_bytes() { return sizeof(PoseRec); }
size_t pcm_fp32_lc_rec_bytes() { return sizeof(LcRec); }
size_t pcm_fp32_anchor_entry_bytes() { return sizeof(AncRec); }

// graph of item t: the last g with off[g] <= t
__device__ __forceinline__ int batch_graph(const long long* __restrict__ off, int n, long long t) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(64) k_block_local_b(BatchFp32Args a) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_blocks) return;
  const int g = batch_graph(a.blk_off, a.G, t);
  const long long p0 = a.pose_off[g], o0 = p0 - g;  // first pose, first odometry row
  block_local_body((int)(t - a.blk_off[g]), (int)(a.pose_off[g + 1] - p0), a.delta + o0 * R64<3>::PWC,
                   a.hinv + o0 * 36, static_cast<PoseRec*>(a.prec) + p0, a.pl64 + p0 * 12,
                   a.pb64 + a.blk_off[g] * 12, a.bend + a.blk_off[g] * 36);
}

__global__ void k_anchor_poses_b(BatchFp32Args a) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.G) return;
  anchor_poses_body((int)(a.blk_off[g + 1] - a.blk_off[g]), a.pose0 + (size_t)g * 12, a.pb64 + a.blk_off[g] * 12,
                    a.ta + a.blk_off[g] * 12);
}

__global__ void __launch_bounds__(256) k_anchor_terms_b(BatchFp32Args a) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_sq) return;
  const int g = batch_graph(a.sq_off, a.G, t);
  anchor_terms_body(t - a.sq_off[g], (int)(a.blk_off[g + 1] - a.blk_off[g]), a.ta + a.blk_off[g] * 12,
                    a.bend + a.blk_off[g] * 36, a.terms + a.sq_off[g] * 36,
                    static_cast<AncRec*>(a.sanc) + a.sq_off[g]);
}

__global__ void k_anchor_suffix_b(BatchFp32Args a) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_blocks) return;
  const int g = batch_graph(a.blk_off, a.G, t);
  anchor_suffix_body((int)(t - a.blk_off[g]), (int)(a.blk_off[g + 1] - a.blk_off[g]), a.terms + a.sq_off[g] * 36,
                     static_cast<AncRec*>(a.sanc) + a.sq_off[g]);
}

__global__ void k_pose_trace_b(BatchFp32Args a) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_poses) return;
  const int g = batch_graph(a.pose_off, a.G, t);
  pose_trace_body((int)(t - a.pose_off[g]), static_cast<const AncRec*>(a.sanc) + a.sq_off[g],
                  static_cast<PoseRec*>(a.prec) + a.pose_off[g]);
}

__global__ void k_lc_records_b(BatchFp32Args a) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_lc) return;
  const int g = batch_graph(a.lc_off, a.G, t);
  const long long o = a.lc_off[g];
  if (t - o >= a.m[g]) return;  // beyond the accepted loop closures
  lc_record_body((int)(t - o), a.c_from + o, a.c_to + o, a.c_lcm + o * R64<3>::PWC,
                 static_cast<LcRec*>(a.lcr) + o);
}

// The pair check of k_pairs32 over every graph: workgroup b belongs to graph
// g = the last with pblk_off[g] <= b and takes four runs of one row tile of
// that graph (tiles numbered from the graph's loop-closure count; tiles at or
// beyond its accepted count exit).  Bits go to the lower triangle only (row
// j, bit i < j; the batch's symmetrize fills the upper one), and queued pairs
// carry batch-wide loop-closure indices.
// The records and offsets come as separate __restrict__ arguments: the older
// loop closure's records are uniform over the wavefront, and the compiler
// loads them into SGPRs only from memory it knows the kernel does not write
// (read through the argument struct they went to VGPRs and spilled).
__global__ void __launch_bounds__(64 * kTileWaves) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_pairs32_b(BatchFp32Args a, const LcRec* __restrict__ lcs_all, const PoseRec* __restrict__ prec_all,
            const long long* __restrict__ pblk_off, const long long* __restrict__ lc_off,
            const long long* __restrict__ pose_off, const long long* __restrict__ blk_off,
            const long long* __restrict__ sq_off, const long long* __restrict__ adj_off,
            const int* __restrict__ m_arr, float thr2, float tau, float kdyn) {
  __shared__ JSide js[64];
  if (*a.no_rotation) return;
  const long long blk = xcd_block();
  const int g = batch_graph(pblk_off, a.G, blk);
  const int S = a.S;
  const long long o = lc_off[g];
  const int n_lc = (int)(lc_off[g + 1] - o);
  const int JTg = (n_lc + 63) / 64;
  const long long local = blk - pblk_off[g];
  int lo = 0, hi = JTg;  // tile_blocks_before(lo) <= local < tile_blocks_before(hi)
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (tile_blocks_before(mid, S) <= local) lo = mid;
    else hi = mid;
  }
  const int jt = lo;
  const int m = m_arr[g];
  if (jt * 64 >= m) return;  // tile beyond the accepted loop closures (whole workgroup)
  const LcRec* __restrict__ lcs = lcs_all + o;
  const PoseRec* __restrict__ prec = prec_all + pose_off[g];
  Anchors an;
  an.anc = static_cast<const AncRec*>(a.sanc) + sq_off[g];
  an.nb = (int)(blk_off[g + 1] - blk_off[g]);
  const int W = (m + 63) / 64;
  unsigned long long* __restrict__ adj = a.adj + adj_off[g];
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const int run = (int)(local - tile_blocks_before(jt, S)) * kTileWaves + wave;
  const int j = jt * 64 + lane;
  const bool jvalid = j < m;
  const int jj = jvalid ? j : m - 1;
  if (wave == 0) js[lane].lj = lcs[jj];
  else if (wave == 1) js[lane].rc = prec[lcs[jj].a];
  else if (wave == 2) js[lane].rd = prec[lcs[jj].b];
  __syncthreads();
  const int iw = run / S, len = 64 / S;
  if (iw > jt) return;  // not reached: a tile's runs fill its blocks
  const int i0 = iw * 64 + (run % S) * len, i1 = min(i0 + len, m);
  unsigned long long rowbits = 0ull;
  unsigned int n_cons = 0;
  for (int i = i0; i < i1; ++i) {
    int off = lane * (int)sizeof(JSide);
    asm volatile("" : "+v"(off));
    const JSide& J = *reinterpret_cast<const JSide*>(reinterpret_cast<const char*>(js) + off);
    const bool active = jvalid && i < j;
    const LcRec& li = lcs[i];
    const PoseRec& ra = prec[li.a];
    const PoseRec& rb = prec[li.b];
    bool cons = false;
    if (active) {
      P32 TA, TX, TY, TW, TB, TL, TBi;
      S6 CA, CX, CW, CB, CL;
      between32(an, li.a, ra, J.lj.a, J.rc, TA, CA);
      compose_meas(TA, CA, J.lj, TX, CX);
      p32_inv(TX, TY);
      compose_meas(TY, CX, li, TW, CW);
      between32(an, li.b, rb, J.lj.b, J.rd, TB, CB);
      p32_mul(TW, TB, TL);
      p32_inv(TB, TBi);
      s6_congruence_add(TBi, CW, CB, CL);
      float xi[6];
      const bool log_ok = se3_log32(TL, xi);
      float d2, mp;
      const bool pd = mahal2_32(CL, xi, &d2, &mp);
      cons = d2 < thr2;
      const float chain = ra.trc + J.rc.trc + rb.trc + J.rd.trc;
      const float band = tau + kdyn * chain / fmaxf(mp, 1e-30f);
      const bool near = fabsf(d2 - thr2) <= band * thr2;
      if (!log_ok || !pd || near || !(d2 == d2) || d2 > 3.0e38f) {
        const unsigned int q = atomicAdd(a.qcount, 1u);
        if ((long long)q < a.qcap) a.queue[q] = make_int2((int)(o + i), (int)(o + j));
      }
    }
    if (cons) rowbits |= 1ull << (i & 63);
    const unsigned long long colbits = __ballot(cons);
    if (lane == 0) n_cons += (unsigned)__popcll(colbits);
  }
  if (jvalid && rowbits) atomicOr(&adj[(size_t)j * W + iw], rowbits);
  if (lane == 0 && n_cons) atomicAdd(a.n_consistent, (unsigned long long)n_cons);
}

int pcm_fp32_batch_launch(hipStream_t st, const BatchFp32Args& a, double lc_thr, double margin,
                          double chain_coef) {
  // a covariance without rotation information anywhere: the FP64 pairs instead
  if (a.n_odom > 0)
    k_any_no_rotation<<<(unsigned)((a.n_odom + 255) / 256), 256, 0, st>>>((int)a.n_odom, a.delta, R64<3>::PWC, 48,
                                                                          a.no_rotation);
  if (a.n_lc > 0)
    k_any_no_rotation<<<(unsigned)((a.n_lc + 255) / 256), 256, 0, st>>>((int)a.n_lc, a.lcm, R64<3>::PWC, 48,
                                                                        a.no_rotation);
  if (a.n_blocks > 0) {
    k_block_local_b<<<(unsigned)((a.n_blocks + 63) / 64), 64, 0, st>>>(a);
    k_anchor_poses_b<<<(unsigned)((a.G + 63) / 64), 64, 0, st>>>(a);
    k_anchor_terms_b<<<(unsigned)((a.n_sq + 255) / 256), 256, 0, st>>>(a);
    k_anchor_suffix_b<<<(unsigned)((a.n_blocks + 63) / 64), 64, 0, st>>>(a);
    k_pose_trace_b<<<(unsigned)((a.n_poses + 255) / 256), 256, 0, st>>>(a);
  }
  if (a.n_lc > 0) k_lc_records_b<<<(unsigned)((a.n_lc + 255) / 256), 256, 0, st>>>(a);
  RPGO_HIP_CHECK(hipGetLastError());
  const float thr2 = (float)(lc_thr * lc_thr);
  const float tau = (float)(margin > 0.0 ? margin : 2e-3);
  const float kdyn = (float)(chain_coef > 0.0 ? chain_coef : 16.0 * 2.220446049250313e-16);
  if (a.n_pair_blocks > 0)
    k_pairs32_b<<<(unsigned)a.n_pair_blocks, 64 * kTileWaves, 0, st>>>(
        a, static_cast<const LcRec*>(a.lcr), static_cast<const PoseRec*>(a.prec), a.pblk_off, a.lc_off, a.pose_off,
        a.blk_off, a.sq_off, a.adj_off, a.m, thr2, tau, kdyn);
  RPGO_HIP_CHECK(hipGetLastError());
  return RPGO_OK;
}

int pcm_count_consistent(rpgo_ctx* ctx) {
  PcmState& s = ctx->pcm;
  const int m = s.m, W = (m + 63) / 64;
  const long long nwords = (long long)m * W;
  if (nwords == 0) return RPGO_OK;
  long long blocks = (nwords + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  k_popcount<<<(unsigned)blocks, 256, 0, ctx->stream>>>(nwords, s.adj.as<unsigned long long>(),
                                                        s.counters.as<unsigned long long>() + 1);
  RPGO_HIP_CHECK(hipGetLastError());
  return RPGO_OK;
}

}  // namespace rpgo
