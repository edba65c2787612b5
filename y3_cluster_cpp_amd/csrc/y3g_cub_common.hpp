// y3g_cub_common.hpp -- pieces of the Genz-Malik cubature shared by the
// batched kernel (y3g_kernels.hip) and the device-wide single-integral path
// (y3g_single.hip): workgroup reductions, LDS staging of a point's tables,
// region evaluation and the error-bin split selection.
#pragma once

#include <hip/hip_runtime.h>

#include "y3g_internal.hpp"

namespace y3g {

// q / n for the task loops (q < 2^20, 0 < n <= 64) by a multiply-high with
// ceil(2^32 / n): exact there, and a few instructions in place of a
// run-time integer division.
__host__ __device__ inline unsigned long long div_magic(int n)
{
  return 0xFFFFFFFFull / (unsigned long long)n + 1ull;
}
__device__ __forceinline__ int div_by(int q, unsigned long long magic)
{
  return (int)(((unsigned long long)(unsigned)q * magic) >> 32);
}

// Workgroup barrier ordering LDS only: each wave waits for its own LDS
// operations (lgkmcnt) but not for its vector-memory ones, so that the next
// chunk's parents fetched straight into LDS (fetch_parents) stay in flight
// across a chunk's evaluation.  For phases that exchange data through LDS
// alone.
__device__ __forceinline__ void lds_barrier()
{
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// threadIdx.x behind an empty asm: the per-thread values derived from it
// (LDS offsets, lane masks, pointer choices) are recomputed where they are
// used, a few VALU operations, instead of being hoisted out of the loops,
// which at the FP32 kernel's 80 VGPRs spills them to scratch and reloads
// them from memory at every use.
__device__ __forceinline__ int tid_here()
{
  int t = (int)threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// ------------------------------------------------------------ reductions
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int wave_sum_i(int v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Reductions within each row of 16 lanes through DPP lane permutations
// (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror):
// after the four steps every lane of a row holds the row's sum / maximum.
// No LDS round trip; the whole wave must be active.
template <int CTRL> __device__ __forceinline__ double dpp_row(double v)
{
  int const lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  int const hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double row16_sum(double v)
{
  v += dpp_row<0xB1>(v);
  v += dpp_row<0x4E>(v);
  v += dpp_row<0x141>(v);
  v += dpp_row<0x140>(v);
  return v;
}
__device__ __forceinline__ double row16_max(double v)
{
  v = fmax(v, dpp_row<0xB1>(v));
  v = fmax(v, dpp_row<0x4E>(v));
  v = fmax(v, dpp_row<0x141>(v));
  v = fmax(v, dpp_row<0x140>(v));
  return v;
}

template <int BLOCK>
__device__ double block_sum(double v, double* red)
{
  constexpr int NW = BLOCK / 64;
  int const lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NW; ++i) s += red[i];
  return s;
}

template <int BLOCK>
__device__ void block_sum2(double& a, double& b, double* red)
{
  constexpr int NW = BLOCK / 64;
  int const lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  a = wave_sum(a);
  b = wave_sum(b);
  __syncthreads();
  if (lane == 0) red[w] = a, red[NW + w] = b;
  __syncthreads();
  double sa = 0.0, sb = 0.0;
#pragma unroll
  for (int i = 0; i < NW; ++i) sa += red[i], sb += red[NW + i];
  a = sa;
  b = sb;
}

template <int BLOCK>
__device__ int block_sum_i(int v, int* red)
{
  constexpr int NW = BLOCK / 64;
  int const lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  v = wave_sum_i(v);
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  int s = 0;
#pragma unroll
  for (int i = 0; i < NW; ++i) s += red[i];
  return s;
}

// Table pointers of precision T.
template <typename T> struct TabSel;
template <> struct TabSel<float> {
  static __device__ const float* hmf(const KernArgs& A) { return A.hmf32; }
  static __device__ const float* hmfc(const KernArgs& A) { return A.hmfc32; }
  static __device__ const float* da(const KernArgs& A) { return A.da32; }
  static __device__ const float* sci(const KernArgs& A) { return A.sci32; }
};
template <> struct TabSel<double> {
  static __device__ const double* hmf(const KernArgs& A) { return A.hmf64; }
  static __device__ const double* hmfc(const KernArgs& A) { return A.hmfc64; }
  static __device__ const double* da(const KernArgs& A) { return A.da64; }
  static __device__ const double* sci(const KernArgs& A) { return A.sci64; }
};

// Stage the tables of point pt into LDS (the HMF table only when s_hmf is
// given).
template <typename T>
__device__ void stage_tables(const KernArgs& A, int pt, T* s_hmf, T* s_da,
                             T* s_sci)
{
  const T* g_hmf = TabSel<T>::hmf(A) + (size_t)pt * A.hmf_stride;
  const T* g_da = TabSel<T>::da(A) + (size_t)pt * A.da_stride;
  const T* g_sci = TabSel<T>::sci(A) + (size_t)pt * A.sci_stride;
  if (s_hmf)
    for (int k = threadIdx.x; k < A.hmf_stride; k += blockDim.x) s_hmf[k] = g_hmf[k];
  for (int k = threadIdx.x; k < A.da_stride; k += blockDim.x) s_da[k] = g_da[k];
  for (int k = threadIdx.x; k < A.sci_stride; k += blockDim.x) s_sci[k] = g_sci[k];
}

// ------------------------------------------------------------- cubature
//
// Region algorithm (identical to the oracle's y3o_cubature restatement):
// initial k^d uniform partition; per region R7, R5, err = |R7 - R5| and the
// split axis (largest fourth difference; widest relative edge if all zero);
// each step splits the largest-error regions holding half of E (and at least
// max(64, N/8) of them, at most kmax; select_bins) at the midpoint of their
// axis, child A in the parent slot, child B appended; if
// |R7_A + R7_B - R7_parent| exceeds the children's combined error, their
// errors are scaled up to that difference; stop when
// E <= max(epsabs, epsrel |I|) (status 0) or when the region or evaluation
// budget would be exceeded (status 1).

constexpr int kChunk = 32;                      // regions per evaluation chunk

// Pieces of the initial partition along axis a: k = init_split (or 4 for
// d <= 3, 2 above) on every axis, and init_lnm along the lnM axis of the
// synthetic kinds.  The MOR and the lambda_obs kernel confine their
// integrands to bands 0.1 - 0.4 wide in lnM, and a region still spanning
// half the mass range can hold such a band between its nodes, report a tiny
// |R7 - R5| and never be split.  Which regions miss a band depends on the
// partition: on the first 64 stream points (12,288 integrals, smoothed HMF)
// 2, 4, 8 and 16 pieces left 6, 2, 3 and 10 reported errors below the true
// ones; 4 is the default, and with err_scale 1.5 none remains.
__host__ __device__ inline int lnm_axis_of(int kind)
{
  return kind == Y3G_COUNT3D ? 2 : kind == Y3G_COUNT5D ? 4 : kind == Y3G_DSIGMA6D ? 3 : -1;
}
__host__ __device__ inline int init_pieces(int kind, int d, int init_split, int init_lnm, int a)
{
  int const k = init_split > 0 ? init_split : (d <= 3 ? 4 : 2);
  return (init_lnm > 0 && a == lnm_axis_of(kind)) ? init_lnm : k;
}

// Region error from the degree-7, -5 and -3 rules: E1 = |R7 - R5| is the
// error of R5; when the rule sequence converges (r = E1 / |R5 - R3| small)
// the error of R7 is a fraction of E1, taken as E1 * clamp(c r, smin, 1).
__host__ __device__ inline double rule_error(double r7, double r5, double r3,
                                             double c, double smin)
{
  double const e1 = fabs(r7 - r5), e2 = fabs(r5 - r3);
  if (!(c > 0.0) || !(e2 > 0.0)) return e1;
  double f = c * e1 / e2;
  f = f < smin ? smin : (f > 1.0 ? 1.0 : f);
  return e1 * f;
}
constexpr double kAlpha = 0.5;                  // error mass split per step
constexpr int kMinSplit = 64;

// Error bin: binary exponent of E minus that of err (0 = largest), 63 for
// err == 0.
__host__ __device__ inline int err_bin(double err, double E, int eE)
{
  if (!(err > 0.0)) return 63;
  int e;
  (void)frexp(err, &e);
  int const b = eE - e;
  return b < 0 ? 0 : (b > 62 ? 62 : b);
}

// err / E in 2^-40 fixed point, with qs = 2^40 / E (err_qscale) computed
// once per step: one multiply per region instead of a division.
__host__ __device__ inline double err_qscale(double E) { return 1099511627776.0 / E; }
__host__ __device__ inline unsigned long long err_q(double err, double qs)
{
  double const q = floor(err * qs);
  return q > 0.0 ? (unsigned long long)q : 0ull;
}

// Split bin B: the first bin where the cumulative mass reaches kAlpha and
// the first where the cumulative count reaches max(kMinSplit, N/8) (capped
// by N and room); all of bins < B up to room (lim0), then bin B (lim1).
__host__ __device__ inline void select_bins(const int* hc,
                                            const unsigned long long* hq,
                                            int N, int room, int& B, int& lim0,
                                            int& lim1)
{
  unsigned long long const qa =
    (unsigned long long)(kAlpha * 1099511627776.0);
  int km = N / 8 > kMinSplit ? N / 8 : kMinSplit;
  if (km > room) km = room;
  if (km > N) km = N;
  int Ba = 63, Bk = 63;
  unsigned long long cq = 0;
  int cc = 0;
  for (int b = 0; b < 64; ++b) {
    cq += hq[b];
    cc += hc[b];
    if (Ba == 63 && cq >= qa) Ba = b;
    if (Bk == 63 && cc >= km) Bk = b;
  }
  B = Ba > Bk ? Ba : Bk;
  int below = 0;
  for (int b = 0; b < B; ++b) below += hc[b];
  lim0 = below < room ? below : room;
  int const inb = B < 63 ? hc[B] : 0;
  lim1 = room - lim0 < inb ? room - lim0 : inb;
}
constexpr int kMaxNodes = gm_num_nodes(kMaxDim);  // 149
constexpr int kKmax = 1024;
// distinct (zt, lnM) offset pairs among a region's Genz-Malik nodes: the
// centre, +-l2, +-l3, +-l4 on either axis alone, (+-l4, +-l4), (+-l5, +-l5)
constexpr int kMaxPairs = 21;
// distinct offsets on one axis: 0, +-l2, +-l3 (= +-l4), +-l5
constexpr int kMaxAxisOffsets = 7;
// distinct (lnM, u, theta) offset triples among the nodes of a 6-D region:
// the centre, +-l2 and +-l3 on one of the three axes, (+-l4, +-l4) on two of
// them, (+-l5, +-l5, +-l5)
constexpr int kMaxProf = 1 + 6 + 6 + 12 + 8;
// nodes read by the fourth differences: the centre, +-l2 and +-l3 per axis
constexpr int kFourNodes = 1 + 4 * kMaxDim;
// distinct (lo, t, zt, lnM) offset quadruples among the nodes of a 5-D or 6-D
// region: the centre, +-l2 and +-l3 on one of the four axes, (+-l4, +-l4) on
// two of them, (+-l5, ..., +-l5)
constexpr int kMaxQuads = 1 + 8 + 8 + 24 + 16;

// One chunk's parents as fetched from the region slab, 512 dwords: lo and hi
// [parent][axis], est, err, raw [parent], and the dword holding each
// parent's split axis byte.
struct ParentStage {
  double lo[kChunk / 2][kMaxDim], hi[kChunk / 2][kMaxDim];
  double est[kChunk / 2], err[kChunk / 2], raw[kChunk / 2];
  unsigned axw[kChunk / 2];
  unsigned pad[512 - 2 * (2 * (kChunk / 2) * kMaxDim + 3 * (kChunk / 2)) - kChunk / 2];
};
static_assert(sizeof(ParentStage) == 2048, "ParentStage is 512 dwords");

// DES data of a point staged in the FP64 workgroup's LDS up to this many
// doubles (the synthetic SIG_MAX tables of the bench take 887)
constexpr int kDesLds = 2048;

template <typename T, int BLOCK> struct CubShared {
  // the HMF table stays in global memory (L2) for both precisions: the FP32
  // path reads it only for the pair factors, and in LDS it would take 51 KB
  // (FP64) of the workgroup, one workgroup per CU instead of two (FP64) or
  // three (FP32)
  static constexpr bool kHmfLds = false;
  T hmf[kHmfLds ? kMaxHmf : 1];
  const T* hmf_tab;  // the table the evaluation reads (LDS or global)
  const T* hmfc_tab;  // its spline coefficients (global), null: bilinear HMF
  const double* des;  // the point's DES data (real-data kinds): desl or global
  // the point's DES data staged in LDS when it fits, and the MOR_DES tables
  // (FP64 only)
  double desl[sizeof(T) == 8 ? kDesLds : 1];
  double mort[sizeof(T) == 8 ? kMorDesWords : 1];
  T da[kMaxDa];
  T sci[kMaxSci];
  // node values: the FP32 path keeps only those the fourth differences read
  T f[kChunk * (sizeof(T) == 4 ? kFourNodes : kMaxNodes)];
  double c[kChunk][kMaxDim], h[kChunk][kMaxDim];
  double est[kChunk], err[kChunk];
  int ax[kChunk];
  int idx[kChunk];
  double pest[kChunk / 2], perr[kChunk / 2], praw[kChunk / 2];
  int sel[kKmax];
  double red[2 * (BLOCK / 64)];
  int ired[BLOCK / 64];
  int ired1[BLOCK / 64];
  int hc[64];
  unsigned long long hq[64];
  int wc[BLOCK / 64][64];  // per-wave error-bin counts
  int selB, lim0, lim1;
  double w7[6], w5[6], w3[6];
  y3g_integral ig;
  PointDev P;
  int item, point;
  unsigned pgen;           // streamed: the table generation of `point`
  long long sitem;         // streamed: the claimed ring entry (-1: none)
  int job, sslot;          // streamed: the entry's job and table slot
  unsigned sgen;           //   and the generation it needs or builds
  int setup_kind;  // kind x 16 + ndim the node maps (pair_setup) hold, -1 none
  // FP32 path, per chunk: the pairs' factors and (Delta-Sigma) the profile
  // of each (lnM, u, theta) triple of the chunk's regions; the zt and lnM
  // factors they are made of (zmbuf: ZFac [kChunk][7], then MFac [kChunk][7])
  static constexpr bool kF32 = sizeof(T) == 4;
  PairC pc[kF32 ? kChunk : 1][kMaxPairs];
  float prof[kF32 ? kChunk : 1][kMaxProf];
  alignas(16) unsigned char
    zmbuf[kF32 ? (sizeof(ZFac<float>) + sizeof(MFac<float>)) * kChunk * kMaxAxisOffsets : 1];
  // FP32 path, per integral: the pairs' unit offsets, their distinct zt and
  // lnM offsets, the profile triples' offsets, and per node its pair, class,
  // triple and unit offsets
  double poff[kMaxPairs][2];
  double zoffs[kMaxAxisOffsets], moffs[kMaxAxisOffsets];  // distinct zt / lnM offsets
  unsigned char pzi[kMaxPairs], pmi[kMaxPairs];            // each pair's two
  int nzo, nmo;
  float tuo[kMaxProf], tto[kMaxProf];                      // triple: u, theta offsets
  unsigned char tmi[kMaxProf];                             //   and its lnM offset
  int nprofs;            // (5-D counts: the (zo, zt) offset pairs in tuo / tto)
  float qo0[kMaxQuads], qo1[kMaxQuads];                    // quadruple: lo, t offsets
  unsigned char qk[kMaxQuads];                             //   and its pair
  int nquads;
  unsigned char nquad[kMaxNodes];                          // quadruple of each node
  unsigned short nqp[kMaxNodes];                           // quadruple | second << 8
  NfwKnot nfwtab[kF32 ? kNfwN : 1];                        // NFW knot table, LDS copy
  unsigned long long dv_zo, dv_mo, dv_np, dv_nt, dv_nq;    // div_magic of the counts
  unsigned char npair[kMaxNodes];
  unsigned char nclass[kMaxNodes];                    // gm_class of each node
  unsigned char nprof[kMaxNodes];                     // triple of each node
  float noff[kF32 ? kMaxNodes : 1][kMaxDim];          // unit offsets, float
  int npairs;
};

// The distinct (zt, lnM) unit offsets of the current integral's nodes and
// each node's index among them (one thread).
template <typename T, int BLOCK>
__device__ void pair_setup(CubShared<T, BLOCK>& S)
{
  int it, iz, im;
  kind_axes<T>(S.ig.kind, it, iz, im);
  int const d = S.ig.ndim;
  int const nn = gm_num_nodes(d);
  int np = 0;
  for (int j = 0; j < nn; ++j) {
    double const oz = gm_offset(d, j, iz), om = gm_offset(d, j, im);
    int k = 0;
    while (k < np && !(S.poff[k][0] == oz && S.poff[k][1] == om)) ++k;
    if (k == np && np < kMaxPairs) {
      S.poff[np][0] = oz;
      S.poff[np][1] = om;
      ++np;
    }
    S.npair[j] = (unsigned char)(k < kMaxPairs ? k : kMaxPairs - 1);
    S.nclass[j] = (unsigned char)gm_class(d, j);
    if constexpr (sizeof(T) == 4)
      for (int a = 0; a < kMaxDim; ++a) S.noff[j][a] = a < d ? float(gm_offset(d, j, a)) : 0.0f;
  }
  S.npairs = np;
  // the pairs' distinct zt and lnM offsets
  int nz = 0, nm = 0;
  for (int k = 0; k < np; ++k) {
    int a = 0;
    while (a < nz && S.zoffs[a] != S.poff[k][0]) ++a;
    if (a == nz && nz < kMaxAxisOffsets) S.zoffs[nz++] = S.poff[k][0];
    S.pzi[k] = (unsigned char)(a < kMaxAxisOffsets ? a : kMaxAxisOffsets - 1);
    int b = 0;
    while (b < nm && S.moffs[b] != S.poff[k][1]) ++b;
    if (b == nm && nm < kMaxAxisOffsets) S.moffs[nm++] = S.poff[k][1];
    S.pmi[k] = (unsigned char)(b < kMaxAxisOffsets ? b : kMaxAxisOffsets - 1);
  }
  S.nzo = nz;
  S.nmo = nm;
  // Delta-Sigma: the distinct (lnM, u, theta) offset triples
  int nt = 0;
  if (S.ig.kind == Y3G_DSIGMA6D) {
    for (int j = 0; j < nn; ++j) {
      double const om = gm_offset(d, j, im), ou = gm_offset(d, j, 4), ot = gm_offset(d, j, 5);
      int b = 0;
      while (b < nm && S.moffs[b] != om) ++b;
      int t = 0;
      while (t < nt && !(S.tmi[t] == b && S.tuo[t] == float(ou) && S.tto[t] == float(ot))) ++t;
      if (t == nt && nt < kMaxProf) {
        S.tmi[nt] = (unsigned char)(b < kMaxAxisOffsets ? b : kMaxAxisOffsets - 1);
        S.tuo[nt] = float(ou);
        S.tto[nt] = float(ot);
        ++nt;
      }
      S.nprof[j] = (unsigned char)(t < kMaxProf ? t : kMaxProf - 1);
    }
  }
  // 5-D counts: the distinct (zo, zt) offset pairs
  if (S.ig.kind == Y3G_COUNT5D) {
    for (int j = 0; j < nn; ++j) {
      float const ozo = float(gm_offset(d, j, 1)), ozt = float(gm_offset(d, j, iz));
      int t = 0;
      while (t < nt && !(S.tuo[t] == ozo && S.tto[t] == ozt)) ++t;
      if (t == nt && nt < kMaxProf) {
        S.tuo[nt] = ozo;
        S.tto[nt] = ozt;
        ++nt;
      }
      S.nprof[j] = (unsigned char)(t < kMaxProf ? t : kMaxProf - 1);
    }
  }
  S.nprofs = nt;
  // 5-D counts and Delta-Sigma: the distinct (lo, t, zt, lnM) quadruples,
  // each with its (zt, lnM) pair
  int nq = 0;
  if (S.ig.kind != Y3G_COUNT3D) {
    for (int j = 0; j < nn; ++j) {
      float const o0 = float(gm_offset(d, j, 0)), ot = float(gm_offset(d, j, it));
      int const k = S.npair[j];
      int q = 0;
      while (q < nq && !(S.qo0[q] == o0 && S.qo1[q] == ot && S.qk[q] == k)) ++q;
      if (q == nq && nq < kMaxQuads) {
        S.qo0[nq] = o0;
        S.qo1[nq] = ot;
        S.qk[nq] = (unsigned char)k;
        ++nq;
      }
      S.nquad[j] = (unsigned char)(q < kMaxQuads ? q : kMaxQuads - 1);
    }
  }
  S.nquads = nq;
  for (int j = 0; j < nn; ++j) S.nqp[j] = (unsigned short)(S.nquad[j] | (S.nprof[j] << 8));
  S.dv_zo = div_magic(nz > 0 ? nz : 1);
  S.dv_mo = div_magic(nm > 0 ? nm : 1);
  S.dv_np = div_magic(np > 0 ? np : 1);
  S.dv_nt = div_magic(nt > 0 ? nt : 1);
  S.dv_nq = div_magic(nq > 0 ? nq : 1);
}

// Factors of m regions' nodes (FP32 path), one per thread in three phases:
// the zt and lnM factors of each region's 7 + 7 distinct values; their
// combinations for the 21 (zt, lnM) pairs and, for Delta-Sigma, the profiles
// of the 33 (lnM, u, theta) triples; then (5-D counts, Delta-Sigma) the
// factor of the 57 (lo, t, zt, lnM) quadruples and, for the counts, the
// photo-z pdf of the 17 (zo, zt) pairs.  A node's value is then a product of
// two table entries.
template <typename T, int BLOCK, bool SM = false>
__device__ void eval_pairs(CubShared<T, BLOCK>& S, int m, Tabs<T> tb)
{
  int iz, im, it;
  kind_axes<T>(S.ig.kind, it, iz, im);
  ZFac<T>* zf = reinterpret_cast<ZFac<T>*>(S.zmbuf);
  MFac<T>* mf = reinterpret_cast<MFac<T>*>(zf + kChunk * kMaxAxisOffsets);
  int const nzo = S.nzo, nmo = S.nmo;
  // all zt tasks, then all lnM tasks: a wave takes one kind (but at the seam)
  for (int q = tid_here(); q < m * (nzo + nmo); q += BLOCK) {
    if (q < m * nzo) {
      int const r = div_by(q, S.dv_zo), k = q - r * nzo;
      zf[r * kMaxAxisOffsets + k] =
        z_factors<T>(S.ig, S.P, tb, S.c[r][iz] + S.h[r][iz] * S.zoffs[k]);
    } else {
      int const qm = q - m * nzo;
      int const r = div_by(qm, S.dv_mo), k = qm - r * nmo;
      mf[r * kMaxAxisOffsets + k] =
        m_factors<T>(S.ig, S.P, S.c[r][im] + S.h[r][im] * S.moffs[k], S.nfwtab);
    }
  }
  lds_barrier();
  int const np = S.npairs, nt = S.nprofs;
  int const ntp = S.ig.kind == Y3G_DSIGMA6D ? nt : 0;  // profiles
  NodeConstF const kc = node_const_f(S.ig, S.P, S.nfwtab);
  for (int q = tid_here(); q < m * (np + ntp); q += BLOCK) {
    if (q < m * np) {
      int const r = div_by(q, S.dv_np), k = q - r * np;
      S.pc[r][k] = pair_c(pair_combine<T, SM>(S.ig, S.P, tb, zf[r * kMaxAxisOffsets + S.pzi[k]],
                                          mf[r * kMaxAxisOffsets + S.pmi[k]]));
    } else {
      int const qt = q - m * np;
      int const r = div_by(qt, S.dv_nt), t = qt - r * nt;
      MFac<T> const& g = mf[r * kMaxAxisOffsets + S.tmi[t]];
      float const u = float(S.c[r][4]) + float(S.h[r][4]) * S.tuo[t];
      float const w = float(1.0 - S.c[r][4]) - float(S.h[r][4]) * S.tuo[t];
      float const th = float(S.c[r][5]) + float(S.h[r][5]) * S.tto[t];
      S.prof[r][t] = profile_f(kc, g.irs, g.amp, g.cds, u, w, th);
    }
  }
  // 5-D counts and Delta-Sigma: the (lo, t, zt, lnM) factor of each
  // quadruple (over the zt / lnM factors, no longer needed) and, for the
  // counts, the photo-z pdf of each (zo, zt) pair
  if (S.ig.kind == Y3G_COUNT3D) return;
  static_assert(sizeof(float) * kChunk * kMaxQuads <= sizeof(S.zmbuf),
                "quadruple factors exceed their buffer");
  lds_barrier();
  float* Q = reinterpret_cast<float*>(S.zmbuf);
  int const nq = S.nquads;
  int const nz2 = S.ig.kind == Y3G_COUNT5D ? nt : 0;
  for (int q = tid_here(); q < m * (nq + nz2); q += BLOCK) {
    if (q < m * nq) {
      int const r = div_by(q, S.dv_nq), k = q - r * nq;
      float const lo = float(S.c[r][0]) + float(S.h[r][0]) * S.qo0[k];
      float const yt = float(S.c[r][it]) + float(S.h[r][it]) * S.qo1[k];
      Q[r * kMaxQuads + k] = quad_f(kc, S.pc[r][S.qk[k]], lo, yt);
    } else {
      int const qz = q - m * nq;
      int const r = div_by(qz, S.dv_nt), t = qz - r * nz2;
      float const zo = float(S.c[r][1]) + float(S.h[r][1]) * S.tuo[t];
      float const zt = float(S.c[r][iz]) + float(S.h[r][iz]) * S.tto[t];
      S.prof[r][t] = zo_std<float>(zo, zt, kc.zo_lo, kc.zo_hi, kc.sigz0);
    }
  }
}

// The dwords of parents sel[pb, pb + np) in ParentStage order, 512 / BLOCK
// per thread (dword i = k BLOCK + threadIdx.x): plain loads whose results
// are first used when stored to the stage, so that they stay in flight across
// whatever the caller does in between.  Lanes past np fetch parent sel[pb]
// again, so every address lies in the slab (the split-axis array is padded by
// 64 bytes for the dword holding its last byte).
template <int BLOCK> struct StageWords {
  unsigned w[512 / BLOCK];
};
template <int BLOCK>
__device__ __forceinline__ StageWords<BLOCK> fetch_parents(const int* sel, int pb, int np, int cap,
                                                          const double* lo, const double* hi,
                                                          const double* est, const double* err,
                                                          const double* raw, const int8_t* axs)
{
  constexpr int kPar = kChunk / 2, kLoW = 2 * kPar * kMaxDim;
  StageWords<BLOCK> out;
#pragma unroll
  for (int k = 0; k < 512 / BLOCK; ++k) {
    int const i = k * BLOCK + tid_here();
    const unsigned* src;
    if (i < 2 * kLoW) {
      int const half = i & 1, jd = (i % kLoW) >> 1;
      int const j = jd / kMaxDim, a = jd - j * kMaxDim;
      int const p = sel[pb + (j < np ? j : 0)];
      const double* base = i < kLoW ? lo : hi;
      src = reinterpret_cast<const unsigned*>(base + (size_t)a * cap + p) + half;
    } else if (i < 2 * kLoW + 6 * kPar) {
      int const q = i - 2 * kLoW, f = q / (2 * kPar), j = (q % (2 * kPar)) >> 1;
      int const p = sel[pb + (j < np ? j : 0)];
      const double* base = f == 0 ? est : (f == 1 ? err : raw);
      src = reinterpret_cast<const unsigned*>(base + p) + (q & 1);
    } else if (i < 2 * kLoW + 7 * kPar) {
      int const j = i - 2 * kLoW - 6 * kPar;
      int const p = sel[pb + (j < np ? j : 0)];
      src = reinterpret_cast<const unsigned*>(reinterpret_cast<uintptr_t>(axs + p) & ~uintptr_t(3));
    } else {
      src = reinterpret_cast<const unsigned*>(lo);
    }
    // through the global address space: a flat load would also count on
    // lgkmcnt and hold up every LDS wait until it returns
    out.w[k] = *(const __attribute__((address_space(1))) unsigned*)src;
  }
  return out;
}
template <int BLOCK>
__device__ __forceinline__ void store_stage(ParentStage* st, const StageWords<BLOCK>& v)
{
#pragma unroll
  for (int k = 0; k < 512 / BLOCK; ++k)
    reinterpret_cast<unsigned*>(st)[k * BLOCK + threadIdx.x] = v.w[k];
}

// FP32 nodes of region r of a 5-D count or Delta-Sigma integral, by the 16
// lanes of a row: each node's value is the product of its quadruple's factor
// and its profile (Delta-Sigma) or photo-z pdf (counts); this lane's partial
// rule sums, and into S.f the values the fourth differences read.
template <typename T, int BLOCK, int K>
__device__ __forceinline__ void node_row_q(CubShared<T, BLOCK>& S, int r, int sub, double& a7,
                                           double& a5, double& a3)
{
  // the nodes class by class (centre, +-l2 and +-l3 per axis, the l4 pairs,
  // the l5 corners: consecutive node indices), so that each loop has its
  // class's rule weights as compile-time constants and skips the zero ones
  constexpr int d = KindAxes<K>::nd;
  constexpr int c4 = 2 * d * (d - 1), c5 = 1 << d;
  constexpr GmWeights W = gm_weights_of(d);
  const float* Qr = reinterpret_cast<const float*>(S.zmbuf) + r * kMaxQuads;
  const float* F2r = S.prof[r];
  T* fr = S.f + r * kFourNodes;
  auto value = [&](int j) {
    unsigned const qp = S.nqp[j];
    return Qr[qp & 0xFFu] * F2r[qp >> 8];
  };
  if (sub == 0) {
    float const fv = value(0);
    fr[0] = fv;
    double const fd = fv;
    a7 += W.w7[1] * fd;
    a5 += W.w5[1] * fd;
    a3 += W.w3[1] * fd;
  }
  if (sub < 2 * d) {
    float const fv2 = value(1 + sub), fv3 = value(1 + 2 * d + sub);
    fr[1 + sub] = fv2;
    fr[1 + 2 * d + sub] = fv3;
    double const f2 = fv2, f3 = fv3;
    a7 += W.w7[2] * f2 + W.w7[3] * f3;
    a5 += W.w5[2] * f2 + W.w5[3] * f3;
    a3 += W.w3[3] * f3;
  }
  double s4 = 0.0, s5 = 0.0;
  for (int k = sub; k < c4; k += 16) s4 += double(value(1 + 4 * d + k));
  for (int k = sub; k < c5; k += 16) s5 += double(value(1 + 4 * d + c4 + k));
  a7 += W.w7[4] * s4 + W.w7[5] * s5;
  a5 += W.w5[4] * s4;
}

// FP32 nodes of region r for integrand kind K, by the 16 lanes of a row
// (lane sub takes nodes sub, sub + 16, ...): this lane's partial sums of the
// degree-7, -5 and -3 rules, and into S.f the values the fourth differences
// read.  The region's centre and half-widths stay in registers across its
// nodes; each node coordinate is one FMA from them (the (zt, lnM)-dependent
// factors come from the pair, the Delta-Sigma profile from the triple, both
// from FP64 coordinates).
template <typename T, int BLOCK, int K>
__device__ __forceinline__ void node_row(CubShared<T, BLOCK>& S, int r, int sub,
                                         const NodeConstF& kc, double& a7, double& a5,
                                         double& a3)
{
  constexpr int d = KindAxes<K>::nd;
  constexpr int nn = gm_num_nodes(d);
  float cf[d], hf[d];
#pragma unroll
  for (int a = 0; a < d; ++a) cf[a] = float(S.c[r][a]), hf[a] = float(S.h[r][a]);
  const PairC* pcr = S.pc[r];
  const float* profr = S.prof[r];
  T* fr = S.f + r * kFourNodes;
  for (int j = sub; j < nn; j += 16) {
    float yf[d];
#pragma unroll
    for (int a = 0; a < d; ++a) yf[a] = cf[a] + hf[a] * S.noff[j][a];
    float prof = 1.0f;
    if constexpr (K == Y3G_DSIGMA6D) prof = profr[S.nprof[j]];
    float const fv = integrand_node_f<K>(kc, pcr[S.npair[j]], yf, prof);
    if (j <= 4 * d) fr[j] = fv;
    int const g = S.nclass[j];
    double const fd = fv;
    a7 += S.w7[g] * fd;
    a5 += S.w5[g] * fd;
    a3 += S.w3[g] * fd;
  }
}

// Node sweep of m regions evaluating the integrand directly at every node
// (the FP64 path, arithmetic identical to the oracle's).
template <typename T, int BLOCK, int K>
__device__ void eval_nodes_direct(CubShared<T, BLOCK>& S, int m, Tabs<T> tb)
{
  constexpr int d = KindAxes<K>::nd;
  constexpr int nn = gm_num_nodes(d);
  for (int q = tid_here(); q < m * nn; q += BLOCK) {
    int const r = q / nn, j = q - r * nn;
    double y[d];
#pragma unroll
    for (int a = 0; a < d; ++a) y[a] = S.c[r][a] + S.h[r][a] * gm_offset(d, j, a);
    if constexpr (is_des_kind(K))
      S.f[q] = integrand_des_k<K>(S.ig, S.P, tb, S.des, y, mor_des_tabs_at(S.mort));
    else
      S.f[q] = integrand_k<T, K>(S.ig, S.P, tb, y);
  }
}

// Two-level scale of a split region's children: their |R7-R5| times
// clamp(c d / e, smin, 1), where d = |R7(A) + R7(B) - R7(parent)| measures
// the parent's degree-7 error and e is the parent's |R7-R5|, i.e. the
// children's degree-5 differences rescaled by how much the parent's
// overstated its degree-7 error (c = 0: no scaling).  Restated by the
// oracle (y3o_quad.c two_level_scale).
__host__ __device__ inline double two_level_scale(double d, double e, double c, double smin)
{
  if (!(c > 0.0) || !(e > 0.0)) return 1.0;
  double const s = c * d / e;
  return s < smin ? smin : (s > 1.0 ? 1.0 : s);
}

// Budget extension: an integral whose call limit `lim` would cut its next
// refinement step while its reported error `err` is still above ratio x its
// tolerance `tol` gets its limit doubled, up to ext_eval (while lim <
// ext_eval); it then refines exactly as a run with its final limit from the
// start.  Such integrals are the ones
// the base budget leaves far from converged; those near their tolerance,
// whose |R7-R5| sum over-states the true error, stop at the base budget.
// Restated by the oracle (y3o_quad.c ext_budget).
__host__ __device__ inline bool ext_budget(long long lim, long long ext_eval, double ratio,
                                           double err, double tol)
{
  return ratio > 0.0 && ext_eval > lim && err > ratio * tol;
}

// Evaluate the m regions whose centre / half-widths are in S.c / S.h:
// S.est, S.err, S.ax.
template <typename T, int BLOCK, bool SM = false>
__device__ __forceinline__ void eval_chunk(CubShared<T, BLOCK>& S, int m, double err_c, double err_smin)
{
  int const d = S.ig.ndim;
  int const nn = gm_num_nodes(d);
  Tabs<T> tb{S.hmf_tab, S.da, S.sci, S.hmfc_tab};
  if constexpr (sizeof(T) == 4) {
    eval_pairs<T, BLOCK, SM>(S, m, tb);
  } else {
    switch (S.ig.kind) {
      case Y3G_COUNT3D: eval_nodes_direct<T, BLOCK, Y3G_COUNT3D>(S, m, tb); break;
      case Y3G_COUNT5D: eval_nodes_direct<T, BLOCK, Y3G_COUNT5D>(S, m, tb); break;
      case Y3G_DSIGMA6D_DES: eval_nodes_direct<T, BLOCK, Y3G_DSIGMA6D_DES>(S, m, tb); break;
      case Y3G_DSIGMA3D_DES: eval_nodes_direct<T, BLOCK, Y3G_DSIGMA3D_DES>(S, m, tb); break;
      case Y3G_COUNT3D_DES: eval_nodes_direct<T, BLOCK, Y3G_COUNT3D_DES>(S, m, tb); break;
      case Y3G_COUNT5D_DES: eval_nodes_direct<T, BLOCK, Y3G_COUNT5D_DES>(S, m, tb); break;
      case Y3G_COUNT4D_CARD: eval_nodes_direct<T, BLOCK, Y3G_COUNT4D_CARD>(S, m, tb); break;
      case Y3G_POP_P1: eval_nodes_direct<T, BLOCK, Y3G_POP_P1>(S, m, tb); break;
      case Y3G_POP_I1: eval_nodes_direct<T, BLOCK, Y3G_POP_I1>(S, m, tb); break;
      case Y3G_POP_I2: eval_nodes_direct<T, BLOCK, Y3G_POP_I2>(S, m, tb); break;
      case Y3G_POP_J: eval_nodes_direct<T, BLOCK, Y3G_POP_J>(S, m, tb); break;
      default: eval_nodes_direct<T, BLOCK, Y3G_DSIGMA6D>(S, m, tb); break;
    }
  }
  lds_barrier();
  // rule sums: a row of 16 lanes per region (four regions per wave), partial
  // sums over nodes j = sub, sub + 16, ... reduced within the row by DPP; in
  // the FP32 path the row evaluates its region's nodes as it sums them
  constexpr int NG = BLOCK / 16;
  for (int r0 = 0; r0 < m; r0 += NG) {
    int const tt = tid_here();
    int const sub = tt & 15;
    int const r = r0 + (tt >> 4);
    bool const live = r < m;
    const T* fr = S.f + (live ? r : 0) * (sizeof(T) == 4 ? kFourNodes : nn);
    double a7 = 0.0, a5 = 0.0, a3 = 0.0;
    if constexpr (sizeof(T) == 4) {
      if (live) {
        NodeConstF const kc = node_const_f(S.ig, S.P, S.nfwtab);
        switch (S.ig.kind) {
          case Y3G_COUNT3D: node_row<T, BLOCK, Y3G_COUNT3D>(S, r, sub, kc, a7, a5, a3); break;
          case Y3G_COUNT5D: node_row_q<T, BLOCK, Y3G_COUNT5D>(S, r, sub, a7, a5, a3); break;
          default: node_row_q<T, BLOCK, Y3G_DSIGMA6D>(S, r, sub, a7, a5, a3); break;
        }
      }
      // the fourth differences read node values written by lanes of the
      // same row, hence of the same wave: a wave-level fence suffices
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
      if (live) {
        for (int j = sub; j < nn; j += 16) {
          int const g = gm_class(d, j);
          double const fv = double(fr[j]);
          a7 += S.w7[g] * fv;
          a5 += S.w5[g] * fv;
          a3 += S.w3[g] * fv;
        }
      }
    }
    a7 = row16_sum(a7);
    a5 = row16_sum(a5);
    a3 = row16_sum(a3);
    // fourth differences, one lane per axis; split axis = the lowest axis
    // holding the largest
    double four = -1.0;
    if (live && sub < d) {
      double const f0 = double(fr[0]);
      double const f2 = double(fr[1 + 2 * sub]) + double(fr[2 + 2 * sub]);
      double const f3 =
        double(fr[1 + 2 * d + 2 * sub]) + double(fr[2 + 2 * d + 2 * sub]);
      four = fabs(f2 - 2.0 * f0 - (f3 - 2.0 * f0) / 7.0);
    }
    double const bv = row16_max(four);
    unsigned long long const hit = __ballot(sub < d && four == bv);
    unsigned const row = (unsigned)(hit >> (tt & 48)) & 0xFFFFu;
    int best = row ? __builtin_ctz(row) : 0;
    if (live && sub == 0) {
      double vol = 1.0;
      for (int a = 0; a < d; ++a) vol *= 2.0 * S.h[r][a];
      if (!(bv > 0.0)) {
        double bw = -1.0;
        best = 0;
        for (int a = 0; a < d; ++a) {
          double const rel = 2.0 * S.h[r][a] / (S.ig.highs[a] - S.ig.lows[a]);
          if (rel > bw) bw = rel, best = a;
        }
      }
      S.est[r] = a7 * vol;
      S.err[r] = rule_error(a7, a5, a3, err_c, err_smin) * vol;
      S.ax[r] = best;
    }
  }
  lds_barrier();
}

}  // namespace y3g
