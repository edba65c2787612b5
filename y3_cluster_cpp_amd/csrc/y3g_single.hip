// y3g_single.hip -- the plain path: one integral at a time, its regions
// spread over the whole device, one launch sequence per cubature step with
// the host driving the iteration (the pattern of one
// MultiDimensionalIntegrator::integrate call per (volume, grid point),
// src/utils/CosmoSISSICUDAModule.hh:282-335, with PAGANI's device-wide
// region evaluation inside it).  Same region algorithm as the batched
// kernel; up to kSingleKmax regions split per step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "y3g_cub_common.hpp"
#include "y3g_single.hpp"

namespace y3g {

constexpr int kTile = 4096;  // regions per block in the compaction kernels

// initial partition (init_pieces per axis; one region when one == 1)
__global__ void k_s_init(SingleArgs A, int init_split, int init_lnm, int one, int N)
{
  int const r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= N) return;
  y3g_integral const& ig = *A.ig;
  int rr = r;
  for (int a = 0; a < ig.ndim; ++a) {
    int const k = one ? 1 : init_pieces(ig.kind, ig.ndim, init_split, init_lnm, a);
    int const q = rr % k;
    rr /= k;
    double const L = ig.lows[a], H = ig.highs[a];
    double const w = (H - L) / k;
    A.lo[(size_t)a * A.cap + r] = q == 0 ? L : L + q * w;
    A.hi[(size_t)a * A.cap + r] = q == k - 1 ? H : L + (q + 1) * w;
  }
  A.list[r] = r;
}

// Evaluate the regions A.list[0 .. nlist) (pairs: [A0, B0, A1, B1, ...]
// with the parents' est/err in A.pest/A.perr).
template <typename T, int BLOCK, bool SM>
__global__ __launch_bounds__(BLOCK) void k_s_eval(SingleArgs A, int nlist, int pairs)
{
  extern __shared__ char smem_raw[];
  CubShared<T, BLOCK>& S = *reinterpret_cast<CubShared<T, BLOCK>*>(smem_raw);
  int const tid = threadIdx.x;
  if (tid == 0) {
    S.ig = *A.ig;
    S.P = A.K.points[S.ig.point];
    gm_weights(S.ig.ndim, S.w7, S.w5, S.w3);
    if constexpr (sizeof(T) == 4) pair_setup(S);
  }
  __syncthreads();
  if (tid == 0) {
    S.hmf_tab = S.kHmfLds ? S.hmf : TabSel<T>::hmf(A.K) + (size_t)S.ig.point * A.K.hmf_stride;
    const T* hc = TabSel<T>::hmfc(A.K);
    S.hmfc_tab = hc ? hc + (size_t)S.ig.point * A.K.hmfc_stride : nullptr;
  }
  if constexpr (sizeof(T) == 4)
    for (int k = tid; k < kNfwN; k += BLOCK) S.nfwtab[k] = kNfwTab[k];
  stage_tables<T>(A.K, S.ig.point, S.kHmfLds ? S.hmf : nullptr, S.da, S.sci);
  __syncthreads();
  int const d = S.ig.ndim;
  int const nchunks = (nlist + kChunk - 1) / kChunk;
  for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
    int const base = c * kChunk;
    int const m = min(kChunk, nlist - base);
    for (int q = tid; q < m * d; q += BLOCK) {
      int const r = q / d, a = q - r * d;
      int const g = A.list[base + r];
      double const l = A.lo[(size_t)a * A.cap + g];
      double const h = A.hi[(size_t)a * A.cap + g];
      S.c[r][a] = 0.5 * (l + h);
      S.h[r][a] = 0.5 * (h - l);
    }
    __syncthreads();
    eval_chunk<T, BLOCK, SM>(S, m, A.K.err_c, A.K.err_smin);
    if (!pairs) {
      if (tid < m) {
        int const g = A.list[base + tid];
        A.est[g] = S.est[tid];
        A.err[g] = S.err[tid];
        A.ax[g] = (int8_t)S.ax[tid];
      }
    } else if (tid < m / 2) {
      int const j = base / 2 + tid;
      double ea = S.err[2 * tid], eb = S.err[2 * tid + 1];
      double const delta = fabs(S.est[2 * tid] + S.est[2 * tid + 1] - A.pest[j]);
      double const se = ea + eb;
      if (delta > se) {
        if (se > 0.0) {
          double const f = delta / se;
          ea *= f;
          eb *= f;
        } else {
          ea = eb = 0.5 * delta;
        }
      }
      int const ga = A.list[base + 2 * tid], gb = A.list[base + 2 * tid + 1];
      A.est[ga] = S.est[2 * tid];
      A.err[ga] = ea;
      A.ax[ga] = (int8_t)S.ax[2 * tid];
      A.est[gb] = S.est[2 * tid + 1];
      A.err[gb] = eb;
      A.ax[gb] = (int8_t)S.ax[2 * tid + 1];
    }
    __syncthreads();
  }
}

// Per-block partial sums of est and err (fixed order), then one block
// adds the partials in block order.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_s_reduce(SingleArgs A, int N)
{
  __shared__ double red[2 * (BLOCK / 64)];
  double a = 0.0, b = 0.0;
  int const per = (N + gridDim.x - 1) / gridDim.x;
  int const r0 = blockIdx.x * per, r1 = min(N, r0 + per);
  for (int r = r0 + threadIdx.x; r < r1; r += BLOCK) a += A.est[r], b += A.err[r];
  block_sum2<BLOCK>(a, b, red);
  if (threadIdx.x == 0) A.partial[2 * blockIdx.x] = a, A.partial[2 * blockIdx.x + 1] = b;
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_s_final(SingleArgs A, int nblocks)
{
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int i = 0; i < nblocks; ++i) a += A.partial[2 * i], b += A.partial[2 * i + 1];
    A.IE[0] = a;
    A.IE[1] = b;
    // histogram reset for the next kernel
  }
  if (threadIdx.x < 64) A.hc[threadIdx.x] = 0, A.hq[threadIdx.x] = 0ull;
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_s_hist(SingleArgs A, int N, double E, int eE)
{
  __shared__ int hc[64];
  __shared__ unsigned long long hq[64];
  if (threadIdx.x < 64) hc[threadIdx.x] = 0, hq[threadIdx.x] = 0ull;
  __syncthreads();
  double const qs = err_qscale(E);
  for (int r = blockIdx.x * BLOCK + threadIdx.x; r < N; r += gridDim.x * BLOCK) {
    int const b = err_bin(A.err[r], E, eE);
    atomicAdd(&hc[b], 1);
    atomicAdd(&hq[b], err_q(A.err[r], qs));
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    if (hc[threadIdx.x]) atomicAdd(&A.hc[threadIdx.x], hc[threadIdx.x]);
    if (hq[threadIdx.x]) atomicAdd(&A.hq[threadIdx.x], hq[threadIdx.x]);
  }
}

// counts of (bin < B) and (bin == B) per tile
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_s_count(SingleArgs A, int N, double E, int eE, int B)
{
  __shared__ int red[BLOCK / 64];
  int c0 = 0, c1 = 0;
  int const r0 = blockIdx.x * kTile, r1 = min(N, r0 + kTile);
  for (int r = r0 + threadIdx.x; r < r1; r += BLOCK) {
    int const b = err_bin(A.err[r], E, eE);
    c0 += (b < B);
    c1 += (b == B);
  }
  c0 = block_sum_i<BLOCK>(c0, red);
  c1 = block_sum_i<BLOCK>(c1, red);
  if (threadIdx.x == 0) A.blockcnt[2 * blockIdx.x] = c0, A.blockcnt[2 * blockIdx.x + 1] = c1;
}

// ordered write of the selected indices: tile t takes its first take0[t]
// regions with bin < B at offset off0[t] and its first take1[t] regions with
// bin == B at offset off1[t]
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_s_compact(SingleArgs A, int N, double E, int eE, int B)
{
  __shared__ int wcnt[BLOCK / 64];
  __shared__ int run;
  int const t = blockIdx.x;
  int const r0 = t * kTile, r1 = min(N, r0 + kTile);
  for (int pass = 0; pass < 2; ++pass) {
    int const take = A.blockoff[4 * t + 2 * pass + 1];
    int const off = A.blockoff[4 * t + 2 * pass];
    if (threadIdx.x == 0) run = 0;
    __syncthreads();
    for (int base = r0; base < r1; base += BLOCK) {
      int const r = base + threadIdx.x;
      bool f = false;
      if (r < r1) {
        int const b = err_bin(A.err[r], E, eE);
        f = pass == 0 ? (b < B) : (b == B);
      }
      unsigned long long const bal = __ballot(f);
      int const lane = threadIdx.x & 63, w = threadIdx.x >> 6;
      int const before = __popcll(bal & ((1ull << lane) - 1ull));
      if (lane == 0) wcnt[w] = __popcll(bal);
      __syncthreads();
      int o = run;
      for (int i = 0; i < w; ++i) o += wcnt[i];
      if (f && o + before < take) A.sel[off + o + before] = r;
      __syncthreads();
      if (threadIdx.x == 0) {
        int s = 0;
        for (int i = 0; i < BLOCK / 64; ++i) s += wcnt[i];
        run += s;
      }
      __syncthreads();
      if (run >= take) break;
    }
    __syncthreads();
  }
}

// children of the selected regions: A in the parent slot, B at N + j
__global__ void k_s_split(SingleArgs A, int N, int ns)
{
  int const j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ns) return;
  int const d = A.ig->ndim;
  int const p = A.sel[j];
  int const ax = A.ax[p];
  for (int a = 0; a < d; ++a) {
    double const l = A.lo[(size_t)a * A.cap + p];
    double const h = A.hi[(size_t)a * A.cap + p];
    if (a == ax) {
      double const mid = 0.5 * (l + h);
      A.lo[(size_t)a * A.cap + N + j] = mid;
      A.hi[(size_t)a * A.cap + N + j] = h;
      A.hi[(size_t)a * A.cap + p] = mid;
    } else {
      A.lo[(size_t)a * A.cap + N + j] = l;
      A.hi[(size_t)a * A.cap + N + j] = h;
    }
  }
  A.pest[j] = A.est[p];
  A.perr[j] = A.err[p];
  A.list[2 * j] = p;
  A.list[2 * j + 1] = N + j;
}

template <typename T, bool SM>
static hipError_t eval_sm(const SingleArgs& A, int nlist, int pairs, int grid, hipStream_t s)
{
  size_t const smem = sizeof(CubShared<T, 512>);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_s_eval<T, 512, SM>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  if (e != hipSuccess) return e;
  int const nch = (nlist + kChunk - 1) / kChunk;
  hipLaunchKernelGGL((k_s_eval<T, 512, SM>), dim3(std::max(1, std::min(grid, nch))), dim3(512),
                     smem, s, A, nlist, pairs);
  return hipGetLastError();
}
template <typename T>
static hipError_t eval(const SingleArgs& A, int nlist, int pairs, int grid, hipStream_t s)
{
  if constexpr (sizeof(T) == 4)
    if (A.K.hmfc32) return eval_sm<T, true>(A, nlist, pairs, grid, s);
  return eval_sm<T, false>(A, nlist, pairs, grid, s);
}

#define SCHK(x)                       \
  do {                                \
    hipError_t const e_ = (x);        \
    if (e_ != hipSuccess) return e_;  \
  } while (0)

hipError_t run_single(SingleArgs A, const y3g_integral& ig, int precision, int kmax,
                      int init_split, int init_lnm, int eval_grid, hipStream_t s,
                      y3g_result* out, long long* nevals)
{
  int const d = ig.ndim;
  int const nn = gm_num_nodes(d);
  int const cap = A.cap;
  long long N0 = 1;
  for (int a = 0; a < d; ++a) N0 *= init_pieces(ig.kind, d, init_split, init_lnm, a);
  int const one = N0 > cap;
  int N = one ? 1 : (int)N0;
  auto ev = [&](int nlist, int pairs) {
    return precision == Y3G_FP64 ? eval<double>(A, nlist, pairs, eval_grid, s)
                                 : eval<float>(A, nlist, pairs, eval_grid, s);
  };
  hipLaunchKernelGGL(k_s_init, dim3((N + 255) / 256), dim3(256), 0, s, A, init_split, init_lnm,
                     one, N);
  SCHK(hipGetLastError());
  SCHK(ev(N, 0));
  long long neval = (long long)N * nn;
  int status = 1;
  double IE[2] = {0.0, 0.0};
  double const sb = smooth_bound(ig.kind, A.K.hmfc32 != nullptr);
  double const es = sb > 0.0 ? A.K.err_scale : 1.0;
  // call limit, doubled up to ext_eval by the budget extension (ext_budget)
  long long lim = ig.max_eval;
  long long const ext_eval = A.K.ext_regions > 0 ? 2LL * A.K.ext_regions * nn : 0;
  int const nred = 256;
  std::vector<int> hc(64), cnt;
  std::vector<unsigned long long> hq(64);
  std::vector<int> off;
  for (;;) {
    hipLaunchKernelGGL(k_s_reduce<256>, dim3(nred), dim3(256), 0, s, A, N);
    hipLaunchKernelGGL(k_s_final<64>, dim3(1), dim3(64), 0, s, A, nred);
    SCHK(hipGetLastError());
    SCHK(hipMemcpyAsync(IE, A.IE, sizeof(IE), hipMemcpyDeviceToHost, s));
    SCHK(hipStreamSynchronize(s));
    double const I = IE[0], E = IE[1];
    double const tol = std::max(ig.eps_abs, ig.eps_rel * std::fabs(I));
    if (es * E + sb * std::fabs(I) <= tol) {
      status = 0;
      break;
    }
    int const room0 = std::min(kmax, cap - N);
    long long calls_left = (lim - neval) / (2LL * nn);
    while (calls_left < room0 &&
           ext_budget(lim, ext_eval, A.K.ext_ratio, es * E + sb * std::fabs(I), tol)) {
      lim = std::min(2 * lim, ext_eval);
      calls_left = (lim - neval) / (2LL * nn);
    }
    int const room = (int)std::min((long long)room0, calls_left);
    if (room <= 0) break;
    int eE;
    (void)std::frexp(E, &eE);
    int const gh = std::min(1024, (N + 255) / 256);
    hipLaunchKernelGGL(k_s_hist<256>, dim3(gh), dim3(256), 0, s, A, N, E, eE);
    SCHK(hipGetLastError());
    SCHK(hipMemcpyAsync(hc.data(), A.hc, 64 * sizeof(int), hipMemcpyDeviceToHost, s));
    SCHK(hipMemcpyAsync(hq.data(), A.hq, 64 * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, s));
    SCHK(hipStreamSynchronize(s));
    int B, lim0, lim1;
    select_bins(hc.data(), hq.data(), N, room, B, lim0, lim1);
    int const ns = lim0 + lim1;
    if (ns <= 0 || neval + 2LL * ns * nn > lim) break;
    int const nt = (N + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_s_count<256>, dim3(nt), dim3(256), 0, s, A, N, E, eE, B);
    SCHK(hipGetLastError());
    cnt.resize(2 * nt);
    SCHK(hipMemcpyAsync(cnt.data(), A.blockcnt, 2 * nt * sizeof(int), hipMemcpyDeviceToHost, s));
    SCHK(hipStreamSynchronize(s));
    off.assign(4 * nt, 0);
    int o = 0, left = lim0;
    for (int t = 0; t < nt; ++t) {
      int const take = std::min(cnt[2 * t], left);
      off[4 * t] = o, off[4 * t + 1] = take;
      o += take, left -= take;
    }
    left = lim1;
    for (int t = 0; t < nt; ++t) {
      int const take = std::min(cnt[2 * t + 1], left);
      off[4 * t + 2] = o, off[4 * t + 3] = take;
      o += take, left -= take;
    }
    SCHK(hipMemcpyAsync(A.blockoff, off.data(), 4 * nt * sizeof(int), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_s_compact<256>, dim3(nt), dim3(256), 0, s, A, N, E, eE, B);
    hipLaunchKernelGGL(k_s_split, dim3((ns + 255) / 256), dim3(256), 0, s, A, N, ns);
    SCHK(hipGetLastError());
    SCHK(ev(2 * ns, 1));
    N += ns;
    neval += 2LL * ns * nn;
  }
  out->value = IE[0];
  out->error = es * IE[1] + sb * std::fabs(IE[0]);
  out->prob = 0.0;
  out->neval = neval;
  out->nregions = N;
  out->status = status;
  *nevals += neval;
  return hipSuccess;
}

}  // namespace y3g
