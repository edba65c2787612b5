// y3g_kernels.hip -- gfx950 kernels of the cluster-observable integration
// path: per-point table builder, the persistent workgroup-per-integral
// Genz-Malik cubature, VEGAS with m-Cubes stratification, and the functor
// evaluation kernel used by the parity tests.
//
// Work distribution (both integrators): one launch, one workgroup per
// resident slot, integrals taken from a device work queue (atomic counter)
// in the order given by the host (largest expected cost first).  Each
// integral is processed start to finish by one workgroup, with a fixed
// reduction order, so its result does not depend on what else is in the
// batch.  Sums, variances, error estimates, the split decision and the
// VEGAS grid refinement are FP64; the integrand runs in T.
#include <hip/hip_runtime.h>

#include "y3g_cub_common.hpp"

namespace y3g {

// --------------------------------------------------------- table builder
__constant__ double c_gl8_x[4] = {0.1834346424956498, 0.5255324099163290,
                                  0.7966664774136267, 0.9602898564975363};
__constant__ double c_gl8_w[4] = {0.3626837833783620, 0.3137066458778873,
                                  0.2223810344533745, 0.1012285362903763};

// chi(z) [Mpc/h], composite 8-point Gauss-Legendre on 64 panels (the same
// sum, in the same order, as the oracle's y3o_comoving_distance).
__device__ double comoving_distance(double om, double z)
{
  int const np = 64;
  double const dz = z / np;
  double s = 0.0;
  for (int ip = 0; ip < np; ++ip) {
    double const c = (ip + 0.5) * dz, hw = 0.5 * dz;
    double ps = 0.0;
    for (int k = 0; k < 4; ++k) {
      double const zp = c + hw * c_gl8_x[k], zm = c - hw * c_gl8_x[k];
      double const ep = sqrt(om * (1 + zp) * (1 + zp) * (1 + zp) + (1 - om));
      double const em = sqrt(om * (1 + zm) * (1 + zm) * (1 + zm) + (1 - om));
      ps += c_gl8_w[k] * (1.0 / ep + 1.0 / em);
    }
    s += ps * hw;
  }
  return kDh * s;
}

__device__ __forceinline__ double lin_knot(double a, double b, int n, int i)
{
  return a + double(i) * ((b - a) / double(n - 1));
}

// The cubic-spline coefficients of point pt's HMF table (spline_coefs), by
// the nt threads of one workgroup after the table is written: one thread per
// row along lnM, then one per column along z, then the FP32 copy.
__device__ void hmf_coefs(const TableArgs& A, int pt, int tid, int nt)
{
  int const nm = A.spec.n_lnm, nz = A.spec.n_z, ld = nm + 4;
  const double* hmf64 = A.hmf64 + (size_t)pt * A.hmf_stride;
  double* c64 = A.hmfc64 + (size_t)pt * A.hmfc_stride;
  float* c32 = A.hmfc32 + (size_t)pt * A.hmfc_stride;
  for (int j = tid; j < nz; j += nt) spline_coefs(hmf64 + (size_t)j * nm, nm, 1, c64 + (size_t)(j + 2) * ld, 1);
  __syncthreads();
  for (int k = tid; k < ld; k += nt) spline_coefs(c64 + 2 * ld + k, nz, ld, c64 + k, ld);
  __syncthreads();
  for (int k = tid; k < (nz + 4) * ld; k += nt) c32[k] = float(c64[k]);
}

// One workgroup per point: the HMF coefficients of uploaded tables.
__global__ __launch_bounds__(256) void k_hmf_coefs(TableArgs A, int pt0)
{
  hmf_coefs(A, pt0 + blockIdx.x, threadIdx.x, blockDim.x);
}

// The tables of one parameter point, by the nt threads of one workgroup:
// HMF (n_z x n_lnM), D_A (n_da), Sigma_crit^-1 (n_sci), FP64 and FP32
// copies.
__device__ void build_point_tables(const TableArgs& A, int pt, const PointDev& P, int tid,
                                   int nt)
{
  y3g_table_spec const& s = A.spec;
  double* hmf64 = A.hmf64 + (size_t)pt * A.hmf_stride;
  float* hmf32 = A.hmf32 + (size_t)pt * A.hmf_stride;
  for (int k = tid; k < s.n_lnm * s.n_z; k += nt) {
    int const i = k % s.n_lnm, j = k / s.n_lnm;
    double const lnM = lin_knot(s.lnm_lo, s.lnm_hi, s.n_lnm, i);
    double const z = lin_knot(s.z_lo, s.z_hi, s.n_z, j);
    double const sig = P.p.hmf_amp * exp(-0.25 * (lnM - kLnPivotHmf)) *
                       pow(1.0 + z, -0.8);
    double const v = P.p.hmf_A * (pow(sig, -P.p.hmf_a) + 1.0) *
                     exp(-P.p.hmf_c / (sig * sig));
    hmf64[k] = v;
    hmf32[k] = float(v);
  }
  // the HMF's cubic-spline coefficients (hmf_coefs): rows along lnM, then
  // columns along z in place
  __syncthreads();
  if (s.n_lnm >= kSmoothMinKnots && s.n_z >= kSmoothMinKnots) hmf_coefs(A, pt, tid, nt);
  double* da64 = A.da64 + (size_t)pt * A.da_stride;
  float* da32 = A.da32 + (size_t)pt * A.da_stride;
  for (int i = tid; i < s.n_da; i += nt) {
    double const z = lin_knot(0.0, s.da_zmax, s.n_da, i);
    double const v = comoving_distance(P.p.omega_m, z) / P.p.h / (1.0 + z);
    da64[i] = v;
    da32[i] = float(v);
  }
  double const chis = comoving_distance(P.p.omega_m, P.p.z_source);
  double* sci64 = A.sci64 + (size_t)pt * A.sci_stride;
  float* sci32 = A.sci32 + (size_t)pt * A.sci_stride;
  for (int i = tid; i < s.n_sci; i += nt) {
    double const z = lin_knot(0.0, s.sci_zmax, s.n_sci, i);
    double const chil = comoving_distance(P.p.omega_m, z);
    double const v =
      chil < chis ? kSciConst * (1.0 + z) * chil * (chis - chil) / chis : 0.0;
    sci64[i] = v;
    sci32[i] = float(v);
  }
}

// One workgroup per parameter point.
__global__ __launch_bounds__(256) void k_tables_build(TableArgs A)
{
  int const pt = blockIdx.x;
  build_point_tables(A, pt, A.points[pt], threadIdx.x, blockDim.x);
}

// ------------------------------------------------- continuous batching
//
// The streaming form of the cubature (k_cubature_stream) takes its work from
// a ring in coherent pinned host memory instead of a fixed list: the host
// writes ring entries, then publishes them by raising StreamCtl::tail.  A
// workgroup claims an entry only once it is published (compare-and-swap on
// the device's head counter), so no claim is ever held on an entry that
// does not exist yet.  One idle workgroup at a time (a try-lock) copies the
// host's tail and close flag into device memory, where the others read
// them.  Every wait is bounded.  The launch ends as a whole: once no
// workgroup has claimed or finished an entry for idle_ticks of the
// constant-rate wall clock and none is busy, the first to notice sets
// StreamDev::quit and every workgroup exits (a workgroup leaving on its own
// would shrink the launch under light load); the host relaunches when it
// publishes more.  An integral whose tables are not ready within wait_ticks
// returns status Y3G_STATUS_NO_TABLES.
//
// The polling uses relaxed atomics of the right scope only (device memory
// at agent scope, host memory at system scope): each such access is
// coherent by itself, and none of them writes back or invalidates the L2,
// which an acquire or release would do on every poll -- under the
// workgroups that are integrating.  Ordering rests on dependencies: a
// workgroup reads a ring entry (system-scope loads, which bypass the
// caches) only after it has read a tail beyond it, and the host wrote the
// entry before raising the tail.
//
// Entries are integrals or table builds.  A table build fills one table
// slot from the PointDev the host wrote for it, writes it back
// (__threadfence) and then stores the slot's generation number; an integral
// on that slot waits for its generation, then acquires once (agent scope)
// before it reads the tables -- once per slot and generation per
// workgroup.  Builds are claimed before the integrals that follow them in
// the ring, by workgroups that are running, so the wait always ends.

template <typename W>
__device__ __forceinline__ W ld_agent(const W* p)
{
  return __hip_atomic_load(const_cast<W*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename W>
__device__ __forceinline__ W ld_system(const W* p)
{
  return __hip_atomic_load(const_cast<W*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Pinned host memory to dst, one 64-bit system-scope load at a time.
__device__ __forceinline__ void load_host_words(void* dst, const void* src, int words)
{
  auto* d = reinterpret_cast<unsigned long long*>(dst);
  auto* s = reinterpret_cast<const unsigned long long*>(src);
  for (int k = 0; k < words; ++k) d[k] = ld_system(s + k);
}

// Sleep about nap x 8128 cycles.
__device__ __forceinline__ void nap_for(int nap)
{
  for (int k = 0; k < nap; ++k) __builtin_amdgcn_s_sleep(127);
}

// The next published entry, or -1 (closed and drained, or idle too long).
__device__ long long stream_claim(const StreamArgs& Z)
{
  StreamDev* D = Z.dev;
  long long const t0 = wall_clock64();
  int nap = 1;
  for (;;) {
    if (ld_agent(&D->quit)) return -1;
    unsigned long long h = ld_agent(&D->head);
    unsigned long long const t = ld_agent(&D->tail);
    if (h < t) {
      // busy from before the claim, so that no workgroup sees an idle device
      // while a claim is under way
      __hip_atomic_fetch_add(&D->busy, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (__hip_atomic_compare_exchange_strong(&D->head, &h, h + 1, __ATOMIC_RELAXED,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        __hip_atomic_fetch_max(&D->last, (long long)wall_clock64(), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        return (long long)h;
      }
      __hip_atomic_fetch_sub(&D->busy, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      continue;
    }
    if (ld_agent(&D->close)) return -1;
    unsigned z = 0u;
    if (__hip_atomic_compare_exchange_strong(&D->lock, &z, 1u, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
      // close first: the host raises it after its last tail
      unsigned const cl = ld_system(&Z.ctl->close);
      unsigned long long const th = ld_system(&Z.ctl->tail);
      __hip_atomic_fetch_max(&D->tail, th, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (cl) __hip_atomic_store(&D->close, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&D->lock, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (th > h) {
        nap = 1;
        continue;
      }
    }
    long long const now = wall_clock64();
    long long const last = ld_agent(&D->last);
    if (ld_agent(&D->busy) == 0u && now - (last > t0 ? last : t0) > Z.idle_ticks) {
      __hip_atomic_store(&D->quit, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return -1;
    }
    nap_for(nap);
    nap = nap < 8 ? 2 * nap : 8;
  }
}

// An entry is finished: the device was active now, one fewer busy.
__device__ void stream_done(const StreamArgs& Z)
{
  __hip_atomic_fetch_max(&Z.dev->last, (long long)wall_clock64(), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_fetch_sub(&Z.dev->busy, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait (bounded) until the slot's tables are those of generation gen, then
// acquire them.
__device__ bool stream_wait_tables(const StreamArgs& Z, int slot, unsigned gen)
{
  long long const t0 = wall_clock64();
  while (ld_agent(Z.ready + slot) != gen) {
    if (wall_clock64() - t0 > Z.wait_ticks) return false;
    __builtin_amdgcn_s_sleep(64);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return true;
}

// Write an entry's result (system-scope stores to pinned host memory, past
// the caches), wait until they are done, then store its sequence number
// (entry + 1), which the host polls.
__device__ void stream_publish(const StreamArgs& Z, long long item, const y3g_result& r)
{
  size_t const k = (size_t)(item % (long long)Z.ring_cap);
  static_assert(sizeof(y3g_result) % 8 == 0, "results are written as 64-bit words");
  auto const* w = reinterpret_cast<const unsigned long long*>(&r);
  auto* d = reinterpret_cast<unsigned long long*>(Z.res + k);
  for (int i = 0; i < (int)(sizeof(y3g_result) / 8); ++i)
    __hip_atomic_store(d + i, w[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __builtin_amdgcn_s_waitcnt(0);  // every vector memory operation of the wave done
  __hip_atomic_store(Z.seq + k, (unsigned long long)item + 1ull, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_SYSTEM);
}

#ifndef Y3G_CUB_MIN_WAVES
#define Y3G_CUB_MIN_WAVES 6
#endif
// FP32: six waves per SIMD (at most 80 VGPRs), so that three 512-thread
// workgroups share a CU (their LDS fits too); FP64: the compiler's choice
// The cubature's work loop: integrals from the device queue (k_cubature) or
// from the host ring (STREAM, k_cubature_stream); the region algorithm of
// an integral is the same code in both.
template <typename T, int BLOCK, bool STREAM, bool SM>
__device__ __forceinline__ void cubature_items(const KernArgs& A, const StreamArgs& Z)
{
  extern __shared__ char smem_raw[];
  CubShared<T, BLOCK>& S = *reinterpret_cast<CubShared<T, BLOCK>*>(smem_raw);
  // the parents of the chunk being split
  __shared__ ParentStage stage;
  int const tid = threadIdx.x;
  int const slot = blockIdx.x;
  int const cap = A.cap;
  double* lo = A.slab_lo + (size_t)slot * kMaxDim * cap;
  double* hi = A.slab_hi + (size_t)slot * kMaxDim * cap;
  double* est = A.slab_est + (size_t)slot * cap;
  double* err = A.slab_err + (size_t)slot * cap;
  double* raw = A.slab_raw + (size_t)slot * cap;
  int8_t* axs = A.slab_ax + (size_t)slot * cap;
  if (tid == 0) S.point = -1, S.pgen = 0u, S.setup_kind = -1;
  if constexpr (sizeof(T) == 4)
    for (int k = tid; k < kNfwN; k += BLOCK) S.nfwtab[k] = kNfwTab[k];
  if constexpr (sizeof(T) == 8) {
    MorDesTabs const c = mor_des_const_tabs();
    for (int k = tid; k < kMorDesWords; k += BLOCK)
      S.mort[k] = k < 14 ? c.si[k] : k < 56 ? c.ls[k - 14] : k < 644 ? c.sg[k - 56] : c.sk[k - 644];
  }
  for (;;) {
    __syncthreads();
    int iid = 0;
    if constexpr (STREAM) {
      if (tid == 0) {
        long long const it = stream_claim(Z);
        S.sitem = it;
        if (it >= 0) {
          StreamItem const* e = Z.ring + (size_t)(it % (long long)Z.ring_cap);
          StreamItem h;
          load_host_words(&h, e, sizeof(StreamItem) / 8);
          S.job = h.job;
          S.sslot = h.slot;
          S.sgen = h.gen;
          S.ig = h.ig;
          S.ig.point = h.slot;
        }
      }
      __syncthreads();
      if (S.sitem < 0) break;
      if (S.job == kStreamJobTables) {
        // build the slot's tables, then release its generation
        if (tid == 0) load_host_words(&S.P, Z.hpoints + S.sslot, sizeof(PointDev) / 8);
        __syncthreads();
        build_point_tables(Z.tab, S.sslot, S.P, tid, BLOCK);
        if (tid == 0) Z.points[S.sslot] = S.P;
        __threadfence();
        __syncthreads();
        if (tid == 0) {
          __hip_atomic_store(Z.ready + S.sslot, S.sgen, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
          y3g_result r{};
          stream_publish(Z, S.sitem, r);
          stream_done(Z);
          S.point = -1;  // S.P now holds this slot's constants: reload
        }
        continue;
      }
      // tables this workgroup has staged already are ready
      if (tid == 0)
        S.job = (S.point == S.sslot && S.pgen == S.sgen) ||
                    stream_wait_tables(Z, S.sslot, S.sgen)
                  ? 0
                  : -1;
      __syncthreads();
      if (S.job < 0) {
        if (tid == 0) {
          y3g_result r{};
          r.status = Y3G_STATUS_NO_TABLES;
          stream_publish(Z, S.sitem, r);
          stream_done(Z);
        }
        continue;
      }
    } else {
      if (tid == 0) S.item = atomicAdd(A.queue, 1);
      __syncthreads();
      int const item = S.item;
      if (item >= A.n) break;
      iid = A.order ? A.order[item] : item;
    }
    if (tid == 0) {
      if constexpr (!STREAM) S.ig = A.igs[iid];
      gm_weights(S.ig.ndim, S.w7, S.w5, S.w3);
      // the node maps depend on the kind and dimension only; the queue's
      // longest-first order keeps a workgroup on one kind for long stretches
      if constexpr (sizeof(T) == 4) {
        int const key = S.ig.kind * 16 + S.ig.ndim;
        if (S.setup_kind != key) pair_setup(S), S.setup_kind = key;
      }
    }
    __syncthreads();
    int const pt = S.ig.point;
    // a streamed slot's tables change with its generation
    unsigned const pgen = STREAM ? S.sgen : 0u;
    if (pt != S.point || pgen != S.pgen) {
      if (tid == 0) {
        S.P = A.points[pt];
        S.hmf_tab = S.kHmfLds ? S.hmf : TabSel<T>::hmf(A) + (size_t)pt * A.hmf_stride;
        const T* hc = TabSel<T>::hmfc(A);
        S.hmfc_tab = hc ? hc + (size_t)pt * A.hmfc_stride : nullptr;
      }
      if (A.des) {
        long long const o0 = A.des_off[pt], nd = A.des_off[pt + 1] - o0;
        bool const lds = sizeof(T) == 8 && nd <= kDesLds;
        if constexpr (sizeof(T) == 8)
          if (lds)
            for (long long k = tid; k < nd; k += BLOCK) S.desl[k] = A.des[o0 + k];
        if (tid == 0) S.des = lds ? S.desl : A.des + o0;
      } else if (tid == 0) {
        S.des = nullptr;
      }
      stage_tables<T>(A, pt, S.kHmfLds ? S.hmf : nullptr, S.da, S.sci);
      __syncthreads();
      if (tid == 0) S.point = pt, S.pgen = pgen;
    }
    int const d = S.ig.ndim;
    int const nn = gm_num_nodes(d);
    long long N0 = 1;
    for (int a = 0; a < d; ++a) N0 *= init_pieces(S.ig.kind, d, A.init_split, A.init_lnm, a);
    bool const one = N0 > cap;
    int N = one ? 1 : (int)N0;
    // initial partition, region r -> digits per axis (axis 0 fastest)
    for (int r = tid; r < N; r += BLOCK) {
      int rr = r;
      for (int a = 0; a < d; ++a) {
        int const k = one ? 1 : init_pieces(S.ig.kind, d, A.init_split, A.init_lnm, a);
        int const q = rr % k;
        rr /= k;
        double const L = S.ig.lows[a], H = S.ig.highs[a];
        double const w = (H - L) / k;
        lo[(size_t)a * cap + r] = q == 0 ? L : L + q * w;
        hi[(size_t)a * cap + r] = q == k - 1 ? H : L + (q + 1) * w;
      }
    }
    __syncthreads();
    // evaluate the initial regions
    for (int base = 0; base < N; base += kChunk) {
      int const m = min(kChunk, N - base);
      for (int q = tid; q < m * d; q += BLOCK) {
        int const r = q / d, a = q - r * d;
        double const l = lo[(size_t)a * cap + base + r];
        double const h = hi[(size_t)a * cap + base + r];
        S.c[r][a] = 0.5 * (l + h);
        S.h[r][a] = 0.5 * (h - l);
      }
      __syncthreads();
      eval_chunk<T, BLOCK, SM>(S, m, A.err_c, A.err_smin);
      if (tid < m) {
        est[base + tid] = S.est[tid];
        err[base + tid] = S.err[tid];
        raw[base + tid] = S.err[tid];
        axs[base + tid] = (int8_t)S.ax[tid];
      }
      __syncthreads();
    }
    long long neval = (long long)N * nn;
    int status = 1;
    // totals: summed once here and at the end, updated by each step's
    // (children - parent) differences in between
    double I = 0.0, E = 0.0;
    double const sb = smooth_bound(S.ig.kind, S.hmfc_tab != nullptr);
    double const es = sb > 0.0 ? A.err_scale : 1.0;  // margin with the smoothed HMF
    {
      double a = 0.0, b = 0.0;
      for (int r = tid; r < N; r += BLOCK) a += est[r], b += err[r];
      block_sum2<BLOCK>(a, b, S.red);
      I = a;
      E = b;
    }
    // the call limit; doubled, up to ext_eval, each time it would cut a step
    // with the estimate still above ext_ratio x the tolerance (ext_budget)
    long long lim = S.ig.max_eval;
    long long const ext_eval = A.ext_regions > 0 ? 2LL * A.ext_regions * nn : 0;
    for (;;) {
      double const tol = fmax(S.ig.eps_abs, S.ig.eps_rel * fabs(I));
      if (es * E + sb * fabs(I) <= tol) {
        status = 0;
        break;
      }
      // selection: error bins by binary exponent relative to E (bin 0 holds
      // the largest errors), error mass per bin in 2^-40 E fixed point;
      // split bins 0..B-1 and, in index order, bin B, where B covers at
      // least kAlpha of E and at least max(64, N/8) regions, at most `room`
      // at most kmax splits, within the slab and the call limit
      int const room0 = min(A.kmax, cap - N);
      long long left = (lim - neval) / (2LL * nn);
      while (left < room0 && ext_budget(lim, ext_eval, A.ext_ratio, es * E + sb * fabs(I), tol)) {
        lim = min(2 * lim, ext_eval);
        left = (lim - neval) / (2LL * nn);
      }
      int const room = (int)min((long long)room0, left);
      if (room <= 0) break;
      int eE;
      (void)frexp(E, &eE);
      // wave w takes regions [sb, se), contiguous and in wave order, so
      // that index order is kept without block-wide steps: per-wave bin
      // counts (summed into the histogram), then each wave's share of the
      // selection from their exclusive prefix
      constexpr int NW = BLOCK / 64;
      int const lane = tid & 63, wv = tid >> 6;
      int const seg = (N + NW - 1) / NW;
      int const sb = min(N, wv * seg), se = min(N, sb + seg);
      for (int k = tid; k < NW * 64; k += BLOCK) (&S.wc[0][0])[k] = 0;
      if (tid < 64) S.hq[tid] = 0ull;
      __syncthreads();
      {
        double const qs = err_qscale(E);
        int r = sb + lane;
        // four loads in flight per lane before their bins are counted
        for (; r + 192 < se; r += 256) {
          double ev[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) ev[u] = err[r + 64 * u];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            int const b = err_bin(ev[u], E, eE);
            atomicAdd(&S.wc[wv][b], 1);
            atomicAdd(&S.hq[b], err_q(ev[u], qs));
          }
        }
        for (; r < se; r += 64) {
          double const e = err[r];
          int const b = err_bin(e, E, eE);
          atomicAdd(&S.wc[wv][b], 1);
          atomicAdd(&S.hq[b], err_q(e, qs));
        }
      }
      __syncthreads();
      if (tid < 64) {
        int c = 0;
        for (int v = 0; v < NW; ++v) c += S.wc[v][tid];
        S.hc[tid] = c;
      }
      __syncthreads();
      if (tid == 0) select_bins(S.hc, S.hq, N, room, S.selB, S.lim0, S.lim1);
      __syncthreads();
      int const B = S.selB, lim0 = S.lim0, lim1 = S.lim1;
      int const ns = lim0 + lim1;
      if (ns <= 0 || neval + 2LL * ns * nn > lim) break;
      // ordered compaction: bins < B in index order to sel[0, lim0), bin B
      // in index order to sel[lim0, lim0 + lim1)
      {
        int c0 = 0;
        for (int b = lane; b < B; b += 64) c0 += S.wc[wv][b];
        c0 = wave_sum_i(c0);
        if (lane == 0) S.ired[wv] = c0, S.ired1[wv] = S.wc[wv][B];
      }
      __syncthreads();
      {
        int off0 = 0, off1 = 0;
        for (int v = 0; v < wv; ++v) off0 += S.ired[v], off1 += S.ired1[v];
        unsigned long long const below = (1ull << (tid_here() & 63)) - 1ull;
        for (int base = sb; base < se && (off0 < lim0 || off1 < lim1); base += 64) {
          int const r = base + lane;
          bool f0 = false, f1 = false;
          if (r < se) {
            int const b = err_bin(err[r], E, eE);
            f0 = b < B;
            f1 = b == B;
          }
          unsigned long long const bal0 = __ballot(f0), bal1 = __ballot(f1);
          int const o0 = off0 + __popcll(bal0 & below), o1 = off1 + __popcll(bal1 & below);
          if (f0 && o0 < lim0) S.sel[o0] = r;
          if (f1 && o1 < lim1) S.sel[lim0 + o1] = r;
          off0 += __popcll(bal0);
          off1 += __popcll(bal1);
        }
      }
      __syncthreads();
      // split and evaluate, kChunk/2 pairs at a time; the next chunk's
      // parents are loaded while this one is evaluated
      double dI = 0.0, dE = 0.0;
      store_stage<BLOCK>(&stage, fetch_parents<BLOCK>(S.sel, 0, min(kChunk / 2, ns), cap, lo,
                                                      hi, est, err, raw, axs));
      for (int pb = 0; pb < ns; pb += kChunk / 2) {
        int const np = min(kChunk / 2, ns - pb);
        bool const more = pb + kChunk / 2 < ns;
        __syncthreads();
        ParentStage const& st = stage;
        for (int q = tid; q < np * d; q += BLOCK) {
          int const j = q / d, a = q - j * d;
          int const p = S.sel[pb + j];
          int const ax = (int)(int8_t)(st.axw[j] >> (8 * (int)(reinterpret_cast<uintptr_t>(axs + p) & 3)));
          double const l = st.lo[j][a];
          double const h = st.hi[j][a];
          if (a == ax) {
            double const mid = 0.5 * (l + h);
            S.c[2 * j][a] = 0.5 * (l + mid);
            S.h[2 * j][a] = 0.5 * (mid - l);
            S.c[2 * j + 1][a] = 0.5 * (mid + h);
            S.h[2 * j + 1][a] = 0.5 * (h - mid);
            lo[(size_t)a * cap + N + pb + j] = mid;
            hi[(size_t)a * cap + N + pb + j] = h;
            hi[(size_t)a * cap + p] = mid;
          } else {
            S.c[2 * j][a] = S.c[2 * j + 1][a] = 0.5 * (l + h);
            S.h[2 * j][a] = S.h[2 * j + 1][a] = 0.5 * (h - l);
            lo[(size_t)a * cap + N + pb + j] = l;
            hi[(size_t)a * cap + N + pb + j] = h;
          }
        }
        if (tid < np) {
          S.pest[tid] = st.est[tid];
          S.perr[tid] = st.err[tid];
          S.praw[tid] = st.raw[tid];
        }
        // the next chunk's parents: loaded now, stored to the stage after
        // this chunk's evaluation
        StageWords<BLOCK> next;
        if (more)
          next = fetch_parents<BLOCK>(S.sel, pb + kChunk / 2, min(kChunk / 2, ns - pb - kChunk / 2),
                                      cap, lo, hi, est, err, raw, axs);
        lds_barrier();
        eval_chunk<T, BLOCK, SM>(S, 2 * np, A.err_c, A.err_smin);
        if (more) store_stage<BLOCK>(&stage, next);
        // results; the next chunk's first barrier orders these LDS reads
        // before its writes
        if (tid < np) {
          int const p = S.sel[pb + tid];
          double const ra = S.err[2 * tid], rb = S.err[2 * tid + 1];
          double const delta =
            fabs(S.est[2 * tid] + S.est[2 * tid + 1] - S.pest[tid]);
          double const sc = two_level_scale(delta, S.praw[tid], A.tl_c, A.tl_smin);
          double ea = ra * sc, eb = rb * sc;
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
          dI += S.est[2 * tid] + S.est[2 * tid + 1] - S.pest[tid];
          dE += ea + eb - S.perr[tid];
          est[p] = S.est[2 * tid];
          err[p] = ea;
          raw[p] = ra;
          axs[p] = (int8_t)S.ax[2 * tid];
          est[N + pb + tid] = S.est[2 * tid + 1];
          err[N + pb + tid] = eb;
          raw[N + pb + tid] = rb;
          axs[N + pb + tid] = (int8_t)S.ax[2 * tid + 1];
        }
      }
      __syncthreads();
      block_sum2<BLOCK>(dI, dE, S.red);
      I += dI;
      E += dE;
      N += ns;
      neval += 2LL * ns * nn;
    }
    {
      double a = 0.0, b = 0.0;
      for (int r = tid; r < N; r += BLOCK) a += est[r], b += err[r];
      block_sum2<BLOCK>(a, b, S.red);
      I = a;
      E = b;
    }
    if (tid == 0) {
      y3g_result res;
      res.value = I;
      res.error = es * E + sb * fabs(I);
      res.prob = 0.0;
      res.neval = neval;
      res.nregions = N;
      res.status = status;
      if constexpr (STREAM) {
        stream_publish(Z, S.sitem, res);
        stream_done(Z);
      } else {
        A.out[iid] = res;
      }
      atomicAdd(A.nevals, (unsigned long long)neval);
    }
  }
}

// SM: the FP32 pair factors read the smoothed HMF (the launchers pick it
// when the arguments carry the coefficients); each form has its own
// register allocation.  The FP64 path reads tb.hmfc at run time.
template <typename T, int BLOCK, bool SM = false>
__global__ __launch_bounds__(BLOCK, sizeof(T) == 4 ? Y3G_CUB_MIN_WAVES : 1) void k_cubature(KernArgs A)
{
  cubature_items<T, BLOCK, false, SM>(A, StreamArgs{});
}

template <typename T, int BLOCK, bool SM = false>
__global__ __launch_bounds__(BLOCK, sizeof(T) == 4 ? Y3G_CUB_MIN_WAVES : 1) void k_cubature_stream(
  KernArgs A, StreamArgs Z)
{
  cubature_items<T, BLOCK, true, SM>(A, Z);
}

// ------------------------------------------------------------------ VEGAS
//
// m-Cubes stratification + VEGAS importance grid (identical to the oracle's
// y3o_vegas restatement): g = floor((calls/2)^(1/d)) cubes per axis,
// p = max(2, calls/g^d) samples per cube, Philox4x32-10 with key = seed and
// counter (cube, sample, iteration, block), 53-bit uniforms (FP64) or 24-bit
// uniforms (FP32 path); Lepage refinement with 3-point smoothing and damping
// alpha; iterations < itskip only adapt; inverse-variance combination; the
// reported error is kVegasNsig standard deviations (inflated by
// sqrt(chi2/dof) when that exceeds 1) so that it bounds the true error.

constexpr int kNbMax = 128;
constexpr double kVegasNsig = 3.0;

template <typename T, int BLOCK> struct VegShared {
  T hmf[kMaxHmf];
  T da[kMaxDa];
  T sci[kMaxSci];
  double xg[kMaxDim][kNbMax + 1];
  double hist[kMaxDim][kNbMax];
  double d1[kMaxDim][kNbMax];
  double xn[kMaxDim][kNbMax + 1];  // the refined grid, before it replaces xg
  double red[2 * (BLOCK / 64)];
  y3g_integral ig;
  PointDev P;
  const double* des;  // the point's DES data (Y3G_DSIGMA6D_DES)
  const T* hmfc;      // the point's HMF coefficients (global), null: bilinear
  int item, point, stop;
};

template <typename T> struct Uniform;
template <> struct Uniform<double> {
  static constexpr int words_per_dim = 2;
  static __device__ double get(const uint32_t* w, int k)
  {
    return (double(w[2 * k] >> 5) * 67108864.0 + double(w[2 * k + 1] >> 6)) *
           (1.0 / 9007199254740992.0);
  }
};
template <> struct Uniform<float> {
  static constexpr int words_per_dim = 1;
  static __device__ double get(const uint32_t* w, int k)
  {
    return double(w[k] >> 8) * (1.0 / 16777216.0) + (0.5 / 16777216.0);
  }
};

// One integral of kind K by VEGAS + m-Cubes (the whole workgroup).
template <typename T, int BLOCK, int K>
__device__ __forceinline__ void vegas_one(VegShared<T, BLOCK>& S, const KernArgs& A, int iid,
                                          unsigned long long& nev_total)
{
  int const tid = threadIdx.x;
  int const nb = A.vegas_nbins;
  constexpr int d = KindAxes<K>::nd;
  Tabs<T> tb{S.hmf, S.da, S.sci, S.hmfc};
  for (int q = tid; q < d * (nb + 1); q += BLOCK) {
    int const a = q / (nb + 1), b = q - a * (nb + 1);
    S.xg[a][b] = double(b) / nb;
  }
  // calls per iteration, fewer when the integral's call limit could not
  // hold all iterations (a small max_eval still runs every iteration)
  long long calls = A.vegas_calls;
  long long const per_it = S.ig.max_eval / (A.vegas_itmax > 0 ? A.vegas_itmax : 1);
  if (per_it < calls) calls = per_it;
  if (calls < 2) calls = 2;
  int g = (int)floor(pow(calls / 2.0, 1.0 / d));
  if (g < 1) g = 1;
  long long ncubes = 1;
  for (int a = 0; a < d; ++a) ncubes *= g;
  long long p = calls / ncubes;
  if (p < 2) p = 2;
  double range[d], vol = 1.0;
  for (int a = 0; a < d; ++a)
    range[a] = S.ig.highs[a] - S.ig.lows[a], vol *= range[a];
  uint32_t const k0 = (uint32_t)S.ig.seed, k1 = (uint32_t)(S.ig.seed >> 32);
  int const nblk = (Uniform<T>::words_per_dim * d + 3) / 4;
  double wsum = 0.0, isum = 0.0, i2sum = 0.0, I = 0.0, sig = 0.0;
  int nacc = 0, status = 1;
  long long neval = 0;
  if (tid == 0) S.stop = 0;
  __syncthreads();
  for (int it = 0; it < A.vegas_itmax; ++it) {
    for (int q = tid; q < d * nb; q += BLOCK) S.hist[q / nb][q % nb] = 0.0;
    __syncthreads();
    double sI = 0.0, sV = 0.0;
    // each wave takes a contiguous run of cubes (axis 0 fastest), its lanes
    // 64 adjacent cubes at a time: a lane's cube digits advance by 64 with a
    // carry instead of 64-bit divisions, and its samples stay in the same
    // grid bin along every axis but the first for long stretches, so a lane
    // sums its contribution to a bin and adds it to the shared histogram
    // only when the bin changes
    // (the FP64 kernel, at its register limit, adds every sample directly:
    // the lane sums would live in scratch memory there)
    constexpr bool kLocalHist = sizeof(T) == 4;
    constexpr int NW = BLOCK / 64;
    long long const perw = ((ncubes + NW - 1) / NW + 63) / 64 * 64;
    long long const wbeg = min(ncubes, (long long)(tid >> 6) * perw);
    long long const wend = min(ncubes, wbeg + perw);
    int ck[d];
    {
      long long cc = wbeg + (tid & 63);
      for (int a = 0; a < d; ++a) ck[a] = (int)(cc % g), cc /= g;
    }
    int hb[d];
    double hs[d];
    for (int a = 0; a < d; ++a) hb[a] = -1, hs[a] = 0.0;
    for (long long c = wbeg + (tid & 63); c < wend; c += 64) {
      if (c >= wbeg + 64) {
        ck[0] += 64;
        for (int a = 0; a < d - 1 && ck[a] >= g; ++a) {
          int const q = ck[a] / g;
          ck[a] -= q * g;
          ck[a + 1] += q;
        }
      }
      double s1 = 0.0, s2 = 0.0;
      for (long long j = 0; j < p; ++j) {
        uint32_t w[12];
        for (int blk = 0; blk < nblk; ++blk) {
          uint32_t ctr[4] = {(uint32_t)c, (uint32_t)j, (uint32_t)it,
                             (uint32_t)blk};
          philox(ctr, k0, k1);
          w[4 * blk] = ctr[0], w[4 * blk + 1] = ctr[1];
          w[4 * blk + 2] = ctr[2], w[4 * blk + 3] = ctr[3];
        }
        double y[d], jac = vol;
        int bins[d];
        for (int a = 0; a < d; ++a) {
          double const u = Uniform<T>::get(w, a);
          double const z = (ck[a] + u) / g * nb;
          int b = (int)z;
          if (b > nb - 1) b = nb - 1;
          double const wb = S.xg[a][b + 1] - S.xg[a][b];
          y[a] = S.ig.lows[a] + (S.xg[a][b] + (z - b) * wb) * range[a];
          jac *= nb * wb;
          bins[a] = b;
        }
        double fx;
        if constexpr (is_des_kind(K))
          fx = integrand_des_k<K>(S.ig, S.P, tb, S.des, y) * jac;
        else
          fx = double(integrand_k<T, K>(S.ig, S.P, tb, y)) * jac;
        s1 += fx;
        s2 += fx * fx;
        for (int a = 0; a < d; ++a) {
          if constexpr (kLocalHist) {
            if (bins[a] != hb[a]) {
              if (hb[a] >= 0) atomicAdd(&S.hist[a][hb[a]], hs[a]);
              hb[a] = bins[a];
              hs[a] = 0.0;
            }
            hs[a] += fx * fx;
          } else {
            atomicAdd(&S.hist[a][bins[a]], fx * fx);
          }
        }
      }
      double const m = s1 / p;
      double const v = (s2 - s1 * m) / (double(p) * (p - 1));
      sI += m;
      sV += v > 0.0 ? v : 0.0;
    }
    if constexpr (kLocalHist)
      for (int a = 0; a < d; ++a)
        if (hb[a] >= 0) atomicAdd(&S.hist[a][hb[a]], hs[a]);
    block_sum2<BLOCK>(sI, sV, S.red);
    neval += ncubes * p;
    double const Ii = sI / ncubes;
    double const Vi = sV / (double(ncubes) * ncubes);
    if (it >= A.vegas_itskip) {
      double const wi = Vi > 0.0 ? 1.0 / Vi : 1e300;
      wsum += wi;
      isum += wi * Ii;
      i2sum += wi * Ii * Ii;
      ++nacc;
      I = isum / wsum;
      sig = sqrt(1.0 / wsum);
      double const dof = nacc - 1;
      double const chi = dof > 0 ? (i2sum - wsum * I * I) / dof : 0.0;
      double const sigr = kVegasNsig * (chi > 1.0 ? sig * sqrt(chi) : sig) +
                          smooth_bound(S.ig.kind, S.hmfc != nullptr) * fabs(I);
      if (sigr <= fmax(S.ig.eps_abs, S.ig.eps_rel * fabs(I))) {
        status = 0;
        break;
      }
    }
    if (neval + ncubes * p > S.ig.max_eval) break;
    if (it + 1 < A.vegas_itmax) {
      // Lepage refinement, one thread per axis
      if (tid < d) {
        int const a = tid;
        double* hs = S.hist[a];
        double* dd = S.d1[a];
        double sum = 0.0;
        if (nb == 1) {
          dd[0] = hs[0];
        } else {
          dd[0] = 0.5 * (hs[0] + hs[1]);
          dd[nb - 1] = 0.5 * (hs[nb - 2] + hs[nb - 1]);
          for (int b = 1; b < nb - 1; ++b)
            dd[b] = (hs[b - 1] + hs[b] + hs[b + 1]) / 3.0;
        }
        for (int b = 0; b < nb; ++b) sum += dd[b];
        if (sum > 0.0) {
          double rs = 0.0;
          for (int b = 0; b < nb; ++b) {
            double const q = dd[b] / sum;
            dd[b] = (q > 0.0 && q < 1.0) ? pow((q - 1.0) / log(q), A.vegas_alpha)
                                         : 0.0;
            rs += dd[b];
          }
          if (rs > 0.0) {
            double const target = rs / nb;
            double* xn = S.xn[a];
            int kk = -1;
            double dr = 0.0;
            xn[0] = 0.0;
            for (int i = 1; i < nb; ++i) {
              while (dr < target && kk < nb - 1) {
                ++kk;
                dr += dd[kk];
              }
              dr -= target;
              xn[i] = S.xg[a][kk + 1] -
                      (S.xg[a][kk + 1] - S.xg[a][kk]) * dr / dd[kk];
            }
            xn[nb] = 1.0;
            for (int i = 0; i <= nb; ++i) S.xg[a][i] = xn[i];
          }
        }
      }
      __syncthreads();
    }
  }
  if (tid == 0) {
    double const dof = nacc - 1;
    double const chi = dof > 0 ? (i2sum - wsum * I * I) / dof : 0.0;
    y3g_result res;
    res.value = I;
    res.error = kVegasNsig * (chi > 1.0 ? sig * sqrt(chi) : sig) +
                smooth_bound(S.ig.kind, S.hmfc != nullptr) * fabs(I);
    res.prob = chi;
    res.neval = neval;
    res.nregions = 0;
    res.status = status;
    A.out[iid] = res;
    nev_total += (unsigned long long)neval;
  }
}

template <typename T, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_vegas(KernArgs A)
{
  extern __shared__ char smem_raw[];
  VegShared<T, BLOCK>& S = *reinterpret_cast<VegShared<T, BLOCK>*>(smem_raw);
  int const tid = threadIdx.x;
  if (tid == 0) S.point = -1;
  unsigned long long nev_total = 0;
  for (;;) {
    __syncthreads();
    if (tid == 0) S.item = atomicAdd(A.queue, 1);
    __syncthreads();
    int const item = S.item;
    if (item >= A.n) break;
    int const iid = A.order ? A.order[item] : item;
    if (tid == 0) S.ig = A.igs[iid];
    __syncthreads();
    int const pt = S.ig.point;
    if (pt != S.point) {
      if (tid == 0) {
        S.P = A.points[pt];
        S.des = A.des ? A.des + A.des_off[pt] : nullptr;
        const T* hc = TabSel<T>::hmfc(A);
        S.hmfc = hc ? hc + (size_t)pt * A.hmfc_stride : nullptr;
      }
      stage_tables<T>(A, pt, S.hmf, S.da, S.sci);
      __syncthreads();
      if (tid == 0) S.point = pt;
    }
    switch (S.ig.kind) {
      case Y3G_COUNT3D: vegas_one<T, BLOCK, Y3G_COUNT3D>(S, A, iid, nev_total); break;
      case Y3G_COUNT5D: vegas_one<T, BLOCK, Y3G_COUNT5D>(S, A, iid, nev_total); break;
      case Y3G_DSIGMA6D_DES:
        if constexpr (sizeof(T) == 8) vegas_one<T, BLOCK, Y3G_DSIGMA6D_DES>(S, A, iid, nev_total);
        break;
      case Y3G_DSIGMA3D_DES:
        if constexpr (sizeof(T) == 8) vegas_one<T, BLOCK, Y3G_DSIGMA3D_DES>(S, A, iid, nev_total);
        break;
      case Y3G_COUNT3D_DES:
        if constexpr (sizeof(T) == 8) vegas_one<T, BLOCK, Y3G_COUNT3D_DES>(S, A, iid, nev_total);
        break;
      case Y3G_COUNT5D_DES:
        if constexpr (sizeof(T) == 8) vegas_one<T, BLOCK, Y3G_COUNT5D_DES>(S, A, iid, nev_total);
        break;
      case Y3G_COUNT4D_CARD:
        if constexpr (sizeof(T) == 8) vegas_one<T, BLOCK, Y3G_COUNT4D_CARD>(S, A, iid, nev_total);
        break;
      case Y3G_POP_P1:
        if constexpr (sizeof(T) == 8) vegas_one<T, BLOCK, Y3G_POP_P1>(S, A, iid, nev_total);
        break;
      case Y3G_POP_I1:
        if constexpr (sizeof(T) == 8) vegas_one<T, BLOCK, Y3G_POP_I1>(S, A, iid, nev_total);
        break;
      case Y3G_POP_I2:
        if constexpr (sizeof(T) == 8) vegas_one<T, BLOCK, Y3G_POP_I2>(S, A, iid, nev_total);
        break;
      case Y3G_POP_J:
        if constexpr (sizeof(T) == 8) vegas_one<T, BLOCK, Y3G_POP_J>(S, A, iid, nev_total);
        break;
      default: vegas_one<T, BLOCK, Y3G_DSIGMA6D>(S, A, iid, nev_total); break;
    }
  }
  if (tid == 0 && nev_total) atomicAdd(A.nevals, nev_total);
}

// ------------------------------------------------------- functor eval
template <typename T>
__global__ __launch_bounds__(256) void k_eval_functor(FunctorArgs F)
{
  size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= F.n) return;
  const double* x = F.in + i * F.nin;
  const double* q = F.par;
  PointDev const& P = F.points[0];
  Tabs<T> tb{F.hmf ? (const T*)F.hmf : nullptr, (const T*)F.da, (const T*)F.sci};
  T r = T(0);
  switch (F.which) {
    case Y3G_F_ROFFSET_GAMMA: r = roffset_gamma<T>(T(x[0]), T(q[0])); break;
    case Y3G_F_LO_LC:
      r = lo_lc<T>(T(x[0]), T(x[1]), T(x[2]), T(q[0]), T(q[1]), T(q[2]), T(q[3]));
      break;
    case Y3G_F_ZO_ZT_DES:
      r = gaussian<T>(T(x[0]), T(x[1]), sigma_photoz_des<T>(T(x[1])));
      break;
    case Y3G_F_INT_ZO_ZT_DES:
      r = int_zo_zt_des<T>(T(x[0]), T(x[1]), T(x[2]));
      break;
    case Y3G_F_SIGMA_PHOTOZ: r = sigma_photoz_des<T>(T(x[0])); break;
    case Y3G_F_INT_LC_LT_DES: r = T(int_lc_lt_des(q, x[0], x[1], x[2])); break;
    case Y3G_F_MOR_DES: r = T(mor_des(q, x[0], x[1], x[2])); break;
    case Y3G_F_MOR_DES_LOG: r = T(mor_des_log(q, x[0], x[1], x[2])); break;
    case Y3G_F_EMG_DES_PDF: r = T(emg_des_pdf(q, x[0], x[1], x[2])); break;
    case Y3G_F_EMG_DES_CDF: r = T(emg_des_cdf(q, x[0], x[1], x[2])); break;
    case Y3G_F_SIG_MAX: r = T(sig_max(q, x[0], x[1], x[2])); break;
    case Y3G_F_LC_LT: r = T(lc_lt_tab(x[0], x[1], x[2])); break;
    case Y3G_F_SCI_AVERAGE: {
      int const n = int(q[0]);
      r = T(emg_des_curve(q + 1, q + 1 + n, n, x[0]));
      break;
    }
    case Y3G_F_SCI_RZ: {
      int const nr = int(q[0]), nz = int(q[1]);
      r = T(interp2d_clamp(q + 2, nr, q + 2 + nr, nz, q + 2 + nr + nz, x[0], x[1]));
      break;
    }
    case Y3G_F_NFW_DSIGMA_MIS:
      r = T(nfw_dsigma_mis_tab(F.nfwmis, F.nfw_nx, F.nfw_nm, x[0], x[1], x[2], q[0], q[1]));
      break;
    case Y3G_F_DV_DO_DZ: r = dv_do_dz<T>(P, tb.da, x[0]); break;
    case Y3G_F_HMF: r = hmf<T>(P, tb.hmf, x[0], x[1]); break;
    case Y3G_F_MOR: {
      T const s = T(P.p.mor_sigma_ln);
      T const lnlbar = T(P.p.mor_lnl0) + T(P.p.mor_slope) * T(x[1] - kLnPivotMor) +
                       T(P.p.mor_zslope) * M<T>::log((T(1) + T(x[2])) * T(1.0 / 1.35));
      T const dd = (M<T>::log(T(x[0])) - lnlbar) / s;
      r = M<T>::exp(T(-0.5) * dd * dd) * T(0.3989422804014327) / (s * T(x[0]));
      break;
    }
    case Y3G_F_LO_LT_PDF:
      r = lo_lt_pdf<T>(T(x[0]), T(x[1]), T(P.p.f_prj), T(P.p.prj_scale));
      break;
    case Y3G_F_LO_LT_BIN:
      r = lo_lt_bin<T>(T(x[0]), T(x[1]), T(x[2]), T(P.p.f_prj), T(P.p.prj_scale));
      break;
    case Y3G_F_NFW_DSIGMA: {
      NfwHalo<T> const H = nfw_halo<T>(x[1], T(x[2]), T(x[3]));
      r = nfw_dsigma<T>(T(x[0]), H);
      break;
    }
    case Y3G_F_DSIGMA_OFFSET: {
      NfwHalo<T> const H = nfw_halo<T>(x[3], T(x[4]), T(x[5]));
      r = dsigma_offset<T>(T(x[0]), T(x[1]), T(x[2]), H);
      break;
    }
    case Y3G_F_SCI: r = lerp_clamp<T>(tb.sci, P.nsci, x[0], P.inv_dzsci); break;
    case Y3G_F_INTEGRAND:
      if constexpr (sizeof(T) == 8) {
        if (is_des_kind(F.ig->kind)) {
          r = integrand_des(*F.ig, P, tb, F.des, x);
          break;
        }
      }
      r = integrand<T>(*F.ig, P, tb, x);
      break;
    case Y3G_F_INTEGRAND_PAIRED: {
      int it, iz, im;
      kind_axes<T>(F.ig->kind, it, iz, im);
      r = integrand_pair<T>(*F.ig, P, pair_factors<T>(*F.ig, P, tb, x[iz], x[im]), x);
      break;
    }
    case Y3G_F_INTEGRAND_NODE: {
      int it, iz, im;
      kind_axes<T>(F.ig->kind, it, iz, im);
      PairF<T> const pf = pair_factors<T>(*F.ig, P, tb, x[iz], x[im]);
      if constexpr (sizeof(T) == 4) {
        float yf[kMaxDim];
        for (int a = 0; a < kMaxDim; ++a) yf[a] = a < F.ig->ndim ? float(x[a]) : 0.0f;
        NodeConstF const kc = node_const_f(*F.ig, P, kNfwTab);
        PairC const pc = pair_c(pf);
        switch (F.ig->kind) {
          case Y3G_COUNT3D: r = integrand_node_f<Y3G_COUNT3D>(kc, pc, yf, 1.0f); break;
          case Y3G_COUNT5D: r = integrand_node_f<Y3G_COUNT5D>(kc, pc, yf, 1.0f); break;
          default: {
            float const prof =
              profile_f(kc, pf.irs, pf.amp, pf.cds, yf[4], float(1.0 - x[4]), yf[5]);
            r = integrand_node_f<Y3G_DSIGMA6D>(kc, pc, yf, prof);
            break;
          }
        }
      } else {
        r = integrand_pair<T>(*F.ig, P, pf, x);
      }
      break;
    }
  }
  F.out[i] = double(r);
}

// ------------------------------------------- fixed-GL counts with selection
// One workgroup per (point, bin), summing the n_lnm x n_z terms
// w_k zfac_q hmf(lnM_k, z_q) S_b(lnM_k, z_q) in FP64
// (n_operator_sel_gl_t.hh:88-118).  Everything that depends on one node
// coordinate only is computed once per workgroup into LDS: per mass node its
// weight, HMF cell / fraction / amplitude factor and selection-axis cell /
// fraction; per redshift node w_q dV/dz Omega (zfac) and its two cells.  A
// term is then two bilinear lookups (HMF table, selection plane: 8 loads,
// both L2-resident) and a few FMAs, with the arithmetic of hmf() and
// sel_function() kept in the same order as the direct per-node evaluation.
struct SelNodeM {
  double w, ht, hfac, st;  // GL weight; HMF lnM fraction, amplitude; sel fraction
  int hi, si;              // HMF lnM cell; sel lnM cell (-1: outside the axis)
};
struct SelNodeZ {
  double zf, hs, su;       // w_q dV/dz Omega; HMF z fraction; sel z fraction
  int hj, sj;              // HMF z cell; sel z cell (-1: outside the axis)
};
constexpr int kSelBlock = 256;
// The point's selection axes are copied to LDS (one round trip, before the
// node descriptors, whose bracket searches then read LDS) when the launch's
// longest axes fit this many doubles (the reference's default grids take
// 60); longer ones are searched in global memory.  The bin's plane stays in
// global memory: its lookups gather across lanes (a wave's lanes share the
// mass node and step through the redshift cells), which the L1 cache serves
// without the bank conflicts the same gathers meet in LDS.
constexpr int kSelAxesMax = 1024;

static size_t selgl_smem(int n_lnm, int n_z, int stage_doubles)
{
  return sizeof(SelNodeM) * (size_t)n_lnm + sizeof(SelNodeZ) * (size_t)n_z +
         sizeof(double) * (size_t)stage_doubles;
}

template <bool kStage>
__global__ __launch_bounds__(kSelBlock) void k_selgl(SelGLArgs A)
{
  extern __shared__ __align__(16) unsigned char sel_smem[];
  __shared__ double red[2 * (kSelBlock / 64)];
  SelNodeM* nm = reinterpret_cast<SelNodeM*>(sel_smem);
  SelNodeZ* nz = reinterpret_cast<SelNodeZ*>(nm + A.n_lnm);
  double* stage = reinterpret_cast<double*>(nz + A.n_z);
  int const pt = blockIdx.x / A.max_bins, b = blockIdx.x - pt * A.max_bins;
  SelPoint const sp = A.sel[pt];
  if (b >= sp.n_bins) {
    if (threadIdx.x == 0) A.out[blockIdx.x] = 0.0;
    return;
  }
  PointDev const& P = A.points[pt];
  const double* htab = A.hmf64 + (size_t)pt * A.hmf_stride;
  const double* da = A.da64 + (size_t)pt * A.da_stride;
  const double* xm = A.nodes;
  const double* wm = xm + A.n_lnm;
  const double* xz = wm + A.n_lnm;
  const double* wz = xz + A.n_z;
  int const npl = sp.n_z * sp.n_lnm;
  const double* lnm = A.data + sp.off_lnm;
  const double* zs = A.data + sp.off_z;
  const double* plane = A.data + sp.off_s + (size_t)b * npl;
  if constexpr (kStage) {
    // [lnM axis | z axis], adjacent in global memory too
    const double* ax = lnm;
    int const nax = sp.n_lnm + sp.n_z;
    for (int i = threadIdx.x; i < nax; i += kSelBlock) stage[i] = ax[i];
    lnm = stage;
    zs = lnm + sp.n_lnm;
    __syncthreads();
  }
  for (int k = threadIdx.x; k < A.n_lnm; k += kSelBlock) {
    double const lnM = xm[k];
    SelNodeM d;
    d.w = wm[k];
    double u = (lnM - P.lnm0) * P.inv_dlnm;  // hmf(): lnM axis
    u = u < 0.0 ? 0.0 : (u > double(P.nm - 1) ? double(P.nm - 1) : u);
    int i = int(u);
    if (i > P.nm - 2) i = P.nm - 2;
    d.hi = i;
    d.ht = u - double(i);
    d.hfac = P.p.hmf_s * (lnM * 0.4342944819 - 13.8124426028) + P.p.hmf_q;
    d.si = -1;                               // sel_function(): lnM axis
    d.st = 0.0;
    if (lnM >= lnm[0] && lnM <= lnm[sp.n_lnm - 1]) {
      int const c = bracket(lnm, sp.n_lnm, lnM);
      d.si = c;
      d.st = (lnM - lnm[c]) / (lnm[c + 1] - lnm[c]);
    }
    nm[k] = d;
  }
  for (int q = threadIdx.x; q < A.n_z; q += kSelBlock) {
    double const z = xz[q];
    SelNodeZ d;
    d.zf = wz[q] * dv_do_dz<double>(P, da, z) * omega_z_des(z);
    double v = (z - P.z0) * P.inv_dz;        // hmf(): z axis
    v = v < 0.0 ? 0.0 : (v > double(P.nz - 1) ? double(P.nz - 1) : v);
    int j = int(v);
    if (j > P.nz - 2) j = P.nz - 2;
    d.hj = j;
    d.hs = v - double(j);
    d.sj = -1;                               // sel_function(): z axis
    d.su = 0.0;
    if (z >= zs[0] && z <= zs[sp.n_z - 1]) {
      int const c = bracket(zs, sp.n_z, z);
      d.sj = c;
      d.su = (z - zs[c]) / (zs[c + 1] - zs[c]);
    }
    nz[q] = d;
  }
  __syncthreads();

  // term t = q * n_lnm + k, stepping by the block size without a division.
  // Mass fastest: a wave's lanes then read one HMF row and one plane row
  // (both lnM-fastest) at neighbouring cells, few cache lines per gather.
  double acc = 0.0;
  int const nt = A.n_lnm * A.n_z;
  int const dq = kSelBlock / A.n_lnm, dk = kSelBlock - dq * A.n_lnm;
  int q = threadIdx.x / A.n_lnm, k = threadIdx.x - q * A.n_lnm;
  for (int t = threadIdx.x; t < nt; t += kSelBlock) {
    SelNodeM const m = nm[k];
    SelNodeZ const z = nz[q];
    if (m.si >= 0 && z.sj >= 0) {            // S_b = 0 outside its axes
      const double* h0 = htab + z.hj * P.nm + m.hi;
      const double* h1 = h0 + P.nm;
      double const ha = h0[0] + m.ht * (h0[1] - h0[0]);
      double const hb = h1[0] + m.ht * (h1[1] - h1[0]);
      double const hm = (ha + z.hs * (hb - ha)) * m.hfac;
      const double* r0 = plane + z.sj * sp.n_lnm + m.si;
      const double* r1 = r0 + sp.n_lnm;
      double const t1 = m.st, u1 = z.su;
      double const sv = (1.0 - t1) * (1.0 - u1) * r0[0] + t1 * (1.0 - u1) * r0[1] +
                        (1.0 - t1) * u1 * r1[0] + t1 * u1 * r1[1];
      acc += m.w * z.zf * hm * sv;
    }
    q += dq;
    k += dk;
    if (k >= A.n_lnm) {
      k -= A.n_lnm;
      ++q;
    }
  }
  acc = block_sum<kSelBlock>(acc, red);
  if (threadIdx.x == 0) A.out[blockIdx.x] = acc;
}

hipError_t launch_selgl(const SelGLArgs& A, int n_points, hipStream_t s)
{
  bool const stage = A.max_axes <= kSelAxesMax;
  size_t const smem = selgl_smem(A.n_lnm, A.n_z, stage ? A.max_axes : 0);
  const void* fn = stage ? reinterpret_cast<const void*>(&k_selgl<true>)
                         : reinterpret_cast<const void*>(&k_selgl<false>);
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  if (e != hipSuccess) return e;
  if (stage)
    hipLaunchKernelGGL(k_selgl<true>, dim3(n_points * A.max_bins), dim3(kSelBlock), smem, s, A);
  else
    hipLaunchKernelGGL(k_selgl<false>, dim3(n_points * A.max_bins), dim3(kSelBlock), smem, s,
                       A);
  return hipGetLastError();
}

// ----------------------------------------------------------- launchers
template <typename T, int BLOCK, bool SM>
static hipError_t launch_cub_sm(const KernArgs& A, int slots, hipStream_t s)
{
  size_t const smem = sizeof(CubShared<T, BLOCK>);
  hipError_t e = hipFuncSetAttribute(
    reinterpret_cast<const void*>(&k_cubature<T, BLOCK, SM>),
    hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_cubature<T, BLOCK, SM>), dim3(slots), dim3(BLOCK), smem,
                     s, A);
  return hipGetLastError();
}
template <typename T, int BLOCK>
static hipError_t launch_cub(const KernArgs& A, int slots, hipStream_t s)
{
  if constexpr (sizeof(T) == 4)
    if (A.hmfc32) return launch_cub_sm<T, BLOCK, true>(A, slots, s);
  return launch_cub_sm<T, BLOCK, false>(A, slots, s);
}

template <typename T, int BLOCK>
static hipError_t launch_veg(const KernArgs& A, int slots, hipStream_t s)
{
  size_t const smem = sizeof(VegShared<T, BLOCK>);
  hipError_t e = hipFuncSetAttribute(
    reinterpret_cast<const void*>(&k_vegas<T, BLOCK>),
    hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_vegas<T, BLOCK>), dim3(slots), dim3(BLOCK), smem, s,
                     A);
  return hipGetLastError();
}

int cubature_blocks_per_cu(int precision, int block)
{
  int n = 0;
  size_t const smem = cubature_smem(precision, block);
  const void* f =
    precision == Y3G_FP64
      ? (block == 256 ? reinterpret_cast<const void*>(&k_cubature<double, 256>)
                      : reinterpret_cast<const void*>(&k_cubature<double, 512>))
      : (block == 256 ? reinterpret_cast<const void*>(&k_cubature<float, 256>)
                      : reinterpret_cast<const void*>(&k_cubature<float, 512>));
  if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem) !=
        hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, block, smem) != hipSuccess)
    return 0;
  return n;
}

size_t cubature_smem(int precision, int block)
{
  if (precision == Y3G_FP64)
    return block == 256 ? sizeof(CubShared<double, 256>)
                        : sizeof(CubShared<double, 512>);
  return block == 256 ? sizeof(CubShared<float, 256>)
                      : sizeof(CubShared<float, 512>);
}

hipError_t launch_cubature(const KernArgs& A, int precision, int block,
                           int slots, hipStream_t s)
{
  if (precision == Y3G_FP64)
    return block == 256 ? launch_cub<double, 256>(A, slots, s)
                        : launch_cub<double, 512>(A, slots, s);
  return block == 256 ? launch_cub<float, 256>(A, slots, s)
                      : launch_cub<float, 512>(A, slots, s);
}

template <typename T, bool SM>
static hipError_t launch_cub_stream_sm(const KernArgs& A, const StreamArgs& Z, int slots,
                                       hipStream_t s)
{
  size_t const smem = sizeof(CubShared<T, 512>);
  hipError_t e = hipFuncSetAttribute(
    reinterpret_cast<const void*>(&k_cubature_stream<T, 512, SM>),
    hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_cubature_stream<T, 512, SM>), dim3(slots), dim3(512), smem, s, A, Z);
  return hipGetLastError();
}
template <typename T>
static hipError_t launch_cub_stream(const KernArgs& A, const StreamArgs& Z, int slots,
                                    hipStream_t s)
{
  if constexpr (sizeof(T) == 4)
    if (A.hmfc32) return launch_cub_stream_sm<T, true>(A, Z, slots, s);
  return launch_cub_stream_sm<T, false>(A, Z, slots, s);
}

hipError_t launch_cubature_stream(const KernArgs& A, const StreamArgs& Z, int precision,
                                  int slots, hipStream_t s)
{
  return precision == Y3G_FP64 ? launch_cub_stream<double>(A, Z, slots, s)
                               : launch_cub_stream<float>(A, Z, slots, s);
}

int cubature_stream_blocks_per_cu(int precision)
{
  int n = 0;
  size_t const smem = cubature_smem(precision, 512);
  const void* f = precision == Y3G_FP64
                    ? reinterpret_cast<const void*>(&k_cubature_stream<double, 512>)
                    : reinterpret_cast<const void*>(&k_cubature_stream<float, 512>);
  if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem) !=
        hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, 512, smem) != hipSuccess)
    return 0;
  return n;
}

hipError_t launch_vegas(const KernArgs& A, int precision, int block, int slots,
                        hipStream_t s)
{
  if (precision == Y3G_FP64)
    return block == 256 ? launch_veg<double, 256>(A, slots, s)
                        : launch_veg<double, 512>(A, slots, s);
  return block == 256 ? launch_veg<float, 256>(A, slots, s)
                      : launch_veg<float, 512>(A, slots, s);
}

// Zero n 32-bit words (counters and queues of memory a caller may capture
// into a graph: a kernel, not a memset node).
__global__ void k_zero_words(unsigned* p, size_t n)
{
  size_t const i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

hipError_t launch_zero_words(void* p, size_t n, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  unsigned const nb = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_zero_words, dim3(nb), dim3(256), 0, s, static_cast<unsigned*>(p), n);
  return hipGetLastError();
}

// NFW_DSIGMA_MIS table, one 256-lane workgroup per entry (xmis_j, x_i):
// the lanes share the outer Gauss-Legendre nodes of G(x_i, xmis_j) (panels
// from lane 0 through LDS), each node a ring average; FP64 block sum, then
// T = ln(G / 2).  tab = {ln x [nx], ln xmis [nm], T [nm][nx]}, axes filled.
__global__ __launch_bounds__(256) void k_nfw_mis_table(double* tab, int nx, int nm)
{
  __shared__ double plo[kNfwMisMaxPanels], phi[kNfwMisMaxPanels];
  __shared__ double red[4];
  __shared__ int npan;
  int const e = blockIdx.x;
  int const i = e % nx, j = e / nx;
  double const x = exp(tab[i]), m = exp(tab[nx + j]);
  if (threadIdx.x == 0) npan = nfw_mis_panels(x, m, plo, phi);
  __syncthreads();
  int const nn = 8 * npan;
  double acc = 0.0;
  for (int q = threadIdx.x; q < nn; q += 256) {
    int const p = q >> 3, k = q & 7;
    double const c = 0.5 * (plo[p] + phi[p]), h = 0.5 * (phi[p] - plo[p]);
    double const u = c + h * gl8_node(k);
    acc += h * gl8_weight(k) * u * exp(-u) * nfw_ring(x, m * u);
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0)
    tab[nx + nm + (size_t)j * nx + i] = log(0.5 * (red[0] + red[1] + red[2] + red[3]));
}

hipError_t launch_nfw_mis_table(double* tab, int nx, int nm, hipStream_t s)
{
  hipLaunchKernelGGL(k_nfw_mis_table, dim3((unsigned)(nx * nm)), dim3(256), 0, s, tab, nx, nm);
  return hipGetLastError();
}

hipError_t launch_tables(const TableArgs& A, int n_points, hipStream_t s)
{
  hipLaunchKernelGGL(k_tables_build, dim3(n_points), dim3(256), 0, s, A);
  return hipGetLastError();
}

hipError_t launch_hmf_coefs(const TableArgs& A, int pt0, int n_points, hipStream_t s)
{
  hipLaunchKernelGGL(k_hmf_coefs, dim3(n_points), dim3(256), 0, s, A, pt0);
  return hipGetLastError();
}

hipError_t launch_functor(const FunctorArgs& F, int precision, hipStream_t s)
{
  unsigned const nb = (unsigned)((F.n + 255) / 256);
  if (precision == Y3G_FP64)
    hipLaunchKernelGGL(k_eval_functor<double>, dim3(nb), dim3(256), 0, s, F);
  else
    hipLaunchKernelGGL(k_eval_functor<float>, dim3(nb), dim3(256), 0, s, F);
  return hipGetLastError();
}

}  // namespace y3g
