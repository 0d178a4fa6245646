// Forward simulation kernels for gfx950: Philox4x32-10 counter-based RNG, the
// EV contraction, the agent partition by regime and the per-agent argmax and
// state transition. Declarations and reference citations: include/pylcm_amd.h.

#include <climits>
#include <mutex>
#include <string>
#include <type_traits>

#include "common.h"

namespace lcm {

constexpr int kSimBlock = 256;
constexpr int kSimChunk = 128;  // actions staged in LDS at a time
constexpr int kSimLanes = 4;    // lanes per agent; they split the agent's actions
constexpr int kSimAgents = kSimBlock / kSimLanes;  // agents per workgroup
constexpr int kPartAgents = 4;  // agents per thread of the partition kernel

// Loop over the agent features unrolled to the static maximum, so that phi[k]
// stays a register (with the runtime bound nf each phi[k] becomes a chain of
// selects).
#define LCM_FOR_FEAT(k) _Pragma("unroll") for (int k = 0; k < LCM_SIM_MAX_FEAT; ++k) if (k < nf)

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplications with Weyl key
// schedule.
// ---------------------------------------------------------------------------
struct U4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    const uint64_t p0 = uint64_t(M0) * c.x;
    const uint64_t p1 = uint64_t(M1) * c.z;
    const uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
    const uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
  }
  return c;
}

// 53-bit uniform in [0, 1) from two 32-bit words.
__device__ __forceinline__ double uniform53(uint32_t a, uint32_t b) {
  return (double(a >> 5) * 67108864.0 + double(b >> 6)) * (1.0 / 9007199254740992.0);
}

// Box-Muller standard normal from one Philox block (first of the pair).
__device__ __forceinline__ double normal_bm(const U4& r) {
  const double u1 = 1.0 - uniform53(r.x, r.y);  // (0, 1]
  const double u2 = uniform53(r.z, r.w);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

constexpr uint32_t kStreamRegime = 0;
constexpr uint32_t kStreamProcess = 1;
constexpr uint32_t kStreamDiscrete = 2;
constexpr uint32_t kStreamTaste = 3;  // counter word 3 = the discrete cell
constexpr double kEulerGamma = 0.5772156649015329;

__global__ void philox_kernel(uint32_t* out, int n, uint32_t c1, uint32_t c2, uint32_t c3,
                              uint32_t k0, uint32_t k1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const U4 r = philox4x32_10(U4{uint32_t(i), c1, c2, c3}, k0, k1);
  out[4 * i + 0] = r.x;
  out[4 * i + 1] = r.y;
  out[4 * i + 2] = r.z;
  out[4 * i + 3] = r.w;
}

// ---------------------------------------------------------------------------
// EV contraction: one workgroup per (descriptor, sd, d) row.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
ev_kernel(const LcmEvDesc* __restrict__ descs, int max_rows) {
  const int di = blockIdx.y;
  const LcmEvDesc E = descs[di];
  const int row = blockIdx.x;
  if (row >= E.n_sd * E.n_d) return;
  const LcmGsDesc* __restrict__ D = E.gs;
  const int n_d = E.n_d, g_n = E.g_n;
  const int sd = row / n_d, d = row % n_d;
  T* out = static_cast<T*>(E.EV) + int64_t(row) * g_n;
  int nonfinite = 0;
  for (int k = threadIdx.x; k < g_n; k += blockDim.x) {
    // accumulated in FP64 and rounded once to T, as the solve's EV rows
    double acc = 0.0;
    for (int r = 0; r < D->n_targets; ++r) {
      const int nj = D->n_sd_next[r];
      const T* __restrict__ Mr = static_cast<const T*>(D->M[r]) + (int64_t(sd) * n_d + d) * nj;
      const T* __restrict__ Vr = static_cast<const T*>(D->Vn[r]);
      for (int j = 0; j < nj; ++j) {
        const T m = Mr[j];
        if (m != T(0)) acc = __builtin_fma(double(m), double(Vr[int64_t(j) * g_n + k]), acc);
      }
    }
    const T v = T(acc);
    out[k] = v;
    nonfinite |= !is_finite(v);
  }
  const bool row_finite = __syncthreads_or(nonfinite) == 0;
  if (threadIdx.x == 0 && E.fin) E.fin[row] = row_finite ? 1 : 0;
  if (threadIdx.x == 0) {
    T cs = T(0);
    for (int s = 0; s < D->n_scalar; ++s) {
      const T m = static_cast<const T*>(D->ms[s])[int64_t(sd) * n_d + d];
      const T v = static_cast<const T*>(D->Vs[s])[0];
      cs += zero_safe(m, v);
    }
    static_cast<T*>(E.CS)[row] = cs;
  }
}

// ---------------------------------------------------------------------------
// Partition agents by regime (order within a list does not affect results).
// ---------------------------------------------------------------------------
// Counts per regime in LDS first, then one global atomic per (workgroup,
// regime) reserves the workgroup's range: contention on the global counters
// stays at one update per workgroup instead of one per agent. The same pass
// carries every agent's state into the next panel (agents not simulated
// this period keep it; simulated ones are overwritten by the simulation
// kernel) and marks the period's actions as not simulated (-1).
template <typename T>
__global__ void __launch_bounds__(256)
partition_kernel(LcmAgents cur, LcmAgents nxt, int32_t* __restrict__ action_out,
                 int n, int period, int n_regimes, int32_t* __restrict__ counts,
                 int32_t* __restrict__ next_counts, int32_t* __restrict__ order) {
  __shared__ int lcount[LCM_SIM_MAX_REGIMES];
  __shared__ int lbase[LCM_SIM_MAX_REGIMES];
  // kPartAgents agents per thread (strided, so every access stays coalesced):
  // fewer workgroups, so fewer global atomics on the per-regime counters
  int r[kPartAgents], local[kPartAgents];
  if (threadIdx.x < LCM_SIM_MAX_REGIMES) lcount[threadIdx.x] = 0;
  // the other half of the counters serves the next period; the kernels that
  // read it last (the previous period's) have finished
  if (blockIdx.x == 0 && threadIdx.x < n_regimes) next_counts[threadIdx.x] = 0;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPartAgents; ++k) {
    const int i = (blockIdx.x * kPartAgents + k) * blockDim.x + threadIdx.x;
    r[k] = -1;
    if (i < n) {
      int rr = cur.regime[i];
      nxt.regime[i] = rr;
      nxt.sdd[i] = cur.sdd[i];
      static_cast<T*>(nxt.w)[i] = static_cast<const T*>(cur.w)[i];
      static_cast<T*>(nxt.z)[i] = static_cast<const T*>(cur.z)[i];
      action_out[i] = -1;
      if (cur.start[i] > period || rr < 0 || rr >= n_regimes) rr = -1;
      r[k] = rr;
    }
    local[k] = r[k] >= 0 ? atomicAdd(&lcount[r[k]], 1) : 0;
  }
  __syncthreads();
  if (threadIdx.x < n_regimes && lcount[threadIdx.x] > 0)
    lbase[threadIdx.x] = atomicAdd(&counts[threadIdx.x], lcount[threadIdx.x]);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPartAgents; ++k) {
    const int i = (blockIdx.x * kPartAgents + k) * blockDim.x + threadIdx.x;
    if (r[k] >= 0) order[int64_t(r[k]) * n + lbase[r[k]] + local[k]] = i;
  }
}

__global__ void clear_kernel(int32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}


// ---------------------------------------------------------------------------
// Per-agent argmax over actions and next-state draw.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T sim_coordinate(int kind, int n, T start, T stop, const T* pts, T x) {
  if (n < 2) return T(0);
  if (kind == LCM_GRID_LIN) {
    const T step = (stop - start) / T(n - 1);
    return (x - start) / step;
  } else if (kind == LCM_GRID_LOG) {
    const T ls = log(start), step = (log(stop) - ls) / T(n - 1);
    const T cl = (log(x) - ls) / step;
    const T rlo = floor(cl);
    const T lower = exp(ls + step * rlo), upper = exp(ls + step * (rlo + T(1)));
    return rlo + (x - lower) / (upper - lower);
  }
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (pts[mid] <= x) lo = mid + 1; else hi = mid;
  }
  const int up = min(max(lo, 1), n - 1), dn = up - 1;
  return T(dn) + (x - pts[dn]) / (pts[up] - pts[dn]);
}

template <typename T, bool FINITE>
__device__ __forceinline__ T interp_at(const T* row, int n, T coord) {
  if (n < 2) return row[0];
  T fl = fmin(fmax(floor(coord), T(0)), T(n - 2));
  const int lo = int(fl);
  const T t = coord - fl;
  if constexpr (FINITE) return (T(1) - t) * row[lo] + t * row[lo + 1];
  return zero_safe(T(1) - t, row[lo]) + zero_safe(t, row[lo + 1]);
}

// Running first-argmax of one lane and its merge with another lane's: the
// larger value wins, equal values keep the smaller flat action index.
template <typename T>
struct SimBest {
  T best;
  int idx;
  bool nan;
};

template <typename T>
__device__ __forceinline__ SimBest<T> sim_merge(SimBest<T> a, SimBest<T> b) {
  a.nan |= b.nan;
  if (b.idx >= 0 && (a.idx < 0 || b.best > a.best || (b.best == a.best && b.idx < a.idx))) {
    a.best = b.best;
    a.idx = b.idx;
  }
  return a;
}

// Merge of the kSimLanes lanes of one agent (butterfly within the aligned group).
template <typename T>
__device__ __forceinline__ SimBest<T> sim_merge_lanes(SimBest<T> mb) {
#pragma unroll
  for (int off = kSimLanes / 2; off >= 1; off >>= 1) {
    SimBest<T> o;
    o.best = __shfl_xor(mb.best, off);
    o.idx = __shfl_xor(mb.idx, off);
    o.nan = __shfl_xor(int(mb.nan), off) != 0;
    mb = sim_merge(mb, o);
  }
  return mb;
}

__device__ __forceinline__ int min_lanes(int v) {
#pragma unroll
  for (int off = kSimLanes / 2; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off));
  return v;
}

template <typename T>
__device__ __forceinline__ T apply_op(int op, T x) {
  switch (op) {
    case 1: return exp(x);
    case 2: return log(x);
    case 3: return sqrt(x);
    case 4: return -x;
    default: return x;
  }
}

// Draw an index from categorical weights p[0..n) with uniform u (CDF walk).
template <typename T>
__device__ __forceinline__ int draw_categorical(const T* p, int n, double u) {
  double total = 0.0;
  for (int i = 0; i < n; ++i) total += double(p[i]);
  double acc = 0.0;
  int last = 0;
  for (int i = 0; i < n; ++i) {
    if (p[i] > T(0)) last = i;
    acc += double(p[i]);
    if (p[i] > T(0) && u * total < acc) return i;
  }
  return last;
}

// Waves per SIMD the float32 kernel is compiled for: 8 (64 VGPRs, 29 more
// spilled to scratch and reloaded in the outer loops) run the 100k-agent panel
// 4 % faster than the unconstrained 94-VGPR build's 5 waves, with the same
// bits; the float64 kernel keeps the default (8 waves would spill 112).
template <typename T>
constexpr int kSimWavesPerEu = sizeof(T) == 4 ? 8 : 1;

template <typename T>
__global__ void __launch_bounds__(kSimBlock) __attribute__((amdgpu_waves_per_eu(kSimWavesPerEu<T>)))
sim_kernel(const LcmSimDesc* __restrict__ descs, int n_regimes, LcmAgents cur, LcmAgents nxt,
           LcmSimOut out, const int32_t* __restrict__ counts, const int32_t* __restrict__ order,
           int n_agents, int period, uint32_t k0, uint32_t k1) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int r = blockIdx.y;
  const LcmSimDesc* __restrict__ D = descs + r;
  if (D->U0 == nullptr) return;
  const int cnt = counts[r];
  // kSimLanes consecutive lanes serve one agent and split its actions
  const int lane = threadIdx.x % kSimLanes;
  const int slot = blockIdx.x * kSimAgents + threadIdx.x / kSimLanes;
  if (int(blockIdx.x) * kSimAgents >= cnt) return;  // whole block idle
  const bool valid = slot < cnt;
  const int agent = valid ? order[int64_t(r) * n_agents + slot] : 0;

  const int n_sdd = D->n_sdd, n_d = D->n_d, n_c = D->n_c, nf = D->n_feat;
  const bool interp = D->interp != 0;
  const int g_n = interp ? D->g_n : 0;
  const int n_nodes = D->n_nodes;
  const int sdd = valid ? cur.sdd[agent] : 0;
  const T w = valid ? static_cast<const T*>(cur.w)[agent] : T(0);
  const T z = valid ? static_cast<const T*>(cur.z)[agent] : T(0);

  // per-agent features
  T phi[LCM_SIM_MAX_FEAT];
#pragma unroll
  for (int k = 0; k < LCM_SIM_MAX_FEAT; ++k) phi[k] = T(0);
  for (int k = 0; k < nf; ++k) {
    const int op = D->feat_op[k];
    T v;
    if (op == 0) {
      v = z;
    } else {
      const T* fc = static_cast<const T*>(D->fcoef[k]) + int64_t(sdd) * LCM_SIM_MAX_FEAT;
      v = static_cast<const T*>(D->fconst[k])[sdd];
#pragma unroll
      for (int i = 0; i < LCM_SIM_MAX_FEAT; ++i) v += (i < k ? fc[i] * phi[i] : T(0));
      v = apply_op(op, v);
    }
#pragma unroll
    for (int i = 0; i < LCM_SIM_MAX_FEAT; ++i) if (i == k) phi[i] = v;
  }

  // process coordinate of z on the nodes: rows row_lo, row_lo + p_stride, weight t
  const T* nodes = static_cast<const T*>(D->nodes);
  int row_lo = D->sd_base[sdd];  // flat discrete state of process node 0
  T tz = T(0);
  if (interp && n_nodes > 1) {
    const T cz = sim_coordinate<T>(LCM_GRID_IRREG, n_nodes, T(0), T(0), nodes, z);
    const T fl = fmin(fmax(floor(cz), T(0)), T(n_nodes - 2));
    tz = cz - fl;
    row_lo += int(fl) * D->p_stride;
  }
  const int row_hi = row_lo + (n_nodes > 1 ? D->p_stride : 0);
  const int n_rows = interp ? D->n_rows : 0;

  // LDS: EV rows of the current d [n_rows][g_n], grid points, action chunk tables
  T* ev = reinterpret_cast<T*>(smem);
  // EV rows in LDS, or read in place from global memory when they do not fit
  const bool ev_in_lds = D->ev_global == 0;
  T* gp = ev + (ev_in_lds && n_rows * g_n > 0 ? n_rows * g_n : 1);
  // staged tables per action: U0 A B0 lo hi LoB HiB Uf.. Bf.. LoF.. HiF..
  const int ntab = 7 + 4 * nf;
  const bool lo_feat = D->LoB != nullptr, hi_feat = D->HiB != nullptr;
  T* tab = gp + (g_n > 0 ? g_n : 1);
  if (interp && D->g_kind == LCM_GRID_IRREG) {
    for (int k = threadIdx.x; k < g_n; k += kSimBlock) gp[k] = static_cast<const T*>(D->g_pts)[k];
  }
  const T g_start = T(D->g_start), g_stop = T(D->g_stop);
  const T beta = T(D->beta);
  const bool lin = interp && D->g_kind == LCM_GRID_LIN && g_n >= 2;
  const T g_step = lin ? (g_stop - g_start) / T(g_n - 1) : T(1);
  const T g_inv = T(1) / g_step;

  T best = -INFINITY;
  int bidx = -1;
  bool nanf = false;
  // taste shocks: per discrete cell the masked max Qc and its first maximiser
  // (the first feasible NaN when Qc is NaN, index 0 of the cell when nothing
  // is feasible); the cell is the first argmax of Qc + noise
  const double taste = D->taste_scale;
  int nan_idx = INT_MAX;  // first feasible NaN of this lane (taste shocks)
  double t_noisy = 0.0;
  T t_qc = T(0);
  int t_idx = -1;
  for (int d = 0; d < n_d; ++d) {
    __syncthreads();
    int nonfinite = 0;
    if (interp && !ev_in_lds && D->ev_fin) {
      // rows read in place: their finiteness was recorded when they were written
      for (int rr = 0; rr < n_rows; ++rr) nonfinite |= D->ev_fin[rr * n_d + d] == 0;
    } else if (interp) {
      const T* EVg = static_cast<const T*>(D->EV);
      for (int rr = 0; rr < n_rows; ++rr) {
        const T* src = EVg + (int64_t(rr) * n_d + d) * g_n;
        for (int k = threadIdx.x; k < g_n; k += kSimBlock) {
          const T v = src[k];
          if (ev_in_lds) ev[rr * g_n + k] = v;
          nonfinite |= !is_finite(v);
        }
      }
    }
    const bool finite = __syncthreads_or(nonfinite) == 0;
    T cs = T(0);
    if (!D->terminal) {
      const T* CS = static_cast<const T*>(D->CS);
      cs = CS ? CS[int64_t(row_lo) * n_d + d] : T(0);
    }
    for (int c0 = 0; c0 < n_c; c0 += kSimChunk) {
      const int cn = min(kSimChunk, n_c - c0);
      __syncthreads();
      // stage tables for all sdd: tab[((s*ntab + j)*kSimChunk) + i]
      for (int s = 0; s < n_sdd; ++s)
      for (int i = threadIdx.x; i < cn; i += kSimBlock) {
        const int64_t e = (int64_t(s) * n_d + d) * n_c + c0 + i;
        T* base = tab + int64_t(s) * ntab * kSimChunk + i;
        base[0 * kSimChunk] = static_cast<const T*>(D->U0)[e];
        base[1 * kSimChunk] = interp ? static_cast<const T*>(D->A)[e] : T(0);
        base[2 * kSimChunk] = interp ? static_cast<const T*>(D->B0)[e] : T(0);
        base[3 * kSimChunk] = static_cast<const T*>(D->lo)[e];
        base[4 * kSimChunk] = static_cast<const T*>(D->hi)[e];
        base[5 * kSimChunk] = lo_feat ? static_cast<const T*>(D->LoB)[e] : T(-INFINITY);
        base[6 * kSimChunk] = hi_feat ? static_cast<const T*>(D->HiB)[e] : T(INFINITY);
        for (int k = 0; k < nf; ++k) {
          base[(7 + k) * kSimChunk] = D->Uf[k] ? static_cast<const T*>(D->Uf[k])[e] : T(0);
          base[(7 + nf + k) * kSimChunk] = (interp && D->Bf[k]) ? static_cast<const T*>(D->Bf[k])[e] : T(0);
          base[(7 + 2 * nf + k) * kSimChunk] = (lo_feat && D->LoF[k]) ? static_cast<const T*>(D->LoF[k])[e] : T(0);
          base[(7 + 3 * nf + k) * kSimChunk] = (hi_feat && D->HiF[k]) ? static_cast<const T*>(D->HiF[k])[e] : T(0);
        }
      }
      __syncthreads();
      // the action loops, instantiated for EV rows in LDS and in global memory
      // (one address space per instantiation keeps the gathers ds_read / global loads)
      auto actions = [&](auto lds_tag) {
        constexpr bool IN_LDS = decltype(lds_tag)::value;
        const T* evb;
        if constexpr (IN_LDS) evb = ev;
        else evb = static_cast<const T*>(D->EV) + int64_t(d) * g_n;
        const int64_t ev_rs = IN_LDS ? int64_t(g_n) : int64_t(n_d) * g_n;
        const T* mine = tab + int64_t(sdd) * ntab * kSimChunk;
        int i0 = lane;
        if (interp && lin && finite && !lo_feat && !hi_feat && !D->terminal && nf <= 1) {
          // The common case (linear next grid, finite EV rows, interval bounds
          // without features, at most one feature) four actions at a time, so
          // that their LDS reads are in flight together; the same arithmetic and
          // update order as the loop below, which takes the remaining actions.
          const T* rl = evb + int64_t(row_lo) * ev_rs;
          const T* rh = evb + int64_t(row_hi) * ev_rs;
          const bool two = n_nodes > 1;
          const T top = T(g_n - 2);
          const T phi0 = phi[0];
          auto quad = [&](auto nf_tag) {
            constexpr int NF = decltype(nf_tag)::value;
            auto q_at = [&](int i) -> T {
              T u = mine[i];
              if constexpr (NF == 1) u += mine[7 * kSimChunk + i] * phi0;
              T x = mine[1 * kSimChunk + i] * w + mine[2 * kSimChunk + i];
              if constexpr (NF == 1) x += mine[8 * kSimChunk + i] * phi0;
              T coord;
              if constexpr (std::is_same<T, double>::value) coord = (x - g_start) / g_step;
              else coord = (x - g_start) * g_inv;
              const T fl = fmin(fmax(floor(coord), T(0)), top);
              const int li = int(fl);
              const T t = coord - fl;
              T v = (T(1) - tz) * ((T(1) - t) * rl[li] + t * rl[li + 1]);
              if (two) v += tz * ((T(1) - t) * rh[li] + t * rh[li + 1]);
              return u + beta * (cs + v);
            };
            auto take = [&](T q, bool feas, int i) {
              if (feas) {
                const bool isn = q != q;
                nanf |= isn;
                if (isn && nan_idx == INT_MAX) nan_idx = d * n_c + c0 + i;
                if (q == q && (bidx < 0 || q > best)) {
                  best = q;
                  bidx = d * n_c + c0 + i;
                }
              }
            };
            for (; i0 + 3 * kSimLanes < cn; i0 += 4 * kSimLanes) {
              bool f[4];
  #pragma unroll
              for (int k = 0; k < 4; ++k) {
                const int i = i0 + k * kSimLanes;
                f[k] = valid && (w >= mine[3 * kSimChunk + i]) && (w <= mine[4 * kSimChunk + i]);
              }
              if (__ballot(f[0] | f[1] | f[2] | f[3]) == 0) continue;
              T q[4];
  #pragma unroll
              for (int k = 0; k < 4; ++k) q[k] = q_at(i0 + k * kSimLanes);
  #pragma unroll
              for (int k = 0; k < 4; ++k) take(q[k], f[k], i0 + k * kSimLanes);
            }
          };
          if (nf == 1) quad(std::integral_constant<int, 1>{});
          else quad(std::integral_constant<int, 0>{});
        }
        for (int i = i0; i < cn; i += kSimLanes) {
          T lo = mine[3 * kSimChunk + i], hi = mine[4 * kSimChunk + i];
          if (lo_feat) {
            T l = mine[5 * kSimChunk + i];
            LCM_FOR_FEAT(k) l += mine[(7 + 2 * nf + k) * kSimChunk + i] * phi[k];
            if (D->lo_strict) l = nextafter(l, T(INFINITY));
            lo = fmax(lo, l);
          }
          if (hi_feat) {
            T h = mine[6 * kSimChunk + i];
            LCM_FOR_FEAT(k) h += mine[(7 + 3 * nf + k) * kSimChunk + i] * phi[k];
            if (D->hi_strict) h = nextafter(h, T(-INFINITY));
            hi = fmin(hi, h);
          }
          const bool feas = valid && (w >= lo) && (w <= hi);
          if (__ballot(feas) == 0) continue;
          T u = mine[i];
          LCM_FOR_FEAT(k) u += mine[(7 + k) * kSimChunk + i] * phi[k];
          T q;
          if (D->terminal) {
            q = u;
          } else if (interp) {
            T x = mine[1 * kSimChunk + i] * w + mine[2 * kSimChunk + i];
            LCM_FOR_FEAT(k) x += mine[(7 + nf + k) * kSimChunk + i] * phi[k];
            T coord;
            if (lin) {
              // linear grid: the same coordinate as the grid-search kernel
              if constexpr (std::is_same<T, double>::value) coord = (x - g_start) / g_step;
              else coord = (x - g_start) * g_inv;
            } else {
              coord = sim_coordinate<T>(D->g_kind, g_n, g_start, g_stop, gp, x);
            }
            T v;
            if (finite) {
              v = (T(1) - tz) * interp_at<T, true>(evb + int64_t(row_lo) * ev_rs, g_n, coord);
              if (n_nodes > 1) v += tz * interp_at<T, true>(evb + int64_t(row_hi) * ev_rs, g_n, coord);
            } else {
              v = zero_safe(T(1) - tz, interp_at<T, false>(evb + int64_t(row_lo) * ev_rs, g_n, coord));
              if (n_nodes > 1) v += zero_safe(tz, interp_at<T, false>(evb + int64_t(row_hi) * ev_rs, g_n, coord));
            }
            q = u + beta * (cs + v);
          } else {
            q = u + beta * cs;
          }
          if (feas) {
            const bool isn = q != q;
            nanf |= isn;
            if (isn && nan_idx == INT_MAX) nan_idx = d * n_c + c0 + i;
            if (q == q && (bidx < 0 || q > best)) {
              best = q;
              bidx = d * n_c + c0 + i;
            }
          }
        }
      };
      if (ev_in_lds) actions(std::true_type{});
      else actions(std::false_type{});
    }
    if (taste > 0.0) {
      const SimBest<T> mb = sim_merge_lanes(SimBest<T>{best, bidx, nanf});
      const int first_nan = min_lanes(nan_idx);
      T qc;
      int ci;
      if (mb.nan) {
        qc = T(NAN);
        ci = first_nan;
      } else if (mb.idx >= 0) {
        qc = mb.best;
        ci = mb.idx;
      } else {
        qc = T(-INFINITY);
        ci = d * n_c;
      }
      const U4 rt = philox4x32_10(U4{uint32_t(agent), uint32_t(period), kStreamTaste, uint32_t(d)}, k0, k1);
      const double u = uniform53(rt.x, rt.y) + 0x1p-54;  // in (0, 1)
      const double noisy = double(qc) + taste * (-log(-log(u)) - kEulerGamma);
      // first argmax; a NaN counts as the maximum (jnp.argmax)
      if (t_idx < 0 || (t_noisy == t_noisy && (noisy != noisy || noisy > t_noisy))) {
        t_noisy = noisy;
        t_qc = qc;
        t_idx = ci;
      }
      best = T(-INFINITY);
      bidx = -1;
      nanf = false;
      nan_idx = INT_MAX;
    }
  }
  if (taste > 0.0) {
    if (!valid || lane != 0) return;
  } else {
    // merge the agent's lanes (butterfly within the aligned group of kSimLanes)
    const SimBest<T> mb = sim_merge_lanes(SimBest<T>{best, bidx, nanf});
    best = mb.best;
    bidx = mb.idx;
    nanf = mb.nan;
    if (!valid || lane != 0) return;
  }

  // record this period
  const int64_t o = int64_t(period) * n_agents + agent;
  out.regime[o] = r;
  if (taste > 0.0) {
    // the drawn cell's maximiser and its value without the noise
    bidx = t_idx;
    out.action[o] = t_idx;
    static_cast<T*>(out.value)[o] = t_qc;
  } else {
    out.action[o] = (bidx >= 0 && !nanf) ? bidx : 0;
    static_cast<T*>(out.value)[o] = nanf ? T(NAN) : (bidx >= 0 ? best : T(-INFINITY));
  }
  static_cast<T*>(out.w)[o] = w;
  static_cast<T*>(out.z)[o] = z;
  out.sdd[o] = sdd;

  // next state; a terminal regime is absorbing with frozen state
  if (D->terminal) {
    nxt.regime[agent] = r;
    return;
  }
  const int a = bidx >= 0 ? bidx : 0;
  const int d = a / n_c, c = a % n_c;
  const int64_t e = (int64_t(sdd) * n_d + d) * n_c + c;
  T x = static_cast<const T*>(D->A)[e] * w + static_cast<const T*>(D->B0)[e];
  LCM_FOR_FEAT(k) x += (D->Bf[k] ? static_cast<const T*>(D->Bf[k])[e] : T(0)) * phi[k];

  const U4 rr = philox4x32_10(U4{uint32_t(agent), uint32_t(period), kStreamRegime, 0u}, k0, k1);
  const T* pr = static_cast<const T*>(D->preg) + (int64_t(sdd) * n_d + d) * D->n_regimes;
  const int nr = draw_categorical(pr, D->n_regimes, uniform53(rr.x, rr.y));
  int nsdd = 0;
  if (nr < LCM_SIM_MAX_REGIMES && D->Wd[nr] != nullptr) {
    const U4 rd = philox4x32_10(U4{uint32_t(agent), uint32_t(period), kStreamDiscrete, 0u}, k0, k1);
    const int nn = D->n_sdd_next[nr];
    const T* wd = static_cast<const T*>(D->Wd[nr]) + (int64_t(sdd) * n_d + d) * nn;
    nsdd = draw_categorical(wd, nn, uniform53(rd.x, rd.y));
  }
  T zn = z;
  if (n_nodes > 0) {
    const U4 rp = philox4x32_10(U4{uint32_t(agent), uint32_t(period), kStreamProcess, 0u}, k0, k1);
    const double nrm = normal_bm(rp);
    const double u = uniform53(rp.z, rp.w);
    const double pa = D->pa, pb = D->pb, pc = D->pc;
    double v;
    switch (D->proc_kind) {
      case 0: v = pa + pb * double(z) + pc * nrm; break;  // AR(1)
      case 1: v = pa + pc * nrm; break;                   // IID normal
      case 2: v = exp(pa + pc * nrm); break;              // IID lognormal
      default: v = pa + (pb - pa) * u; break;             // IID uniform
    }
    zn = T(v);
  }
  static_cast<T*>(nxt.w)[agent] = x;
  static_cast<T*>(nxt.z)[agent] = zn;
  nxt.sdd[agent] = nsdd;
  nxt.regime[agent] = nr;
}

}  // namespace lcm

using namespace lcm;

extern "C" int lcm_abi_sizes_sim(size_t* ev_desc, size_t* sim_desc, size_t* agents, size_t* sim_out) {
  if (ev_desc) *ev_desc = sizeof(LcmEvDesc);
  if (sim_desc) *sim_desc = sizeof(LcmSimDesc);
  if (agents) *agents = sizeof(LcmAgents);
  if (sim_out) *sim_out = sizeof(LcmSimOut);
  return LCM_OK;
}

extern "C" int lcm_philox(lcm_ctx* ctx, uint32_t* d_out, int n, uint32_t c1, uint32_t c2, uint32_t c3,
                          uint32_t k0, uint32_t k1) {
  if (!ctx || !d_out || n <= 0) return fail(LCM_E_ARG, "lcm_philox: bad argument");
  hipLaunchKernelGGL(philox_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, d_out, n, c1, c2,
                     c3, k0, k1);
  LCM_HIP(hipGetLastError());
  return LCM_OK;
}

extern "C" int lcm_ev_launch(lcm_ctx* ctx, int dtype, const LcmEvDesc* d_descs, int n_desc, int max_rows) {
  if (!ctx || !d_descs || n_desc <= 0 || max_rows <= 0) return fail(LCM_E_ARG, "lcm_ev_launch: bad argument");
  dim3 grid(max_rows, n_desc);
  if (dtype == LCM_DTYPE_F32)
    hipLaunchKernelGGL(ev_kernel<float>, grid, dim3(256), 0, ctx->stream, d_descs, max_rows);
  else if (dtype == LCM_DTYPE_F64)
    hipLaunchKernelGGL(ev_kernel<double>, grid, dim3(256), 0, ctx->stream, d_descs, max_rows);
  else
    return fail(LCM_E_ARG, "lcm_ev_launch: unknown dtype");
  LCM_HIP(hipGetLastError());
  return LCM_OK;
}

extern "C" size_t lcm_sim_lds_bytes(int dtype, int max_g, int max_ev_rows, int max_act) {
  const size_t sz = dtype == LCM_DTYPE_F64 ? 8 : 4;
  const size_t ev = size_t(max_ev_rows > 0 ? max_ev_rows : 1) * size_t(max_g > 0 ? max_g : 1);
  return (ev + size_t(max_g > 0 ? max_g : 1) + size_t(max_act) * kSimChunk) * sz + 64;
}

namespace lcm {

// One period of the simulation with the partition scratch at d_counts (two
// halves of n_regimes counters) and d_order (n_regimes * n_agents indices).
static int sim_period(lcm_ctx* ctx, int dtype, const LcmSimDesc* d_sim_descs, int n_regimes,
                      const LcmAgents* cur, const LcmAgents* next, const LcmSimOut* out, int n_agents,
                      int period, uint64_t seed, size_t lds_bytes, int32_t* d_counts, int32_t* d_order) {
  if (!ctx || !d_sim_descs || !cur || !next || !out || !d_counts || !d_order)
    return fail(LCM_E_ARG, "lcm_sim_period: null argument");
  if (n_agents <= 0 || n_regimes <= 0 || n_regimes > LCM_SIM_MAX_REGIMES)
    return fail(LCM_E_ARG, "lcm_sim_period: bad sizes");
  if (lds_bytes > 160 * 1024) return fail(LCM_E_ARG, "lcm_sim_period: LDS request above 160 KiB");
  const int nb = (n_agents + 256 * kPartAgents - 1) / (256 * kPartAgents);
  const int nb_sim = (n_agents + kSimAgents - 1) / kSimAgents;
  // counters: half period % 2 for this period; this period's partition clears
  // the other half for the next one, so only the first period clears both
  if (period == 0) hipLaunchKernelGGL(clear_kernel, dim3(1), dim3(64), 0, ctx->stream, d_counts, 2 * n_regimes);
  int32_t* counts = d_counts + (period & 1) * n_regimes;
  int32_t* next_counts = d_counts + ((period + 1) & 1) * n_regimes;
  const size_t sz = dtype == LCM_DTYPE_F64 ? 8 : 4;
  int32_t* action_out = out->action + int64_t(period) * n_agents;
  if (sz == 8)
    hipLaunchKernelGGL(partition_kernel<double>, dim3(nb), dim3(256), 0, ctx->stream, *cur, *next, action_out,
                       n_agents, period, n_regimes, counts, next_counts, d_order);
  else
    hipLaunchKernelGGL(partition_kernel<float>, dim3(nb), dim3(256), 0, ctx->stream, *cur, *next, action_out,
                       n_agents, period, n_regimes, counts, next_counts, d_order);
  dim3 grid(nb_sim, n_regimes);
  const uint32_t k0 = uint32_t(seed), k1 = uint32_t(seed >> 32);
  if (dtype == LCM_DTYPE_F32) {
    if (lds_bytes > 64 * 1024)
      LCM_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(sim_kernel<float>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_bytes)));
    hipLaunchKernelGGL(sim_kernel<float>, grid, dim3(kSimBlock), lds_bytes, ctx->stream, d_sim_descs,
                       n_regimes, *cur, *next, *out, counts, d_order, n_agents, period, k0, k1);
  } else if (dtype == LCM_DTYPE_F64) {
    if (lds_bytes > 64 * 1024)
      LCM_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(sim_kernel<double>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_bytes)));
    hipLaunchKernelGGL(sim_kernel<double>, grid, dim3(kSimBlock), lds_bytes, ctx->stream, d_sim_descs,
                       n_regimes, *cur, *next, *out, counts, d_order, n_agents, period, k0, k1);
  } else {
    return fail(LCM_E_ARG, "lcm_sim_period: unknown dtype");
  }
  LCM_HIP(hipGetLastError());
  return LCM_OK;
}

// Partition scratch of lcm_sim_period_noalloc, one copy per device for the
// life of the process: kNoallocCountsLen counters (two halves of at most
// LCM_SIM_MAX_REGIMES, padded so the order array starts 256-byte aligned),
// then LCM_SIM_NOALLOC_MAX_ORDER agent indices (16 MiB).
constexpr int kNoallocCountsLen = 64;
static __device__ __align__(256) int32_t scratch[kNoallocCountsLen + LCM_SIM_NOALLOC_MAX_ORDER];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static int32_t* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on ctx's device, looked up with
// hipGetSymbolAddress on the first call for that device (in host code the
// array's name is not a device address).
static int scratch_of(lcm_ctx* ctx, int32_t** out) {
  const int dev = ctx->device;
  if (dev < 0 || dev >= kMaxDevices) return fail(LCM_E_ARG, "lcm_sim_period_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int prev = 0;
    LCM_HIP(hipGetDevice(&prev));
    if (prev != dev) LCM_HIP(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (prev != dev) LCM_HIP(hipSetDevice(prev));
    if (e != hipSuccess) return check_hip(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(LCM_E_LAUNCH, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[dev] = static_cast<int32_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return LCM_OK;
}

}  // namespace lcm

extern "C" int lcm_sim_period_ex(lcm_ctx* ctx, int dtype, const LcmSimDesc* d_sim_descs, int n_regimes,
                                 const LcmAgents* cur, const LcmAgents* next, const LcmSimOut* out,
                                 int n_agents, int period, uint64_t seed, size_t lds_bytes,
                                 int32_t* d_counts, int32_t* d_order) {
  return sim_period(ctx, dtype, d_sim_descs, n_regimes, cur, next, out, n_agents, period, seed, lds_bytes,
                    d_counts, d_order);
}

extern "C" int lcm_sim_period_noalloc(lcm_ctx* ctx, int dtype, const LcmSimDesc* d_sim_descs, int n_regimes,
                                      const LcmAgents* cur, const LcmAgents* next, const LcmSimOut* out,
                                      int n_agents, int period, uint64_t seed, size_t lds_bytes) {
  if (!ctx) return fail(LCM_E_ARG, "lcm_sim_period_noalloc: null ctx");
  if (n_agents <= 0 || n_regimes <= 0 || n_regimes > LCM_SIM_MAX_REGIMES)
    return fail(LCM_E_ARG, "lcm_sim_period_noalloc: bad sizes");
  if (int64_t(n_regimes) * int64_t(n_agents) > int64_t(LCM_SIM_NOALLOC_MAX_ORDER))
    return fail(LCM_E_ARG, "lcm_sim_period_noalloc: n_regimes * n_agents = " +
                               std::to_string(int64_t(n_regimes) * int64_t(n_agents)) + " exceeds the " +
                               std::to_string(LCM_SIM_NOALLOC_MAX_ORDER) +
                               " entries of the static scratch; use lcm_sim_period_ex with a caller buffer");
  int32_t* base = nullptr;
  const int rc = scratch_of(ctx, &base);
  if (rc != LCM_OK) return rc;
  return sim_period(ctx, dtype, d_sim_descs, n_regimes, cur, next, out, n_agents, period, seed, lds_bytes, base,
                    base + kNoallocCountsLen);
}
