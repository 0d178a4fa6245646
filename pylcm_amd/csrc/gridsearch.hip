// Fused grid-search period kernel for gfx950.
//
// One launch evaluates a whole backward step for any number of (problem,
// regime) descriptors: for every state it contracts the next-period value rows
// into EV, interpolates EV at the next continuous state of every action, adds
// utility, masks infeasible actions and reduces to the max and the first
// argmax. It stands in for the compiled core `max_Q_over_a(Q_and_F)`
// (src/_lcm/regime_building/max_Q_over_a.py:30-191,
// src/_lcm/regime_building/Q_and_F.py:58-170), the interpolation
// `map_coordinates` (src/_lcm/regime_building/ndimage.py:28-109) on the grid
// coordinates of src/_lcm/grids/coordinates.py:30-167, the expectation
// (Q_and_F.py:1145-1209), `argmax_and_max` (src/_lcm/regime_building/argmax.py:6-69)
// and, with taste shocks, `logsum_and_softmax` (src/_lcm/logsum.py:32-64).
//
// Work decomposition: a workgroup of 256 threads owns TILE consecutive states
// of one (descriptor, discrete state). Lanes map to states; the 256/TILE waves
// that share a state interleave over the actions and are merged through LDS
// at the end (ties keep the smaller flat index, so the result is the first
// argmax whatever the interleave). Per discrete action the block stages
//   * the contracted EV row (next-period V rows weighted by M), and
//   * chunks of CH actions' operands {A, B, U} and constraint rows {Fa, Fb}
// in LDS. The action loop index is wave-uniform, so every per-action operand is
// an LDS broadcast read; the only per-lane LDS traffic is the two-corner gather
// from the EV row (one ds_read2). Blocks are remapped so that consecutive tasks
// (same problem, same V rows) land on the same XCD and share its L2.

#include <type_traits>

#include "common.h"

namespace lcm {

constexpr int kBlock = 256;
constexpr int kXcds = 8;
constexpr int kActChunk = 256;  // actions staged in LDS at a time

template <typename T>
struct Grid1D {
  int kind;
  int n;
  T start, step, inv_step, log_start, log_step;
  const T* pts;  // LDS copy for irregular grids
};

template <typename T>
__device__ __forceinline__ T coordinate(const Grid1D<T>& g, T x) {
  if (g.kind == LCM_GRID_LIN) {
    // get_linspace_coordinate (coordinates.py:30-44)
    if constexpr (std::is_same<T, double>::value) {
      return (x - g.start) / g.step;
    } else {
      return (x - g.start) * g.inv_step;
    }
  } else if (g.kind == LCM_GRID_LOG) {
    // get_logspace_coordinate (coordinates.py:73-114)
    const T cl = (log(x) - g.log_start) / g.log_step;
    const T rlo = floor(cl);
    const T lower = exp(g.log_start + g.log_step * rlo);
    const T upper = exp(g.log_start + g.log_step * (rlo + T(1)));
    return rlo + (x - lower) / (upper - lower);
  } else {
    // get_irreg_coordinate (coordinates.py:117-167): searchsorted(side="right")
    if (g.n < 2) return T(0);
    int lo = 0, hi = g.n;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (g.pts[mid] <= x) lo = mid + 1; else hi = mid;
    }
    const int up = min(max(lo, 1), g.n - 1);
    const int dn = up - 1;
    return T(dn) + (x - g.pts[dn]) / (g.pts[up] - g.pts[dn]);
  }
}

// Linear interpolation with linear extrapolation: lower index clipped to
// [0, n-2] (ndimage.py:102-109); a zero corner weight drops its term.
template <typename T>
__device__ __forceinline__ T interp_row(const T* row, int n, T coord, bool finite_row) {
  if (n < 2) return row[0];
  T fl = floor(coord);
  fl = fmin(fmax(fl, T(0)), T(n - 2));
  const int lo = static_cast<int>(fl);
  const T t = coord - fl;
  const T w0 = T(1) - t;
  const T v0 = row[lo];
  const T v1 = row[lo + 1];
  if (finite_row) return w0 * v0 + t * v1;
  return zero_safe(w0, v0) + zero_safe(t, v1);
}

struct Best {
  int idx;
  bool seen;
  bool nan;
};

// Merge candidate (vk, k) into (vb, b): a feasible NaN poisons the max; among
// feasible values the larger wins and ties keep the smaller flat index.
template <typename T>
__device__ __forceinline__ void merge(T& vb, Best& b, T vk, const Best& k) {
  b.nan |= k.nan;
  if (k.seen && (!b.seen || vk > vb || (vk == vb && k.idx < b.idx))) {
    vb = vk;
    b.idx = k.idx;
    b.seen = true;
  }
}

template <typename T>
struct alignas(sizeof(T) * 4) Act {
  T A, B, U, pad;
};

template <typename T>
struct alignas(sizeof(T) * 2) Con {
  T a, b;
};

__device__ __forceinline__ size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

// Merge the SPLIT partial results of each state into the chunk-0 thread.
template <typename T, int TILE>
__device__ __forceinline__ void block_merge(T& best, Best& bb, T* red_v, int* red_i, int* red_f,
                                            int tid, int ls, int grp) {
  constexpr int SPLIT = kBlock / TILE;
  if constexpr (SPLIT > 1) {
    __syncthreads();
    red_v[tid] = best;
    red_i[tid] = bb.idx;
    red_f[tid] = int(bb.seen) | (int(bb.nan) << 1);
    __syncthreads();
    if (grp == 0) {
      for (int k = 1; k < SPLIT; ++k) {
        const int j = ls + k * TILE;
        Best o{red_i[j], (red_f[j] & 1) != 0, (red_f[j] & 2) != 0};
        merge(best, bb, red_v[j], o);
      }
    }
  }
}

template <typename T, int TILE, bool WDEP>
__global__ void __launch_bounds__(kBlock)
gs_period_kernel(const LcmGsDesc* __restrict__ descs,
                 const LcmGsTask* __restrict__ tasks, int n_tasks) {
  constexpr int SPLIT = kBlock / TILE;
  constexpr int CH = kActChunk;
  extern __shared__ __align__(16) unsigned char smem[];

  // XCD-aware remap: hardware dispatches block b to XCD b % 8; give each XCD a
  // contiguous run of tasks. n_tasks is padded to a multiple of 8 by the host.
  const int b = blockIdx.x;
  const int per_xcd = n_tasks / kXcds;
  const int task_id = (b % kXcds) * per_xcd + (b / kXcds);
  const LcmGsTask tk = tasks[task_id];
  if (tk.desc < 0) return;  // padding task: whole block exits together

  const LcmGsDesc* __restrict__ D = descs + tk.desc;
  const int tid = threadIdx.x;
  const int ls = tid % TILE;
  // the waves sharing a state interleave over actions; grp is wave-uniform
  const int grp = __builtin_amdgcn_readfirstlane(tid / TILE);
  const int n_w = D->n_w, n_d = D->n_d, n_c = D->n_c, sd = tk.sd;
  const int wi = tk.w0 + ls;
  const bool valid = wi < n_w;
  const int wr = valid ? wi : 0;
  const T w = static_cast<const T*>(D->w_pts)[wr];
  const bool terminal = D->terminal != 0;
  const bool taste = D->has_taste != 0;
  // a stateful continuation needs a next-state grid with at least one point
  const bool interp = !terminal && D->n_targets > 0 && D->g_n > 0;
  const int g_n = interp ? D->g_n : 0;
  const int n_constr = D->n_constr;
  const int strict_mask = D->strict_mask;

  // LDS carve (the host sizes the launch for the largest g_n and n_constr)
  const size_t g_bytes = align16(size_t(g_n > 0 ? g_n : 1) * sizeof(T));
  T* ev = reinterpret_cast<T*>(smem);
  T* gp = reinterpret_cast<T*>(smem + g_bytes);
  Act<T>* act = reinterpret_cast<Act<T>*>(smem + 2 * g_bytes);
  Con<T>* con = reinterpret_cast<Con<T>*>(smem + 2 * g_bytes + CH * sizeof(Act<T>));
  T* red_v = reinterpret_cast<T*>(smem + 2 * g_bytes + CH * sizeof(Act<T>) +
                                  size_t(n_constr) * CH * sizeof(Con<T>));
  int* red_i = reinterpret_cast<int*>(red_v + kBlock);
  int* red_f = red_i + kBlock;

  Grid1D<T> g;
  g.kind = D->g_kind;
  g.n = g_n;
  g.start = T(D->g_start);
  const T g_stop = T(D->g_stop);
  g.step = (g_n > 1) ? (g_stop - g.start) / T(g_n - 1) : T(1);
  g.inv_step = T(1) / g.step;
  g.log_start = (g.kind == LCM_GRID_LOG) ? T(log(T(D->g_start))) : T(0);
  g.log_step = (g.kind == LCM_GRID_LOG && g_n > 1)
                   ? (T(log(g_stop)) - g.log_start) / T(g_n - 1)
                   : T(1);
  g.pts = gp;
  if (interp && g.kind == LCM_GRID_IRREG) {
    const T* src = static_cast<const T*>(D->g_pts);
    for (int k = tid; k < g_n; k += kBlock) gp[k] = src[k];
  }

  const T beta = T(D->beta);
  const int64_t wofs_u = WDEP ? int64_t(wr) * D->U_sw : 0;
  const int64_t wofs_a = WDEP ? int64_t(wr) * D->A_sw : 0;
  const int64_t wofs_b = WDEP ? int64_t(wr) * D->B_sw : 0;

  T best = -INFINITY;
  Best bb{0, false, false};
  T qc[LCM_MAX_TASTE_D];
  int qci[LCM_MAX_TASTE_D];
#pragma unroll
  for (int i = 0; i < LCM_MAX_TASTE_D; ++i) { qc[i] = -INFINITY; qci[i] = 0; }

  for (int d = 0; d < n_d; ++d) {
    T cs = T(0);
    bool finite_row = true;
    if (!terminal) {
      __syncthreads();  // previous readers of ev are done
      // EV[sd, d][k] = sum_r sum_j M_r[sd, d, j] * Vn_r[j, k]   (accumulated in T)
      int local_nonfinite = 0;
      for (int k = tid; k < g_n; k += kBlock) {
        T acc = T(0);
        for (int r = 0; r < D->n_targets; ++r) {
          const int nj = D->n_sd_next[r];
          const T* __restrict__ Mr = static_cast<const T*>(D->M[r]) + (int64_t(sd) * n_d + d) * nj;
          const T* __restrict__ Vr = static_cast<const T*>(D->Vn[r]);
          for (int j = 0; j < nj; ++j) {
            const T m = Mr[j];
            if (m != T(0)) acc += m * Vr[int64_t(j) * g_n + k];
          }
        }
        ev[k] = acc;
        local_nonfinite |= !is_finite(acc);
      }
      finite_row = __syncthreads_or(local_nonfinite) == 0;
      for (int s = 0; s < D->n_scalar; ++s) {
        const T m = static_cast<const T*>(D->ms[s])[int64_t(sd) * n_d + d];
        const T v = static_cast<const T*>(D->Vs[s])[0];
        cs += zero_safe(m, v);
      }
    }

    const T* __restrict__ Ug = static_cast<const T*>(D->U) + int64_t(sd) * D->U_ssd + wofs_u + int64_t(d) * n_c;
    const T* __restrict__ Ag = interp ? static_cast<const T*>(D->A) + int64_t(sd) * D->A_ssd + wofs_a + int64_t(d) * n_c : nullptr;
    const T* __restrict__ Bg = interp ? static_cast<const T*>(D->B) + int64_t(sd) * D->B_ssd + wofs_b + int64_t(d) * n_c : nullptr;

    for (int c0 = 0; c0 < n_c; c0 += CH) {
      const int cn = min(CH, n_c - c0);
      if constexpr (!WDEP) {
        __syncthreads();  // previous chunk consumed
        for (int i = tid; i < cn; i += kBlock) {
          Act<T> a;
          a.U = Ug[c0 + i];
          a.A = interp ? Ag[c0 + i] : T(0);
          a.B = interp ? Bg[c0 + i] : T(0);
          a.pad = T(0);
          act[i] = a;
          for (int k = 0; k < n_constr; ++k) {
            const T* Fa = static_cast<const T*>(D->Fa[k]) + int64_t(sd) * D->Fa_ssd[k] + int64_t(d) * n_c;
            const T* Fb = static_cast<const T*>(D->Fb[k]) + int64_t(sd) * D->Fb_ssd[k] + int64_t(d) * n_c;
            Con<T> f;
            f.a = Fa[c0 + i];
            f.b = Fb[c0 + i];
            con[k * CH + i] = f;
          }
        }
        __syncthreads();
      }
      for (int i = grp; i < cn; i += SPLIT) {
        const int c = c0 + i;
        T Uc, Ac, Bc;
        bool feas = valid;
        if constexpr (!WDEP) {
          const Act<T> a = act[i];
          Uc = a.U;
          Ac = a.A;
          Bc = a.B;
          for (int k = 0; k < n_constr; ++k) {
            const Con<T> f = con[k * CH + i];
            const T v = f.a * w + f.b;
            feas = feas && (((strict_mask >> k) & 1) ? (v > T(0)) : (v >= T(0)));
          }
        } else {
          Uc = Ug[c];
          Ac = interp ? Ag[c] : T(0);
          Bc = interp ? Bg[c] : T(0);
          for (int k = 0; k < n_constr; ++k) {
            const T* Fa = static_cast<const T*>(D->Fa[k]);
            const T* Fb = static_cast<const T*>(D->Fb[k]);
            const int64_t ia = int64_t(sd) * D->Fa_ssd[k] + int64_t(wr) * D->Fa_sw[k] + int64_t(d) * n_c + c;
            const int64_t ib = int64_t(sd) * D->Fb_ssd[k] + int64_t(wr) * D->Fb_sw[k] + int64_t(d) * n_c + c;
            const T v = Fa[ia] * w + Fb[ib];
            feas = feas && (((strict_mask >> k) & 1) ? (v > T(0)) : (v >= T(0)));
          }
        }
        if (__ballot(feas) == 0) continue;  // no lane of this wave can take action c
        T q;
        if (interp) {
          const T x = Ac * w + Bc;
          const T v = interp_row(ev, g_n, coordinate(g, x), finite_row);
          q = Uc + beta * (cs + v);
        } else if (terminal) {
          q = Uc;
        } else {
          q = Uc + beta * cs;
        }
        if (feas) {
          if (q != q) {
            bb.nan = true;
          } else if (!bb.seen || q > best) {
            best = q;
            bb.idx = d * n_c + c;
            bb.seen = true;
          }
        }
      }
    }

    if (taste) {
      // per-discrete-action max over the continuous axis, merged over groups
      block_merge<T, TILE>(best, bb, red_v, red_i, red_f, tid, ls, grp);
      const T qd = bb.nan ? T(NAN) : (bb.seen ? best : T(-INFINITY));
      const int id = bb.seen && !bb.nan ? bb.idx : 0;
#pragma unroll
      for (int i = 0; i < LCM_MAX_TASTE_D; ++i) {
        if (i == d) { qc[i] = qd; qci[i] = id; }
      }
      best = -INFINITY;
      bb = Best{0, false, false};
    }
  }

  if (taste) {
    if (grp == 0 && valid) {
      // logsum_and_softmax (logsum.py:32-64): m + scale * log(sum exp((q - m)/scale))
      const T lam = T(D->taste_scale);
      T m = -INFINITY;
      bool has_nan = false;
#pragma unroll
      for (int i = 0; i < LCM_MAX_TASTE_D; ++i) {
        if (i < n_d) {
          has_nan |= qc[i] != qc[i];
          m = fmax(m, qc[i]);
        }
      }
      const T fm = (m == -INFINITY) ? T(0) : m;
      T s = T(0);
#pragma unroll
      for (int i = 0; i < LCM_MAX_TASTE_D; ++i) {
        if (i < n_d) s += exp((qc[i] - fm) / lam);
      }
      T V = fm + lam * log(s);
      if (has_nan) V = T(NAN);
      const int64_t o = int64_t(sd) * n_w + wi;
      static_cast<T*>(D->V_out)[o] = V;
      for (int i = 0; i < LCM_MAX_TASTE_D; ++i) {
        if (i < n_d) D->pol_out[o * n_d + i] = qci[i];
      }
    }
    return;
  }

  block_merge<T, TILE>(best, bb, red_v, red_i, red_f, tid, ls, grp);
  if (grp == 0 && valid) {
    const int64_t o = int64_t(sd) * n_w + wi;
    static_cast<T*>(D->V_out)[o] = bb.nan ? T(NAN) : (bb.seen ? best : T(-INFINITY));
    D->pol_out[o] = (bb.seen && !bb.nan) ? bb.idx : 0;
  }
}

template <typename T, int TILE>
int launch_tile(lcm_ctx* ctx, const LcmGsDesc* d, const LcmGsTask* t, int n, size_t lds, bool wdep) {
  auto kern = wdep ? gs_period_kernel<T, TILE, true> : gs_period_kernel<T, TILE, false>;
  if (lds > 64 * 1024) {
    LCM_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
  }
  hipLaunchKernelGGL(kern, dim3(n), dim3(kBlock), lds, ctx->stream, d, t, n);
  LCM_HIP(hipGetLastError());
  return LCM_OK;
}

template <typename T>
int launch_t(lcm_ctx* ctx, const LcmGsDesc* d, const LcmGsTask* t, int n, int tile, size_t lds, bool wdep) {
  switch (tile) {
    case 64: return launch_tile<T, 64>(ctx, d, t, n, lds, wdep);
    case 128: return launch_tile<T, 128>(ctx, d, t, n, lds, wdep);
    case 256: return launch_tile<T, 256>(ctx, d, t, n, lds, wdep);
    default: return fail(LCM_E_ARG, "tile must be 64, 128 or 256");
  }
}

}  // namespace lcm

// Dynamic LDS bytes a launch needs for next-state grids of up to max_g points
// and up to max_constr constraints (mirrors the carve in gs_period_kernel).
extern "C" size_t lcm_gs_lds_bytes(int dtype, int max_g, int max_constr) {
  const size_t sz = dtype == LCM_DTYPE_F64 ? 8 : 4;
  const size_t g = (size_t(max_g > 0 ? max_g : 1) * sz + 15) & ~size_t(15);
  return 2 * g + lcm::kActChunk * 4 * sz + size_t(max_constr) * lcm::kActChunk * 2 * sz +
         lcm::kBlock * (sz + 8) + 64;
}

extern "C" int lcm_gs_launch_ex(lcm_ctx* ctx, int dtype, const LcmGsDesc* d_descs,
                                const LcmGsTask* d_tasks, int n_tasks, int tile,
                                size_t lds_bytes, int wdep) {
  using namespace lcm;
  if (!ctx || !d_descs || !d_tasks) return fail(LCM_E_ARG, "lcm_gs_launch: null argument");
  if (n_tasks <= 0 || n_tasks % kXcds != 0)
    return fail(LCM_E_ARG, "lcm_gs_launch: n_tasks must be a positive multiple of 8");
  if (lds_bytes > 160 * 1024) return fail(LCM_E_ARG, "lcm_gs_launch: LDS request above 160 KiB");
  if (dtype == LCM_DTYPE_F32) return launch_t<float>(ctx, d_descs, d_tasks, n_tasks, tile, lds_bytes, wdep != 0);
  if (dtype == LCM_DTYPE_F64) return launch_t<double>(ctx, d_descs, d_tasks, n_tasks, tile, lds_bytes, wdep != 0);
  return fail(LCM_E_ARG, "lcm_gs_launch: unknown dtype");
}

extern "C" int lcm_gs_launch(lcm_ctx* ctx, int dtype, const LcmGsDesc* d_descs,
                             const LcmGsTask* d_tasks, int n_tasks, int tile,
                             size_t lds_bytes) {
  return lcm_gs_launch_ex(ctx, dtype, d_descs, d_tasks, n_tasks, tile, lds_bytes, 1);
}
