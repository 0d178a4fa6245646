/*
 * pylcm_amd — C ABI of the MI355X (gfx950) backward-induction kernels.
 *
 * The reference has no native entry points on this path: its solve calls
 * XLA-compiled cores through `PeriodKernel.__call__`
 * (src/_lcm/solution/contract.py:347-426), once per (regime, period), from the
 * host loop in src/_lcm/solution/backward_induction.py:155-275 via
 * `_run_period_kernel` (:327-388). Each function below replaces one piece of
 * that contract; the comment on each names the reference code it stands in for.
 * Python binds this header with ctypes (pylcm_amd/_native.py).
 *
 * Conventions: every function returns 0 on success or a negative LCM_E* code;
 * lcm_last_error() returns the message of the calling thread's last failure.
 * Device pointers are plain `void*` obtained from lcm_alloc. A context owns one
 * HIP stream; contexts may be used from different threads, one thread each.
 */
#ifndef PYLCM_AMD_H
#define PYLCM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LCM_OK 0
#define LCM_E_ARG (-1)      /* invalid argument (shape, size, null pointer) */
#define LCM_E_OOM (-2)      /* device allocation failed */
#define LCM_E_LAUNCH (-3)   /* kernel launch or HIP runtime failure */
#define LCM_E_DEVICE (-4)   /* no usable device / init failed */

#define LCM_DTYPE_F32 0
#define LCM_DTYPE_F64 1

#define LCM_GRID_LIN 0
#define LCM_GRID_LOG 1
#define LCM_GRID_IRREG 2

#define LCM_MAX_TARGETS 4
#define LCM_MAX_CONSTR 4
#define LCM_MAX_SCALAR 4
#define LCM_MAX_TASTE_D 8
#define LCM_ACT_CHUNK 256 /* actions staged in LDS at a time */

/*
 * One (problem, regime, period) of the grid-search solve. Replaces the
 * compiled `max_Q_over_a` core (src/_lcm/regime_building/max_Q_over_a.py:30-191)
 * built from `Q_and_F` (src/_lcm/regime_building/Q_and_F.py:58-170).
 *
 * Tables hold element (sd, w, d, c) at  sd*X_ssd + w*X_sw + d*n_c + c  where sd
 * is the flat discrete-state index, w the continuous-state index, d the flat
 * discrete-action index and c the continuous-action index; a stride of 0
 * broadcasts. All float pointers are of the model dtype.
 *   Q(sd,w,d,c) = U + beta * CE,  CE = sum_s ms_s[sd,d]*Vs_s
 *                                   + interp(EV[sd,d], A*w_pts[w] + B)
 *   EV[sd,d][k] = sum_r sum_j M_r[(sd*n_d + d)*n_sd_next_r + j] * Vn_r[j*g_n + k]
 *   feasible iff  Fa_k*w + Fb_k >= 0  (> 0 when bit k of strict_mask is set),
 *   or, when use_interval is set, iff  Flo <= w <= Fhi  (the same set, given
 *   exactly when every Fa_k is -1, 0 or 1; tables (sd, d, c), stride Fi_ssd)
 * Outputs: V_out[sd*n_w + w] and pol_out[sd*n_w + w] (flat action index
 * d*n_c + c, first maximiser), or with taste shocks pol_out[(sd*n_w + w)*n_d + d]
 * (the continuous argmax per discrete action) and V the EV1 logsum.
 */
typedef struct LcmGsDesc {
  void* V_out;
  int32_t* pol_out;
  const void* w_pts;
  const void* U;
  const void* A;
  const void* B;
  const void* Fa[LCM_MAX_CONSTR];
  const void* Fb[LCM_MAX_CONSTR];
  const void* Vn[LCM_MAX_TARGETS];
  const void* M[LCM_MAX_TARGETS];
  const void* Vs[LCM_MAX_SCALAR];
  const void* ms[LCM_MAX_SCALAR];
  const void* g_pts;
  const void* Flo;  /* with use_interval: feasible iff Flo <= w <= Fhi */
  const void* Fhi;
  /* with use_interval: per (sd, d, chunk of LCM_ACT_CHUNK actions) the pair
   * (min Flo, max Fhi), so a block skips chunks no state of it can take */
  const void* Fchunk;
  /* with a linear next grid: A / step and (B - start) / step, each rounded
   * once to the model dtype (strides of A and B); the folded loop's
   * coordinate of the next state is Af * w + Bf */
  const void* Af;
  const void* Bf;
  int64_t U_ssd, U_sw, A_ssd, A_sw, B_ssd, B_sw;
  int64_t Fa_ssd[LCM_MAX_CONSTR], Fa_sw[LCM_MAX_CONSTR];
  int64_t Fb_ssd[LCM_MAX_CONSTR], Fb_sw[LCM_MAX_CONSTR];
  int64_t Fi_ssd, Fc_ssd;
  int32_t n_sd_next[LCM_MAX_TARGETS];
  int32_t n_sd, n_w, n_d, n_c;
  int32_t n_constr, strict_mask, n_targets, n_scalar;
  int32_t terminal, g_kind, g_n, has_taste;
  int32_t use_interval, pad0;
  double beta, taste_scale, g_start, g_stop;
} LcmGsDesc;

/* A work item: states [w0, w0 + tile) of discrete state sd of descriptor desc. */
typedef struct LcmGsTask {
  int32_t desc;
  int32_t sd;
  int32_t w0;
  int32_t pad;
} LcmGsTask;

/*
 * Standalone EV contraction of one (problem, regime, period): for every
 * discrete state sd and discrete action d,
 *   EV[(sd*n_d + d)*g_n + k] = sum_r sum_j M_r[(sd*n_d + d)*n_sd_next_r + j] * Vn_r[j*g_n + k]
 *   CS[sd*n_d + d]           = sum_s ms_s[sd*n_d + d] * Vs_s        (zero-safe)
 * reading M, Vn, ms, Vs from the grid-search descriptor gs. This is the
 * expectation over next nodes (Q_and_F.py:1145-1209) done once per row; the
 * forward simulation interpolates its rows at off-grid states.
 */
typedef struct LcmEvDesc {
  const LcmGsDesc* gs; /* device pointer */
  void* EV;
  void* CS;
  int32_t n_sd, n_d, g_n, pad;
  int32_t* fin; /* per EV row: 1 if every entry is finite (written with the row) */
} LcmEvDesc;

#define LCM_SIM_MAX_FEAT 4
#define LCM_SIM_MAX_REGIMES 8

/*
 * One (regime, period) of the forward simulation: per agent, the argmax over
 * the action grid at the agent's own (off-grid) state and the draw of its next
 * state. Stands in for `_simulate_regime_in_period`
 * (src/_lcm/simulation/simulate.py:465-620), `argmax_and_max_Q_over_a`
 * (src/_lcm/regime_building/max_Q_over_a.py:194-321) and
 * `calculate_next_states` / `_draw_random_regime_ids`
 * (src/_lcm/simulation/transitions.py:66-339).
 * Tables are indexed (sdd, d, c) with sdd the agent's discrete (non-process)
 * state: element at (sdd*n_d + d)*n_c + c.
 *   phi_0 = process value z (when n_nodes > 0); phi_k = op_k(sum_i fcoef_k[sdd*F+i]*phi_i + fconst_k[sdd])
 *   U = U0 + sum_k Uf_k*phi_k;   w' = A*w + B0 + sum_k Bf_k*phi_k;   feasible iff lo <= w <= hi
 *   CE = CS[row] + (1-t) interp(EV[row_lo, d], w') + t interp(EV[row_hi, d], w')
 *   with row = sd_base[sdd] + node*p_stride and (node, t) the process coordinate of z.
 */
typedef struct LcmSimDesc {
  const void* U0;
  const void* A;
  const void* B0;
  const void* lo;
  const void* hi;
  const void* Uf[LCM_SIM_MAX_FEAT];
  const void* Bf[LCM_SIM_MAX_FEAT];
  const void* fcoef[LCM_SIM_MAX_FEAT];  /* [n_sdd][LCM_SIM_MAX_FEAT] */
  const void* fconst[LCM_SIM_MAX_FEAT]; /* [n_sdd] */
  const void* EV;        /* [n_sd_full][n_d][g_n] */
  const void* CS;        /* [n_sd_full][n_d] */
  const int32_t* sd_base; /* [n_sdd] */
  const void* nodes;     /* process nodes [n_nodes] */
  const void* g_pts;     /* next continuous grid */
  const void* preg;      /* [n_sdd][n_d][n_regimes] regime probabilities */
  const void* Wd[LCM_SIM_MAX_REGIMES]; /* per target regime [n_sdd][n_d][n_sdd_r] */
  /* one constraint per side may read the agent's features:
   *   lo_eff = max(lo, LoB + sum_k LoF_k*phi_k),  hi_eff = min(hi, HiB + sum_k HiF_k*phi_k)
   * (null: absent; a strict side takes the next representable number) */
  const void* LoB;
  const void* HiB;
  const void* LoF[LCM_SIM_MAX_FEAT];
  const void* HiF[LCM_SIM_MAX_FEAT];
  int32_t lo_strict, hi_strict;
  int32_t n_sdd_next[LCM_SIM_MAX_REGIMES];
  int32_t feat_op[LCM_SIM_MAX_FEAT];
  int32_t n_feat, n_sdd, n_d, n_c;
  int32_t terminal, interp, n_nodes, p_stride;
  int32_t g_kind, g_n, n_regimes, regime_code;
  int32_t proc_kind, n_rows; /* n_rows = rows of EV (flat discrete states incl. process) */
  double beta, g_start, g_stop, pa, pb, pc;
  /* > 0: EV1 taste shocks; the discrete action is the first argmax of
   * Qc_d + taste_scale * (Gumbel - Euler's gamma) over the discrete cells d,
   * Qc_d the masked max over the continuous actions
   * (src/_lcm/regime_building/max_Q_over_a.py:255-346) */
  double taste_scale;
  /* 1: the EV rows of this regime are read from global memory (they do not
   * fit in LDS next to the action tables); 0: staged in LDS per cell */
  int32_t ev_global, pad_;
  /* per EV row (n_rows x n_d): 1 if finite, as written by lcm_ev_launch; read
   * instead of scanning the rows when they stay in global memory */
  const int32_t* ev_fin;
} LcmSimDesc;

/* Agent panel state (device arrays of n agents). */
typedef struct LcmAgents {
  void* w;          /* continuous state, model dtype */
  void* z;          /* process value, model dtype */
  int32_t* sdd;     /* discrete state (flat, within the agent's regime) */
  int32_t* regime;  /* regime code, -1 before the agent's first period */
  int32_t* start;   /* first period of the agent */
} LcmAgents;

/* Per-period outputs (device arrays of n agents). */
typedef struct LcmSimOut {
  int32_t* regime;
  int32_t* action;  /* flat action index d*n_c + c, -1 when not simulated */
  void* value;
  void* w;
  void* z;
  int32_t* sdd;
} LcmSimOut;

typedef struct lcm_ctx lcm_ctx;

/* Select and initialise the device. Replaces JAX backend initialisation. */
int lcm_init(int device);
/* Message of this thread's last failure ("" if none). */
const char* lcm_last_error(void);
/* Compute units, free and total device memory in bytes. */
int lcm_device_info(int* n_cu, size_t* free_bytes, size_t* total_bytes);

/* Context = one stream. Replaces XLA's per-device execution stream. */
int lcm_ctx_create(lcm_ctx** out);
int lcm_ctx_destroy(lcm_ctx* ctx);
int lcm_ctx_sync(lcm_ctx* ctx);

/* Device memory. Replaces XLA-owned device buffers (jax.device_put). */
int lcm_alloc(size_t bytes, void** dptr);
int lcm_free(void* dptr);
int lcm_host_alloc(size_t bytes, void** hptr); /* pinned host memory */
int lcm_host_free(void* hptr);
int lcm_memcpy_h2d(lcm_ctx* ctx, void* dst, const void* src, size_t bytes);
int lcm_memcpy_d2h(lcm_ctx* ctx, void* dst, const void* src, size_t bytes);
int lcm_memcpy_h2d_async(lcm_ctx* ctx, void* dst, const void* src, size_t bytes);
int lcm_memcpy_d2h_async(lcm_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Device-to-device copy on ctx's stream: a value function handed to a period
 * kernel from another buffer is copied into the slot its descriptors read
 * (the reference passes `next_regime_to_V_arr` as XLA buffers,
 * src/_lcm/solution/backward_induction.py:391-440). */
int lcm_memcpy_d2d_async(lcm_ctx* ctx, void* dst, const void* src, size_t bytes);
int lcm_memset_async(lcm_ctx* ctx, void* dst, int value, size_t bytes);

/*
 * Fused grid-search period: one launch evaluates every task (any mix of
 * problems, regimes and grid sizes). Replaces `_GridSearchPeriodKernel.__call__`
 * (src/_lcm/solution/grid_search.py:192-212) for all regimes of a period.
 * tile = states per workgroup (64, 128 or 256); lds_bytes = dynamic LDS as
 * returned by lcm_gs_lds_bytes for the launch's descriptors.
 */
int lcm_gs_launch(lcm_ctx* ctx, int dtype, const LcmGsDesc* d_descs,
                  const LcmGsTask* d_tasks, int n_tasks, int tile,
                  size_t lds_bytes);

/* HIP graph capture of the work submitted to ctx between begin and end.
 * Replaces the host loop's per-period dispatch (backward_induction.py:155-275). */
int lcm_graph_begin(lcm_ctx* ctx);
int lcm_graph_end(lcm_ctx* ctx, void** graph_exec);
int lcm_graph_launch(lcm_ctx* ctx, void* graph_exec);
int lcm_graph_destroy(void* graph_exec);

/* Event timing on the context's stream (milliseconds between two records). */
int lcm_event_create(void** ev);
int lcm_event_destroy(void* ev);
int lcm_event_record(lcm_ctx* ctx, void* ev);
int lcm_event_elapsed(void* ev_start, void* ev_stop, float* ms);
/* 1 when the work before the event's record has completed, 0 if not yet
 * (a non-blocking poll, for the serving loop's retirements). */
int lcm_event_query(void* ev);
int lcm_event_synchronize(void* ev);

/* Plain device-to-device copy kernel of `bytes` (a multiple of 16) on ctx's
 * stream. No counterpart in the reference: a kernel of known traffic whose
 * FETCH_SIZE / WRITE_SIZE counters calibrate the HBM traffic read for the
 * solve kernels. */
int lcm_copy_launch(lcm_ctx* ctx, void* dst, const void* src, size_t bytes);

/* Dynamic LDS bytes lcm_gs_launch needs when the launch's largest next-state
 * grid has max_g points and its descriptors have at most max_constr rows. */
size_t lcm_gs_lds_bytes(int dtype, int max_g, int max_constr);

/* sizeof the structs above, so a binding can check its mirror of the layout. */
int lcm_abi_sizes(size_t* gs_desc, size_t* gs_task);
int lcm_abi_sizes_sim(size_t* ev_desc, size_t* sim_desc, size_t* agents, size_t* sim_out);

/* EV contraction for n_desc descriptors (one workgroup per (desc, sd, d)). */
int lcm_ev_launch(lcm_ctx* ctx, int dtype, const LcmEvDesc* d_descs, int n_desc,
                  int max_rows);

/*
 * One period of forward simulation for all regimes active in it: sim_descs
 * holds one LcmSimDesc per regime code (n_regimes entries; a regime inactive
 * this period has U0 == NULL). Agents whose regime is r and whose start period
 * is <= period take the argmax over the action grid; their next state and
 * next regime are drawn with Philox4x32-10 keyed by seed, counter
 * (agent, period, stream, 0). cur/next are the agent panels of this and the
 * next period; out receives this period's record. d_counts[2 * n_regimes]
 * and d_order[n_regimes * n_agents] are device scratch for the partition of
 * the agents by regime: period p counts in half p % 2 and clears the other
 * half for period p + 1, and the call for period 0 clears both, so periods
 * are simulated in order from 0 with the same d_counts. Replaces the period
 * loop body of
 * src/_lcm/simulation/simulate.py:296-378.
 */
int lcm_sim_period_ex(lcm_ctx* ctx, int dtype, const LcmSimDesc* d_sim_descs, int n_regimes,
                      const LcmAgents* cur, const LcmAgents* next, const LcmSimOut* out,
                      int n_agents, int period, uint64_t seed, size_t lds_bytes,
                      int32_t* d_counts, int32_t* d_order);
size_t lcm_sim_lds_bytes(int dtype, int max_g, int max_ev_rows, int max_act);

/*
 * lcm_sim_period_ex without the caller's scratch: the partition counters and
 * agent order live in a static device array of the library (16 MiB per
 * device, held until the process ends), so the call allocates nothing.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). Returns LCM_E_ARG when n_regimes * n_agents exceeds
 * LCM_SIM_NOALLOC_MAX_ORDER. Every call on a device shares that device's
 * array: periods of one simulation are issued in order from 0 (period 0
 * clears the counters), and simulations that overlap on one device race.
 */
#define LCM_SIM_NOALLOC_MAX_ORDER (1 << 22)
int lcm_sim_period_noalloc(lcm_ctx* ctx, int dtype, const LcmSimDesc* d_sim_descs, int n_regimes,
                           const LcmAgents* cur, const LcmAgents* next, const LcmSimOut* out,
                           int n_agents, int period, uint64_t seed, size_t lds_bytes);

/* Raw Philox4x32-10 blocks: out[4*i .. 4*i+3] for counter (i, c1, c2, c3) and
 * key (k0, k1); for known-answer tests of the device generator. */
int lcm_philox(lcm_ctx* ctx, uint32_t* d_out, int n, uint32_t c1, uint32_t c2, uint32_t c3,
               uint32_t k0, uint32_t k1);

/* lcm_gs_launch with an explicit kernel variant: 0 = operands staged in LDS,
 * linear next-state grids and interval constraints (every descriptor must
 * have use_interval set and g_kind linear); 1 = staged, any grid, constraint
 * rows; 2 = general (tables may vary with the continuous state). One of the
 * LCM_GS_OCC_* flags may be or'ed in to choose the float32 kernel's register
 * budget: DENSE is compiled for 8 waves per SIMD (for launches that fill every
 * CU several times over), SPARSE unconstrained (fewer, faster waves, for
 * small launches); without a flag, launches of LCM_GS_DENSE_MIN_TASKS tasks
 * or more whose lds_bytes leave room for more than 6 workgroups per compute
 * unit run DENSE. Both give the same bits; float64 has one build. */
#define LCM_GS_OCC_DENSE 0x100
#define LCM_GS_OCC_SPARSE 0x200
#define LCM_GS_DENSE_MIN_TASKS 6144
int lcm_gs_launch_ex(lcm_ctx* ctx, int dtype, const LcmGsDesc* d_descs,
                     const LcmGsTask* d_tasks, int n_tasks, int tile,
                     size_t lds_bytes, int variant);

/*
 * One fused launch of a serving loop in a single call, all on ctx's stream:
 * record ev_start; copy n_tasks task records from pinned host memory h_tasks
 * to d_tasks; lcm_gs_launch_ex; record ev_ring; then n_copies device-to-host
 * copies (src[i] -> dst[i], bytes[i]; dst pinned) of results that this launch
 * completes; record ev_end. Events may be null (not recorded). No counterpart
 * in the reference, which has no server: it is for host loops that issue many
 * small launches, which then pay one foreign call per launch instead of six.
 */
int lcm_gs_serve_launch(lcm_ctx* ctx, int dtype, const LcmGsDesc* d_descs, const LcmGsTask* h_tasks,
                        LcmGsTask* d_tasks, int n_tasks, int tile, size_t lds_bytes, int variant,
                        void* ev_start, void* ev_ring, void* const* dst, const void* const* src,
                        const size_t* bytes, int n_copies, void* ev_end);

/* Milliseconds from event ref to each of evs[0..n) (all recorded and
 * complete), in one call; for the busy-time accounting of a serving loop (no
 * counterpart in the reference). */
int lcm_event_elapsed_batch(void* ref, void* const* evs, int n, float* out_ms);

/* ------------------------------------------------------------------------- */
/* Native solve server (continuous batching of grid-search solves)            */
/* ------------------------------------------------------------------------- */
/*
 * None of this has a counterpart in the reference, which solves one model at
 * a time from a host loop (src/_lcm/solution/backward_induction.py:155-275)
 * and has no server. The server's library threads place requests in device
 * memory, admit them to lanes (one HIP stream each), issue one fused
 * grid-search launch per lane holding the next backward step of every
 * request in flight there, and retire requests whose last step has run; the
 * host process only writes request files and collects results.
 *
 * Request file (little-endian): an LcmReqHeader; int64 arrays step_count,
 * step_g, step_k, step_variant (n_steps each: tasks of each backward step,
 * its largest next grid and constraint count for lcm_gs_lds_bytes, its
 * kernel variant), fetch_off, fetch_bytes (n_fetch each: outputs copied
 * back, as byte offsets from the request's base), uint64 ptr_mask (n_desc:
 * bit i set when 8-byte word i of the descriptor record is a pointer
 * relative to the base); then the n_desc descriptor records for base 0, the
 * n_task task rows {desc, sd, w0} as int32 (desc local to the request, steps
 * in order, 256-state tasks), and the table image (image_bytes, placed at
 * base + out_start). A request occupies total_bytes of device memory.
 */
#define LCM_REQ_MAGIC 0x31305145524d434cULL /* "LCMREQ01" */
#define LCM_REQ_HOLD 1 /* flags: keep the request's device memory after its solve (lcm_serve_release) */
#define LCM_REQ_TILE 256                    /* states per task row */
#define LCM_DESC_PTR_WORDS 36               /* leading pointer words of LcmGsDesc */

typedef struct LcmReqHeader {
  uint64_t magic;
  int64_t dtype, n_steps, n_desc, n_task, n_fetch;
  int64_t image_bytes, out_start, total_bytes, tile;
  int64_t flags; /* LCM_REQ_HOLD or 0 */
  int64_t reserved[5];
} LcmReqHeader;

/* Parse and check a request file without a device: info[0..8) = n_steps,
 * n_desc, n_task, n_fetch, image_bytes, total_bytes, sum of fetch_bytes,
 * pointer words flagged in ptr_mask; info[8] = largest step variant,
 * info[9] = largest lcm_gs_lds_bytes of a step. Returns LCM_OK or
 * LCM_E_ARG with the reason in lcm_last_error(). */
int lcm_req_inspect(const char* path, int64_t* info, int n_info);

/* The tile the autotuner picks for a launch of `units` 256-state tasks whose
 * states scan at most n_actions actions: that of the nearest entry (squared
 * distance in log2 units and log2 actions; the first of equals) of the
 * measured table (units[i], actions[i]) -> tiles[i]; with no entries 64 up
 * to 256 units, 128 up to 2048, else 256. Host only. */
int lcm_tile_choose(const int64_t* units, const int64_t* actions, const int32_t* tiles, int n, int64_t u,
                    int64_t a);

/* The `tile`-state tasks (tile 64 or 128) of n 256-state tasks: each task's
 * parts w0 + j * tile (j < 256 / tile) that start below its descriptor's
 * n_w (n_w_of_desc indexed by task desc), in order. Returns the count
 * written to out (room for n * 256 / tile). Host only. */
int64_t lcm_split_tasks(const LcmGsTask* in, int64_t n, const int32_t* n_w_of_desc, int tile, LcmGsTask* out);

/* The server's clock: seconds of CLOCK_MONOTONIC (the clock of the times
 * lcm_serve_fetch reports). */
double lcm_serve_clock(void);

typedef struct lcm_serve lcm_serve;

/* A server on the current device computing in `dtype`. budget_bytes of
 * device memory hold placed requests; max_desc descriptor slots; max_tasks
 * task records per fused launch; max_inflight requests placed at once;
 * result_bytes of pinned host memory hold fetched outputs until collected;
 * `lanes` streams, each with queue_depth task buffers (launches queued
 * ahead); `placers` placement threads. fixed_tile 64/128/256, or 0 to pick
 * each launch's tile with lcm_tile_choose over the n_table entries. flags:
 * LCM_SERVE_LONGEST_FIRST orders each launch's tasks by estimated work
 * (states x actions per state), longest first on every XCD; otherwise tasks
 * run in admission order. Neither changes a result. */
#define LCM_SERVE_LONGEST_FIRST 1
int lcm_serve_create(lcm_serve** out, int dtype, size_t budget_bytes, int max_desc, int max_tasks,
                     int max_inflight, size_t result_bytes, int lanes, int queue_depth, int placers,
                     int fixed_tile, const int64_t* table_units, const int64_t* table_actions,
                     const int32_t* table_tiles, int n_table, int flags);

/* Queue request `rid` (any id unique among those not yet fetched) from a
 * request file; the server removes the file once read. */
int lcm_serve_submit(lcm_serve* s, int64_t rid, const char* path);

/* Wait up to timeout_ms for completed requests; writes up to max_n of them
 * (id, status: LCM_OK or an error code) and returns how many. */
int lcm_serve_wait(lcm_serve* s, int64_t* rids, int32_t* status, int max_n, int timeout_ms);

/* Take a completed request: copies its fetched outputs (in file order,
 * packed) to dst when its status is LCM_OK and nbytes covers them; times[3]
 * = submitted, placed, done (lcm_serve_clock); err receives the message of
 * a failed request. Returns the request's status (LCM_E_ARG for an unknown
 * id). */
int lcm_serve_fetch(lcm_serve* s, int64_t rid, void* dst, size_t nbytes, double* times, char* err,
                    size_t err_len);

/* A completed request submitted with LCM_REQ_HOLD keeps its device memory
 * (value functions, tables and descriptors stay valid, the solve's work is
 * done) until lcm_serve_release: *base = its first byte, *descs = its first
 * descriptor record in the server's pool; for device work on the solved
 * request, such as a forward simulation. */
int lcm_serve_held(lcm_serve* s, int64_t rid, uint64_t* base, uint64_t* descs);
int lcm_serve_release(lcm_serve* s, int64_t rid);

/* Counters: out[0..12) = launches, task records launched, launches at tile
 * 64, 128, 256, nanoseconds the lanes spent assembling and issuing launches,
 * nanoseconds the placers spent placing, requests placed, requests retired,
 * device bytes in use, device bytes of the budget, the most ever in use. */
int lcm_serve_stats(const lcm_serve* s, int64_t* out, int n);

/* Milliseconds the device had at least one fused launch running (the union
 * over lanes of each launch's task upload and kernel, from events). */
int lcm_serve_busy_ms(lcm_serve* s, double* ms);

/* Stop the threads (requests not completed are dropped; their request files
 * removed), wait for the lanes' streams and free everything. */
int lcm_serve_destroy(lcm_serve* s);

/* ------------------------------------------------------------------------- */
/* DC-EGM with the FUES upper envelope (all arithmetic FP64)                  */
/* ------------------------------------------------------------------------- */

#define LCM_EGM_MAX_TARGETS 4
#define LCM_EGM_MAX_ROWS 8      /* discrete-action rows of one child block */
#define LCM_EGM_MAX_CAND 4096   /* constrained + Euler candidates per column */
#define LCM_TAPE_MAX_INS 15     /* instructions of one expression tape */
#define LCM_EGM_MAX_ROWS_PER_COL 64  /* child rows (all targets and nodes) one column reads */

/*
 * One carry row of a child regime: the refined endogenous grid, value and
 * marginal-utility rows (`EGMCarry`, src/_lcm/egm/carry.py:25-60), NaN-padded to
 * `len`; *n_valid counts the non-NaN grid entries (written by the kernel that
 * produced the row, or by the host for a terminal regime's closed-form rows,
 * src/_lcm/egm/terminal.py:85-180).
 */
typedef struct LcmEgmRow {
  const double* grid;
  const double* value;
  const double* marg;
  const int* n_valid;
  int len;
  int pad;
} LcmEgmRow;

/*
 * One (discrete state, discrete action) column of a DC-EGM regime at one
 * period. Replaces the per-combo solve `_get_solve_one_combo`
 * (src/_lcm/egm/step_core.py:130-312) with its continuation
 * (src/_lcm/egm/continuation.py:323-500,959-1140,1353-1534), Euler inversion
 * (src/_lcm/egm/euler.py:18-68), constrained candidates (step_core.py:443-514),
 * FUES refinement (src/_lcm/egm/upper_envelope/fues.py:151-1189) and publish
 * (step_core.py:517-616).
 *
 * Expression tapes (u_prog, i_prog: int16 programs, see pylcm_amd/_tape.py)
 * evaluate the utility u(c) with u'(c) and the inverse marginal utility; u_c
 * and i_c point at this column's constants. Per target t (regime reachable
 * next period): kind 0 adds tprob * tscalar to the value only (stateless
 * terminal); kind 1 reads K child nodes with weights tW[k], each a block of
 * tN rows tRows[k*tN + d'] aggregated by logsum (tscale > 0) or hard max, at
 * the next resources tR[j] with gradient tG[j] for savings node j.
 * Outputs: the refined carry rows (n_pad each), counts[0] = kept envelope
 * points (> n_pad: overflow, rows poisoned), counts[1] = valid grid entries,
 * and the published V on the state grid W. cand is scratch of
 * 4 * (n_con + n_s) + 2 doubles (candidate rows, EV at the first node, and
 * an all-dead flag handed from the envelope pass to the publishing pass).
 */
typedef struct LcmEgmCol {
  const double* A;
  const double* W;
  const short* u_prog;
  const double* u_c;
  const short* i_prog;
  const double* i_c;
  const double* tR[LCM_EGM_MAX_TARGETS];
  const double* tG[LCM_EGM_MAX_TARGETS];
  const double* tW[LCM_EGM_MAX_TARGETS];
  const LcmEgmRow* tRows[LCM_EGM_MAX_TARGETS];
  double tprob[LCM_EGM_MAX_TARGETS];
  double tscalar[LCM_EGM_MAX_TARGETS];
  double tscale[LCM_EGM_MAX_TARGETS];
  int tK[LCM_EGM_MAX_TARGETS];
  int tN[LCM_EGM_MAX_TARGETS];
  int tkind[LCM_EGM_MAX_TARGETS];
  double* out_grid;
  double* out_policy;
  double* out_value;
  double* out_marg;
  int* out_counts;
  double* out_V;
  double* cand;
  double beta;
  double ratio;       /* (1 / 1e-4) ** (1 / max(n_con - 1, 1)) */
  double jump_thresh;
  double w_max;       /* max over W */
  int n_s;
  int n_w;
  int n_con;
  int n_pad;
  int n_scan;         /* FUES window; <= 0 scans exhaustively */
  int n_targets;
  int pad0;
  int pad1;
} LcmEgmCol;

/* V of one discrete-state combo: logsum (scale > 0) or max over its n_d
 * column rows (src/_lcm/egm/step.py:731-738), written in the output dtype. */
typedef struct LcmEgmAgg {
  const double* cols;  /* [n_d * n_w] */
  void* out;           /* [n_w] */
  double scale;
  int n_d;
  int n_w;
} LcmEgmAgg;

/* One call solves n_cols columns: a continuation/Euler kernel over blocks of
 * 64 savings nodes, one workgroup per column for the envelope, then wide
 * kernels (many workgroups per column) for the marginal row and the V row on
 * W. lds_bytes from lcm_egm_lds_bytes(max candidates of any column);
 * max_rows = the most child rows any column reads (sum over its targets of
 * nodes x rows, <= LCM_EGM_MAX_ROWS_PER_COL); max_s = the most savings nodes.
 * Columns of one call must not read each other's outputs. */
int lcm_egm_launch(lcm_ctx* ctx, const LcmEgmCol* d_cols, int n_cols, size_t lds_bytes, int max_rows,
                   int max_s);
/* lcm_egm_launch with the savings nodes per continuation workgroup fixed:
 * 64, 128 or 256 (0 = chosen per launch as lcm_egm_launch does: more nodes
 * per workgroup when the launch has thousands of workgroups). Every choice
 * computes the same values; the choice only moves work between threads. */
int lcm_egm_launch_ex(lcm_ctx* ctx, const LcmEgmCol* d_cols, int n_cols, size_t lds_bytes, int max_rows,
                      int max_s, int nodes_per_block);
int lcm_egm_aggregate(lcm_ctx* ctx, int dtype, const LcmEgmAgg* d_aggs, int n_aggs, int max_w);
size_t lcm_egm_lds_bytes(int n_cand);
int lcm_abi_sizes_egm(size_t* row, size_t* col, size_t* agg);

/* Evaluate an expression tape at n points for testing: out_v[i], out_d[i] =
 * value and derivative at x[i] with constants c. */
int lcm_tape_eval(lcm_ctx* ctx, const short* d_prog, const double* d_c, const double* d_x,
                  double* d_v, double* d_d, int n);

#ifdef __cplusplus
}
#endif

#endif /* PYLCM_AMD_H */
