// Native solve server: continuous batching of grid-search solves with the
// whole serving loop in library threads.
//
//   placers  read a request file, wait for device memory, upload the table
//            image and the relocated descriptor records on their own stream,
//            and hand the request to the lane with the fewest in flight;
//   lanes    (one HIP stream and thread each) assemble one fused launch from
//            the next backward step of every request in flight on the lane,
//            pick the launch's state tile, issue task upload + kernel + the
//            result copies of requests that this launch completes, and
//            retire launches whose end event has fired;
//   the host process writes request files (lowering runs there) and takes
//            completed requests with lcm_serve_wait / lcm_serve_fetch.
//
// A request's steps run in order on one stream, so step s + 1 reads the V
// rows step s wrote; requests never read each other's memory, so what else
// is in flight does not change a result. Declarations and the file format
// are in include/pylcm_amd.h; the reference has no server.

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstddef>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.h"

static_assert(offsetof(LcmGsDesc, U_ssd) == LCM_DESC_PTR_WORDS * 8, "pointer words of LcmGsDesc");
static_assert(sizeof(LcmReqHeader) == 16 * 8, "request header size");

namespace lcm {
namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int64_t now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// First-fit allocator over [0, size) in multiples of `unit`.
class Slab {
 public:
  void init(size_t bytes, size_t unit) {
    unit_ = unit;
    size_ = bytes / unit * unit;
    free_.clear();
    if (size_) free_[0] = size_;
    used_ = peak_ = 0;
  }
  size_t round(size_t n) const { return std::max<size_t>(1, (n + unit_ - 1) / unit_) * unit_; }
  bool alloc(size_t n, size_t* off, size_t* len) {
    const size_t m = round(n);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < m) continue;
      const size_t o = it->first, rest = it->second - m;
      free_.erase(it);
      if (rest) free_[o + m] = rest;
      *off = o;
      *len = m;
      used_ += m;
      peak_ = std::max(peak_, used_);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t len) {
    used_ -= len;
    auto it = free_.emplace(off, len).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t size() const { return size_; }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }

 private:
  std::map<size_t, size_t> free_;
  size_t unit_ = 1, size_ = 0, used_ = 0, peak_ = 0;
};

int fallback_tile(int64_t units) { return units <= 256 ? 64 : (units <= 2048 ? 128 : 256); }

int choose_tile(const int64_t* units, const int64_t* acts, const int32_t* tiles, int n, int64_t u, int64_t a) {
  const double lu = std::log2(double(std::max<int64_t>(u, 1)));
  const double la = std::log2(double(std::max<int64_t>(a, 1)));
  int best = 0;
  double bd = HUGE_VAL;
  for (int i = 0; i < n; ++i) {
    const double du = std::log2(double(std::max<int64_t>(units[i], 1))) - lu;
    const double da = std::log2(double(std::max<int64_t>(acts[i], 1))) - la;
    const double d = du * du + da * da;
    if (d < bd) {
      bd = d;
      best = tiles[i];
    }
  }
  return best ? best : fallback_tile(u);
}

int64_t split_tasks(const LcmGsTask* in, int64_t n, const int32_t* n_w_of_desc, int tile, LcmGsTask* out) {
  const int k = LCM_REQ_TILE / tile;
  int64_t m = 0;
  for (int64_t i = 0; i < n; ++i) {
    const LcmGsTask t = in[i];
    const int32_t n_w = n_w_of_desc[t.desc];
    for (int j = 0; j < k; ++j) {
      const int32_t w = t.w0 + j * tile;
      if (w >= n_w) break;  // parts start in increasing order
      out[m++] = LcmGsTask{t.desc, t.sd, w, 0};
    }
  }
  return m;
}

// The metadata of a request file (everything before the descriptor records).
struct ReqMeta {
  LcmReqHeader h{};
  std::vector<int64_t> count, g, k, variant, fetch_off, fetch_bytes;
  std::vector<uint64_t> ptr_mask;
  std::vector<size_t> lds;
  int64_t meta_bytes = 0;  // header + arrays
  int64_t bulk_bytes = 0;  // descriptors + tasks + image
  int64_t fetch_total = 0;
};

bool read_exact(int fd, void* dst, size_t n, off_t off) {
  char* p = static_cast<char*>(dst);
  while (n) {
    const ssize_t r = pread(fd, p, n, off);
    if (r <= 0) return false;
    p += r;
    off += r;
    n -= size_t(r);
  }
  return true;
}

// Read and check the metadata of an open request file; on failure returns
// false with the reason in `err`.
bool read_meta(int fd, ReqMeta& m, std::string& err) {
  if (!read_exact(fd, &m.h, sizeof(m.h), 0)) return err = "short request header", false;
  const LcmReqHeader& h = m.h;
  if (h.magic != LCM_REQ_MAGIC) return err = "not a request file (bad magic)", false;
  if (h.dtype != LCM_DTYPE_F32 && h.dtype != LCM_DTYPE_F64) return err = "request dtype unknown", false;
  if (h.tile != LCM_REQ_TILE) return err = "request tasks must be 256-state tasks", false;
  if (h.flags & ~int64_t(LCM_REQ_HOLD)) return err = "unknown request flags", false;
  const int64_t lim = int64_t(1) << 31;
  if (h.n_steps < 0 || h.n_steps > lim || h.n_desc < 0 || h.n_desc > lim || h.n_task < 0 ||
      h.n_task > lim || h.n_fetch < 0 || h.n_fetch > lim || h.image_bytes < 0 || h.out_start < 0 ||
      h.total_bytes < 0)
    return err = "request header sizes out of range", false;
  if (h.out_start + h.image_bytes > h.total_bytes) return err = "request image beyond its total bytes", false;
  struct stat st;
  if (fstat(fd, &st) != 0 ||
      st.st_size < int64_t(sizeof(LcmReqHeader)) + 8 * (4 * h.n_steps + 2 * h.n_fetch + h.n_desc))
    return err = "request file shorter than its header says", false;
  const size_t n_arr = size_t(4 * h.n_steps + 2 * h.n_fetch);
  std::vector<int64_t> arr(n_arr);
  m.ptr_mask.resize(size_t(h.n_desc));
  off_t off = sizeof(LcmReqHeader);
  if (n_arr && !read_exact(fd, arr.data(), n_arr * 8, off)) return err = "short request arrays", false;
  off += off_t(n_arr * 8);
  if (h.n_desc && !read_exact(fd, m.ptr_mask.data(), size_t(h.n_desc) * 8, off))
    return err = "short request pointer masks", false;
  off += off_t(h.n_desc * 8);
  m.meta_bytes = off;
  const int64_t* a = arr.data();
  const size_t S = size_t(h.n_steps), F = size_t(h.n_fetch);
  m.count.assign(a, a + S);
  m.g.assign(a + S, a + 2 * S);
  m.k.assign(a + 2 * S, a + 3 * S);
  m.variant.assign(a + 3 * S, a + 4 * S);
  m.fetch_off.assign(a + 4 * S, a + 4 * S + F);
  m.fetch_bytes.assign(a + 4 * S + F, a + 4 * S + 2 * F);
  int64_t sum = 0;
  m.lds.resize(S);
  for (size_t s = 0; s < S; ++s) {
    if (m.count[s] < 0) return err = "negative step task count", false;
    sum += m.count[s];
    if (m.variant[s] < 0 || m.variant[s] > 2) return err = "unknown step variant", false;
    if (m.k[s] < 0 || m.k[s] > LCM_MAX_CONSTR || m.g[s] < 0 || m.g[s] > lim)
      return err = "step LDS sizes out of range", false;
    m.lds[s] = lcm_gs_lds_bytes(int(h.dtype), int(m.g[s]), int(m.k[s]));
    if (m.lds[s] > 160 * 1024)
      return err = "a next-state grid of " + std::to_string(m.g[s]) + " points needs " + std::to_string(m.lds[s]) +
                   " B of LDS per workgroup, above the 163840 B of a gfx950 compute unit",
             false;
  }
  if (sum != h.n_task) return err = "step task counts do not add up to the task rows", false;
  m.fetch_total = 0;
  for (size_t f = 0; f < F; ++f) {
    if (m.fetch_off[f] < 0 || m.fetch_bytes[f] < 0 || m.fetch_off[f] + m.fetch_bytes[f] > h.total_bytes)
      return err = "fetched output outside the request's memory", false;
    m.fetch_total += m.fetch_bytes[f];
  }
  for (uint64_t mask : m.ptr_mask)
    if (mask >> LCM_DESC_PTR_WORDS) return err = "pointer mask flags a non-pointer word", false;
  m.bulk_bytes = h.n_desc * int64_t(sizeof(LcmGsDesc)) + h.n_task * 12 + h.image_bytes;
  if (st.st_size != m.meta_bytes + m.bulk_bytes)
    return err = "request file size does not match its header", false;
  return true;
}

struct Req {
  int64_t rid = 0;
  std::string path;
  int status = LCM_OK;
  std::string error;
  double t_submit = 0, t_placed = 0, t_done = 0;
  // placed
  int n_steps = 0, s = 0;
  std::vector<int64_t> step_off;  // n_steps + 1 offsets into tasks
  std::vector<LcmGsTask> tasks;   // desc already global
  std::vector<size_t> lds;
  std::vector<int> variant;
  std::vector<int64_t> n_act;
  char* base = nullptr;
  size_t slab_off = 0, slab_len = 0, desc_off = 0, desc_len = 0, res_off = 0, res_len = 0;
  bool has_res = false;
  bool hold = false;  // keep the slab and descriptors after the solve
  std::vector<int64_t> fetch_off, fetch_bytes;
  int lane = -1;
};

struct Launch {
  hipEvent_t start = nullptr, end = nullptr;
  std::vector<Req*> done;
  bool timed = false;
};

struct Lane {
  lcm_ctx ctx{};
  std::vector<LcmGsTask*> h_ring, d_ring;
  std::vector<hipEvent_t> ring_ev;
  std::vector<char> ring_used;
  int ring_i = 0;
  std::deque<Req*> incoming;  // guarded by the server mutex
  std::vector<Req*> inflight;
  std::deque<Launch> pending;
  std::vector<hipEvent_t> ev_pool;
  int load = 0;  // requests assigned and not retired (server mutex)
  std::thread th;
  // reused per launch
  std::vector<std::pair<int64_t, int64_t>> keys;
  std::vector<LcmGsTask> order_tmp;
  std::vector<void*> dst;
  std::vector<const void*> src;
  std::vector<size_t> bytes;
};

}  // namespace
}  // namespace lcm

using namespace lcm;

struct lcm_serve {
  int device = 0, dtype = 0;
  int max_tasks = 0, max_inflight = 0, fixed_tile = 0, flags = 0;
  std::vector<int64_t> t_units, t_acts;
  std::vector<int32_t> t_tiles;
  char* slab_dev = nullptr;
  LcmGsDesc* desc_dev = nullptr;
  char* res_host = nullptr;
  Slab slab, desc_slab, res_slab;
  std::vector<int32_t> n_w_of_desc;  // by global descriptor index
  std::vector<int64_t> act_of_desc;  // actions per state, by global descriptor index
  std::vector<Lane*> lanes;
  std::vector<std::thread> placers;
  hipEvent_t t0 = nullptr;

  std::mutex mu;
  std::condition_variable cv_submit, cv_mem, cv_lane, cv_done;
  std::deque<Req*> submitted;
  std::deque<Req*> done_q;
  std::unordered_map<int64_t, Req*> finished;
  struct Held {
    size_t slab_off, slab_len, desc_off, desc_len;
  };
  std::unordered_map<int64_t, Held> held;  // completed requests keeping their memory
  int n_placed = 0;
  bool stop = false;

  std::mutex st_mu;  // statistics
  int64_t launches = 0, tasks = 0, by_tile[3] = {0, 0, 0}, loop_ns = 0, place_ns = 0, placed = 0, retired = 0;
  std::vector<std::pair<float, float>> intervals;

  void fail_req(Req* r, int code, const std::string& msg) {
    r->status = code;
    r->error = msg;
    r->t_done = now_s();
    {
      std::lock_guard<std::mutex> lk(mu);
      done_q.push_back(r);
    }
    cv_done.notify_all();
  }

  void place_main();
  int place(Req* r, char*& stage, size_t& stage_bytes, hipStream_t stream, std::string& err);
  void lane_main(Lane* L);
  void launch(Lane* L);
  bool retire_ready(Lane* L, bool block);
  hipEvent_t take_event(Lane* L);
  void order_longest_first(Lane* L, LcmGsTask* h, int64_t n, int64_t npad, int tile);
};

// Reorder a launch's n tasks (npad with padding) so that the kernel's
// workgroups start the longest ones first on every XCD: rank the tasks by
// estimated work (states x actions per state, then the higher w0, whose
// states usually have more feasible actions), and deal the ranks round-robin
// over the eight contiguous runs the kernel hands to the eight XCDs (rank i
// to run i % 8, position i / 8), so each run is longest-first and the runs
// carry equal shares. Results do not depend on the order of tasks.
void lcm_serve::order_longest_first(Lane* L, LcmGsTask* h, int64_t n, int64_t npad, int tile) {
  auto& keys = L->keys;
  keys.resize(size_t(n));
  for (int64_t i = 0; i < n; ++i) {
    const LcmGsTask& t = h[i];
    const int64_t states = std::min<int64_t>(tile, int64_t(n_w_of_desc[size_t(t.desc)]) - t.w0);
    keys[size_t(i)] = {-(states * act_of_desc[size_t(t.desc)] * 65536 + t.w0), i};
  }
  std::sort(keys.begin(), keys.end());
  L->order_tmp.assign(h, h + npad);
  const int64_t per = npad / 8;
  for (int64_t r = 0; r < npad; ++r) {
    const LcmGsTask& t = r < n ? L->order_tmp[size_t(keys[size_t(r)].second)] : L->order_tmp[size_t(r)];
    h[(r % 8) * per + r / 8] = t;
  }
}

hipEvent_t lcm_serve::take_event(Lane* L) {
  if (!L->ev_pool.empty()) {
    hipEvent_t e = L->ev_pool.back();
    L->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// ---------------------------------------------------------------------------
// placement
// ---------------------------------------------------------------------------

int lcm_serve::place(Req* r, char*& stage, size_t& stage_bytes, hipStream_t stream, std::string& err) {
  const int fd = open(r->path.c_str(), O_RDONLY);
  if (fd < 0) return err = "cannot open request file " + r->path, LCM_E_ARG;
  ReqMeta m;
  if (!read_meta(fd, m, err)) {
    close(fd);
    unlink(r->path.c_str());
    err = r->path + ": " + err;
    return LCM_E_ARG;
  }
  const LcmReqHeader& h = m.h;
  if (h.dtype != dtype) {
    close(fd);
    unlink(r->path.c_str());
    return err = "request dtype differs from the server's", LCM_E_ARG;
  }
  const size_t need_desc = size_t(h.n_desc) * sizeof(LcmGsDesc);
  const size_t need_res = size_t(m.fetch_total);
  {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      if (stop) {
        close(fd);
        unlink(r->path.c_str());
        return err = "server closed", LCM_E_ARG;
      }
      if (size_t(h.total_bytes) > slab.size() || need_desc > desc_slab.size() || need_res > res_slab.size()) {
        close(fd);
        unlink(r->path.c_str());
        err = "request needs " + std::to_string(h.total_bytes) + " B of device memory, " +
              std::to_string(need_desc) + " B of descriptors and " + std::to_string(need_res) +
              " B of results; the server holds " + std::to_string(slab.size()) + ", " +
              std::to_string(desc_slab.size()) + " and " + std::to_string(res_slab.size()) + " B";
        return LCM_E_OOM;
      }
      if (n_placed < max_inflight && slab.alloc(size_t(h.total_bytes), &r->slab_off, &r->slab_len)) {
        if (desc_slab.alloc(std::max<size_t>(need_desc, 1), &r->desc_off, &r->desc_len)) {
          if (need_res == 0 || res_slab.alloc(need_res, &r->res_off, &r->res_len)) {
            r->has_res = need_res > 0;
            break;
          }
          desc_slab.release(r->desc_off, r->desc_len);
        }
        slab.release(r->slab_off, r->slab_len);
      }
      cv_mem.wait_for(lk, std::chrono::milliseconds(20));
    }
    ++n_placed;
  }
  const int64_t t0n = now_ns();
  auto unwind = [&](int code) {
    std::lock_guard<std::mutex> lk(mu);
    slab.release(r->slab_off, r->slab_len);
    desc_slab.release(r->desc_off, r->desc_len);
    if (r->has_res) res_slab.release(r->res_off, r->res_len);
    r->has_res = false;
    --n_placed;
    cv_mem.notify_all();
    return code;
  };
  if (size_t(m.bulk_bytes) > stage_bytes) {
    if (stage) hipHostFree(stage);
    stage = nullptr;
    stage_bytes = 0;
    const size_t want = std::max<size_t>(size_t(m.bulk_bytes), size_t(16) << 20);
    if (hipHostMalloc(reinterpret_cast<void**>(&stage), want, hipHostMallocDefault) != hipSuccess) {
      stage = nullptr;
      close(fd);
      unlink(r->path.c_str());
      err = "pinned staging allocation failed";
      return unwind(LCM_E_OOM);
    }
    stage_bytes = want;
  }
  const bool ok = m.bulk_bytes == 0 || read_exact(fd, stage, size_t(m.bulk_bytes), off_t(m.meta_bytes));
  close(fd);
  unlink(r->path.c_str());
  if (!ok) {
    err = "short read of request file " + r->path;
    return unwind(LCM_E_ARG);
  }
  LcmGsDesc* D = reinterpret_cast<LcmGsDesc*>(stage);
  const int32_t* rows = reinterpret_cast<const int32_t*>(stage + need_desc);
  const char* image = stage + need_desc + size_t(h.n_task) * 12;
  char* base = slab_dev + r->slab_off;
  const int64_t desc_base = int64_t(r->desc_off / sizeof(LcmGsDesc));
  // relocate the records to this request's base; note each descriptor's
  // n_w (task splitting) and actions per state (tile choice)
  std::vector<int64_t> act(size_t(h.n_desc));
  for (int64_t i = 0; i < h.n_desc; ++i) {
    uint64_t* w = reinterpret_cast<uint64_t*>(&D[i]);
    for (uint64_t mask = m.ptr_mask[size_t(i)]; mask; mask &= mask - 1)
      w[__builtin_ctzll(mask)] += reinterpret_cast<uint64_t>(base);
    n_w_of_desc[size_t(desc_base + i)] = D[i].n_w;
    act[size_t(i)] = int64_t(D[i].n_d) * D[i].n_c;
    act_of_desc[size_t(desc_base + i)] = act[size_t(i)];
  }
  hipError_t e = hipSuccess;
  if (h.image_bytes)
    e = hipMemcpyAsync(base + h.out_start, image, size_t(h.image_bytes), hipMemcpyHostToDevice, stream);
  if (e == hipSuccess && need_desc)
    e = hipMemcpyAsync(desc_dev + desc_base, D, need_desc, hipMemcpyHostToDevice, stream);
  // task rows: global descriptor indexes, per-step offsets and figures
  r->n_steps = int(h.n_steps);
  r->tasks.resize(size_t(h.n_task));
  r->step_off.resize(size_t(h.n_steps) + 1);
  r->lds = m.lds;
  r->variant.assign(m.variant.begin(), m.variant.end());
  r->n_act.assign(size_t(h.n_steps), 1);
  bool rows_ok = true;
  int64_t j = 0;
  for (int64_t s = 0; s < h.n_steps; ++s) {
    r->step_off[size_t(s)] = j;
    for (int64_t q = 0; q < m.count[size_t(s)]; ++q, ++j) {
      const int32_t d = rows[3 * j], sd = rows[3 * j + 1], w0 = rows[3 * j + 2];
      if (d < 0 || d >= h.n_desc || sd < 0 || w0 < 0) {
        rows_ok = false;
        break;
      }
      r->tasks[size_t(j)] = LcmGsTask{int32_t(desc_base + d), sd, w0, 0};
      r->n_act[size_t(s)] = std::max(r->n_act[size_t(s)], act[size_t(d)]);
    }
    if (!rows_ok) break;
  }
  r->step_off[size_t(h.n_steps)] = j;
  // the staging buffer is reused by the next request: wait for the uploads
  const hipError_t e2 = hipStreamSynchronize(stream);
  if (e == hipSuccess) e = e2;
  if (!rows_ok) {
    err = "task row outside the request's descriptors";
    return unwind(LCM_E_ARG);
  }
  if (e != hipSuccess) {
    err = std::string("request upload: ") + hipGetErrorString(e);
    return unwind(LCM_E_LAUNCH);
  }
  r->base = base;
  r->hold = (h.flags & LCM_REQ_HOLD) != 0;
  r->fetch_off = std::move(m.fetch_off);
  r->fetch_bytes = std::move(m.fetch_bytes);
  r->t_placed = now_s();
  {
    std::lock_guard<std::mutex> lk(st_mu);
    place_ns += now_ns() - t0n;
    ++placed;
  }
  {
    std::lock_guard<std::mutex> lk(mu);
    Lane* best = lanes[0];
    for (Lane* L : lanes)
      if (L->load < best->load) best = L;
    r->lane = int(std::find(lanes.begin(), lanes.end(), best) - lanes.begin());
    ++best->load;
    best->incoming.push_back(r);
  }
  cv_lane.notify_all();
  return LCM_OK;
}

void lcm_serve::place_main() {
  hipSetDevice(device);
  hipStream_t stream = nullptr;
  if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) stream = nullptr;
  char* stage = nullptr;
  size_t stage_bytes = 0;
  for (;;) {
    Req* r = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_submit.wait(lk, [&] { return stop || !submitted.empty(); });
      if (stop) break;
      r = submitted.front();
      submitted.pop_front();
    }
    std::string err;
    const int rc = stream ? place(r, stage, stage_bytes, stream, err) : LCM_E_LAUNCH;
    if (rc != LCM_OK) {
      if (!stream) {
        err = "placement stream could not be created";
        unlink(r->path.c_str());
      }
      fail_req(r, rc, err);
    }
  }
  if (stage) hipHostFree(stage);
  if (stream) hipStreamDestroy(stream);
}

// ---------------------------------------------------------------------------
// lanes
// ---------------------------------------------------------------------------

bool lcm_serve::retire_ready(Lane* L, bool block) {
  bool any = false;
  while (!L->pending.empty()) {
    Launch& f = L->pending.front();
    hipError_t q = block ? hipEventSynchronize(f.end) : hipEventQuery(f.end);
    if (q == hipErrorNotReady) break;
    if (q == hipSuccess && f.timed) {
      float a = 0.f, b = 0.f;
      if (hipEventElapsedTime(&a, t0, f.start) == hipSuccess && hipEventElapsedTime(&b, t0, f.end) == hipSuccess) {
        std::lock_guard<std::mutex> lk(st_mu);
        intervals.emplace_back(a, b);
      }
    }
    const double td = now_s();
    {
      std::lock_guard<std::mutex> lk(mu);
      for (Req* r : f.done) {
        if (r->hold && q == hipSuccess) {
          held[r->rid] = Held{r->slab_off, r->slab_len, r->desc_off, r->desc_len};
        } else {
          slab.release(r->slab_off, r->slab_len);
          desc_slab.release(r->desc_off, r->desc_len);
          --n_placed;
        }
        --L->load;
        r->t_done = td;
        if (q != hipSuccess) {
          r->status = LCM_E_LAUNCH;
          r->error = std::string("fused launch: ") + hipGetErrorString(q);
        }
        r->tasks = std::vector<LcmGsTask>();
        done_q.push_back(r);
      }
    }
    {
      std::lock_guard<std::mutex> lk(st_mu);
      retired += int64_t(f.done.size());
    }
    if (!f.done.empty()) {
      cv_done.notify_all();
      cv_mem.notify_all();
    }
    L->ev_pool.push_back(f.start);
    L->ev_pool.push_back(f.end);
    L->pending.pop_front();
    any = true;
  }
  return any;
}

void lcm_serve::launch(Lane* L) {
  const int64_t t0n = now_ns();
  const int k = L->ring_i;
  LcmGsTask* h = L->h_ring[size_t(k)];
  // the requests of this launch: in admission order while their tasks fit
  size_t np = 0;
  int64_t coarse = 0, n_act = 1;
  size_t lds = 0;
  int variant = 0;
  for (; np < L->inflight.size(); ++np) {
    Req* r = L->inflight[np];
    const int64_t c = r->step_off[size_t(r->s) + 1] - r->step_off[size_t(r->s)];
    if (coarse + c + 8 > max_tasks) break;
    coarse += c;
    lds = std::max(lds, r->lds[size_t(r->s)]);
    variant = std::max(variant, r->variant[size_t(r->s)]);
    n_act = std::max(n_act, r->n_act[size_t(r->s)]);
  }
  if (np == 0) {  // one step larger than a launch's task buffer
    Req* r = L->inflight.front();
    L->inflight.erase(L->inflight.begin());
    {
      std::lock_guard<std::mutex> lk(mu);
      slab.release(r->slab_off, r->slab_len);
      desc_slab.release(r->desc_off, r->desc_len);
      if (r->has_res) res_slab.release(r->res_off, r->res_len);
      r->has_res = false;
      --n_placed;
      --L->load;
    }
    cv_mem.notify_all();
    fail_req(r, LCM_E_ARG, "a backward step of this request has more tasks than a fused launch holds");
    return;
  }
  int tile = fixed_tile;
  if (!tile)
    tile = choose_tile(t_units.data(), t_acts.data(), t_tiles.data(), int(t_tiles.size()), coarse, n_act);
  if (int64_t(LCM_REQ_TILE / tile) * coarse + 8 > max_tasks) tile = LCM_REQ_TILE;
  int64_t n = 0;
  for (size_t i = 0; i < np; ++i) {
    Req* r = L->inflight[i];
    const int64_t a = r->step_off[size_t(r->s)], c = r->step_off[size_t(r->s) + 1] - a;
    if (tile == LCM_REQ_TILE) {
      std::memcpy(h + n, r->tasks.data() + a, size_t(c) * sizeof(LcmGsTask));
      n += c;
    } else {
      n += split_tasks(r->tasks.data() + a, c, n_w_of_desc.data(), tile, h + n);
    }
  }
  const int64_t npad = (n + 7) / 8 * 8;
  for (int64_t j = n; j < npad; ++j) h[j] = LcmGsTask{-1, 0, 0, 0};
  if ((flags & LCM_SERVE_LONGEST_FIRST) && n > 1) order_longest_first(L, h, n, npad, tile);
  // requests this launch completes: their outputs follow the kernel
  Launch rec;
  L->dst.clear();
  L->src.clear();
  L->bytes.clear();
  for (size_t i = 0; i < np; ++i) {
    Req* r = L->inflight[i];
    if (r->s == r->n_steps - 1) {
      rec.done.push_back(r);
      size_t o = r->res_off;
      for (size_t f = 0; f < r->fetch_off.size(); ++f) {
        L->dst.push_back(res_host + o);
        L->src.push_back(r->base + r->fetch_off[f]);
        L->bytes.push_back(size_t(r->fetch_bytes[f]));
        o += size_t(r->fetch_bytes[f]);
      }
    }
    ++r->s;
  }
  size_t w = 0;
  for (size_t i = 0; i < L->inflight.size(); ++i) {
    Req* r = L->inflight[i];
    if (r->s < r->n_steps) L->inflight[w++] = r;
  }
  L->inflight.resize(w);
  rec.start = take_event(L);
  rec.end = take_event(L);
  int rc = LCM_OK;
  if (!rec.start || !rec.end) {
    rc = LCM_E_LAUNCH;
  } else if (npad > 0) {
    rc = lcm_gs_serve_launch(&L->ctx, dtype, desc_dev, h, L->d_ring[size_t(k)], int(npad), tile, lds, variant,
                             rec.start, L->ring_ev[size_t(k)], L->dst.data(), L->src.data(), L->bytes.data(),
                             int(L->dst.size()), rec.end);
    L->ring_used[size_t(k)] = 1;
    L->ring_i = (k + 1) % int(L->h_ring.size());
    rec.timed = true;
  } else {  // no regime of these requests is active at this step
    hipError_t e = hipEventRecord(rec.start, L->ctx.stream);
    for (size_t i = 0; e == hipSuccess && i < L->dst.size(); ++i)
      if (L->bytes[i]) e = hipMemcpyAsync(L->dst[i], L->src[i], L->bytes[i], hipMemcpyDeviceToHost, L->ctx.stream);
    if (e == hipSuccess) e = hipEventRecord(rec.end, L->ctx.stream);
    if (e != hipSuccess) rc = check_hip(e, "serve: empty step");
  }
  if (rc != LCM_OK) {
    // the stream's earlier work is drained; every request of this launch fails
    const std::string msg = std::string("fused launch failed: ") + lcm_last_error();
    hipStreamSynchronize(L->ctx.stream);
    std::vector<Req*> all(L->inflight.begin(), L->inflight.end());
    for (Req* r : rec.done) all.push_back(r);
    // requests of this lane still in flight after the failure cannot continue
    L->inflight.clear();
    {
      std::lock_guard<std::mutex> lk(mu);
      for (Req* r : all) {
        slab.release(r->slab_off, r->slab_len);
        desc_slab.release(r->desc_off, r->desc_len);
        if (r->has_res) res_slab.release(r->res_off, r->res_len);
        r->has_res = false;
        --n_placed;
        --L->load;
        r->status = LCM_E_LAUNCH;
        r->error = msg;
        r->t_done = now_s();
        done_q.push_back(r);
      }
    }
    if (rec.start) L->ev_pool.push_back(rec.start);
    if (rec.end) L->ev_pool.push_back(rec.end);
    cv_done.notify_all();
    cv_mem.notify_all();
    return;
  }
  L->pending.push_back(std::move(rec));
  std::lock_guard<std::mutex> lk(st_mu);
  ++launches;
  tasks += n;
  if (npad > 0) ++by_tile[tile == 64 ? 0 : (tile == 128 ? 1 : 2)];
  loop_ns += now_ns() - t0n;
}

void lcm_serve::lane_main(Lane* L) {
  hipSetDevice(device);
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu);
      if (L->incoming.empty() && L->inflight.empty()) {
        if (L->pending.empty()) {
          cv_lane.wait(lk, [&] { return stop || !L->incoming.empty(); });
        } else {
          // only completions to wait for: poll them, or take new arrivals
          cv_lane.wait_for(lk, std::chrono::microseconds(50), [&] { return stop || !L->incoming.empty(); });
        }
      }
      if (stop) break;
      while (!L->incoming.empty()) {
        L->inflight.push_back(L->incoming.front());
        L->incoming.pop_front();
      }
    }
    retire_ready(L, false);
    if (L->inflight.empty()) continue;
    const int k = L->ring_i;
    if (L->ring_used[size_t(k)] && hipEventQuery(L->ring_ev[size_t(k)]) == hipErrorNotReady) {
      // every task buffer is queued: the upload two launches back has not run
      std::this_thread::sleep_for(std::chrono::microseconds(20));
      continue;
    }
    launch(L);
  }
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int lcm_tile_choose(const int64_t* units, const int64_t* actions, const int32_t* tiles, int n, int64_t u,
                    int64_t a) {
  if (n < 0 || (n > 0 && (!units || !actions || !tiles))) return fail(LCM_E_ARG, "lcm_tile_choose: bad table");
  return choose_tile(units, actions, tiles, n, u, a);
}

int64_t lcm_split_tasks(const LcmGsTask* in, int64_t n, const int32_t* n_w_of_desc, int tile, LcmGsTask* out) {
  if (n < 0 || (n > 0 && (!in || !n_w_of_desc || !out)) || (tile != 64 && tile != 128))
    return fail(LCM_E_ARG, "lcm_split_tasks: bad argument");
  return split_tasks(in, n, n_w_of_desc, tile, out);
}

double lcm_serve_clock(void) { return now_s(); }

int lcm_req_inspect(const char* path, int64_t* info, int n_info) {
  if (!path || !info || n_info < 10) return fail(LCM_E_ARG, "lcm_req_inspect: bad argument");
  const int fd = open(path, O_RDONLY);
  if (fd < 0) return fail(LCM_E_ARG, std::string("cannot open request file ") + path);
  ReqMeta m;
  std::string err;
  const bool ok = read_meta(fd, m, err);
  close(fd);
  if (!ok) return fail(LCM_E_ARG, err);
  int64_t bits = 0, var = 0;
  size_t lds = 0;
  for (uint64_t mask : m.ptr_mask) bits += __builtin_popcountll(mask);
  for (int64_t v : m.variant) var = std::max(var, v);
  for (size_t b : m.lds) lds = std::max(lds, b);
  const int64_t out[10] = {m.h.n_steps,     m.h.n_desc, m.h.n_task, m.h.n_fetch, m.h.image_bytes,
                           m.h.total_bytes, m.fetch_total, bits,    var,         int64_t(lds)};
  std::memcpy(info, out, sizeof(out));
  return LCM_OK;
}

int lcm_serve_create(lcm_serve** out, int dtype, size_t budget_bytes, int max_desc, int max_tasks,
                     int max_inflight, size_t result_bytes, int lanes, int queue_depth, int placers,
                     int fixed_tile, const int64_t* table_units, const int64_t* table_actions,
                     const int32_t* table_tiles, int n_table, int flags) {
  if (!out) return fail(LCM_E_ARG, "lcm_serve_create: null out");
  if (flags & ~LCM_SERVE_LONGEST_FIRST) return fail(LCM_E_ARG, "lcm_serve_create: unknown flags");
  *out = nullptr;
  if (dtype != LCM_DTYPE_F32 && dtype != LCM_DTYPE_F64) return fail(LCM_E_ARG, "lcm_serve_create: dtype");
  if (max_desc <= 0 || max_tasks < 64 || max_inflight <= 0 || lanes <= 0 || lanes > 16 || queue_depth <= 0 ||
      queue_depth > 64 || placers <= 0 || placers > 16 || budget_bytes < 256)
    return fail(LCM_E_ARG, "lcm_serve_create: sizes out of range");
  if (fixed_tile != 0 && fixed_tile != 64 && fixed_tile != 128 && fixed_tile != 256)
    return fail(LCM_E_ARG, "lcm_serve_create: tile must be 0, 64, 128 or 256");
  if (n_table < 0 || (n_table > 0 && (!table_units || !table_actions || !table_tiles)))
    return fail(LCM_E_ARG, "lcm_serve_create: bad tile table");
  for (int i = 0; i < n_table; ++i)
    if (table_tiles[i] != 64 && table_tiles[i] != 128 && table_tiles[i] != 256)
      return fail(LCM_E_ARG, "lcm_serve_create: tile table holds an unknown tile");
  lcm_serve* s = new lcm_serve();
  s->dtype = dtype;
  s->max_tasks = max_tasks / 8 * 8;
  s->max_inflight = max_inflight;
  s->fixed_tile = fixed_tile;
  s->flags = flags;
  s->t_units.assign(table_units, table_units + n_table);
  s->t_acts.assign(table_actions, table_actions + n_table);
  s->t_tiles.assign(table_tiles, table_tiles + n_table);
  auto cleanup = [&](int code) {
    lcm_serve_destroy(s);
    return code;
  };
  hipError_t e = hipGetDevice(&s->device);
  if (e != hipSuccess) return cleanup(check_hip(e, "hipGetDevice"));
  s->slab.init(budget_bytes, 256);
  s->desc_slab.init(size_t(max_desc) * sizeof(LcmGsDesc), sizeof(LcmGsDesc));
  s->res_slab.init(std::max<size_t>(result_bytes, 256), 256);
  s->n_w_of_desc.assign(size_t(max_desc), 0);
  s->act_of_desc.assign(size_t(max_desc), 0);
  if ((e = hipMalloc(reinterpret_cast<void**>(&s->slab_dev), s->slab.size())) != hipSuccess)
    return cleanup(check_hip(e, "hipMalloc (server memory)"));
  if ((e = hipMalloc(reinterpret_cast<void**>(&s->desc_dev), s->desc_slab.size())) != hipSuccess)
    return cleanup(check_hip(e, "hipMalloc (descriptor pool)"));
  if ((e = hipHostMalloc(reinterpret_cast<void**>(&s->res_host), s->res_slab.size(), hipHostMallocDefault)) !=
      hipSuccess)
    return cleanup(check_hip(e, "hipHostMalloc (results)"));
  if ((e = hipEventCreate(&s->t0)) != hipSuccess) return cleanup(check_hip(e, "hipEventCreate"));
  for (int i = 0; i < lanes; ++i) {
    Lane* L = new Lane();
    s->lanes.push_back(L);
    L->ctx.device = s->device;
    if ((e = hipStreamCreateWithFlags(&L->ctx.stream, hipStreamNonBlocking)) != hipSuccess) {
      L->ctx.stream = nullptr;
      return cleanup(check_hip(e, "hipStreamCreateWithFlags"));
    }
    for (int j = 0; j < queue_depth; ++j) {
      void *hp = nullptr, *dp = nullptr;
      hipEvent_t ev = nullptr;
      const size_t nb = size_t(s->max_tasks) * sizeof(LcmGsTask);
      if ((e = hipHostMalloc(&hp, nb, hipHostMallocDefault)) != hipSuccess)
        return cleanup(check_hip(e, "hipHostMalloc (task ring)"));
      L->h_ring.push_back(static_cast<LcmGsTask*>(hp));
      if ((e = hipMalloc(&dp, nb)) != hipSuccess) return cleanup(check_hip(e, "hipMalloc (task ring)"));
      L->d_ring.push_back(static_cast<LcmGsTask*>(dp));
      if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess)
        return cleanup(check_hip(e, "hipEventCreate (task ring)"));
      L->ring_ev.push_back(ev);
      L->ring_used.push_back(0);
    }
  }
  if ((e = hipEventRecord(s->t0, s->lanes[0]->ctx.stream)) != hipSuccess)
    return cleanup(check_hip(e, "hipEventRecord"));
  if ((e = hipStreamSynchronize(s->lanes[0]->ctx.stream)) != hipSuccess)
    return cleanup(check_hip(e, "hipStreamSynchronize"));
  for (Lane* L : s->lanes) L->th = std::thread([s, L] { s->lane_main(L); });
  for (int i = 0; i < placers; ++i) s->placers.emplace_back([s] { s->place_main(); });
  *out = s;
  return LCM_OK;
}

int lcm_serve_submit(lcm_serve* s, int64_t rid, const char* path) {
  if (!s || !path) return fail(LCM_E_ARG, "lcm_serve_submit: null argument");
  Req* r = new Req();
  r->rid = rid;
  r->path = path;
  r->t_submit = now_s();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
      delete r;
      return fail(LCM_E_ARG, "lcm_serve_submit: server closed");
    }
    s->submitted.push_back(r);
  }
  s->cv_submit.notify_one();
  return LCM_OK;
}

int lcm_serve_wait(lcm_serve* s, int64_t* rids, int32_t* status, int max_n, int timeout_ms) {
  if (!s || !rids || !status || max_n <= 0) return fail(LCM_E_ARG, "lcm_serve_wait: bad argument");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait_for(lk, std::chrono::milliseconds(std::max(timeout_ms, 0)),
                      [&] { return s->stop || !s->done_q.empty(); });
  int n = 0;
  while (n < max_n && !s->done_q.empty()) {
    Req* r = s->done_q.front();
    s->done_q.pop_front();
    s->finished[r->rid] = r;
    rids[n] = r->rid;
    status[n] = r->status;
    ++n;
  }
  return n;
}

int lcm_serve_fetch(lcm_serve* s, int64_t rid, void* dst, size_t nbytes, double* times, char* err,
                    size_t err_len) {
  if (!s) return fail(LCM_E_ARG, "lcm_serve_fetch: null server");
  Req* r = nullptr;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->finished.find(rid);
    if (it == s->finished.end()) return fail(LCM_E_ARG, "lcm_serve_fetch: no completed request of this id");
    r = it->second;
    s->finished.erase(it);
  }
  int rc = r->status;
  if (rc == LCM_OK && r->has_res) {
    size_t total = 0;
    for (int64_t b : r->fetch_bytes) total += size_t(b);
    if (dst && nbytes >= total) {
      std::memcpy(dst, s->res_host + r->res_off, total);
    } else if (dst) {
      rc = fail(LCM_E_ARG, "lcm_serve_fetch: destination smaller than the request's outputs");
    }
  }
  if (times) {
    times[0] = r->t_submit;
    times[1] = r->t_placed;
    times[2] = r->t_done;
  }
  if (err && err_len) {
    const size_t n = std::min(err_len - 1, r->error.size());
    std::memcpy(err, r->error.data(), n);
    err[n] = 0;
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (r->has_res) s->res_slab.release(r->res_off, r->res_len);
  }
  s->cv_mem.notify_all();
  delete r;
  return rc;
}

int lcm_serve_held(lcm_serve* s, int64_t rid, uint64_t* base, uint64_t* descs) {
  if (!s || !base || !descs) return fail(LCM_E_ARG, "lcm_serve_held: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->held.find(rid);
  if (it == s->held.end()) return fail(LCM_E_ARG, "lcm_serve_held: no held request of this id");
  *base = reinterpret_cast<uint64_t>(s->slab_dev + it->second.slab_off);
  *descs = reinterpret_cast<uint64_t>(reinterpret_cast<char*>(s->desc_dev) + it->second.desc_off);
  return LCM_OK;
}

int lcm_serve_release(lcm_serve* s, int64_t rid) {
  if (!s) return fail(LCM_E_ARG, "lcm_serve_release: null server");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->held.find(rid);
    if (it == s->held.end()) return fail(LCM_E_ARG, "lcm_serve_release: no held request of this id");
    s->slab.release(it->second.slab_off, it->second.slab_len);
    s->desc_slab.release(it->second.desc_off, it->second.desc_len);
    --s->n_placed;
    s->held.erase(it);
  }
  s->cv_mem.notify_all();
  return LCM_OK;
}

int lcm_serve_stats(const lcm_serve* s, int64_t* out, int n) {
  if (!s || !out || n < 12) return fail(LCM_E_ARG, "lcm_serve_stats: bad argument");
  lcm_serve* m = const_cast<lcm_serve*>(s);
  int64_t v[12];
  {
    std::lock_guard<std::mutex> lk(m->st_mu);
    v[0] = m->launches;
    v[1] = m->tasks;
    v[2] = m->by_tile[0];
    v[3] = m->by_tile[1];
    v[4] = m->by_tile[2];
    v[5] = m->loop_ns;
    v[6] = m->place_ns;
    v[7] = m->placed;
    v[8] = m->retired;
  }
  {
    std::lock_guard<std::mutex> lk(m->mu);
    v[9] = int64_t(m->slab.used());
    v[10] = int64_t(m->slab.size());
    v[11] = int64_t(m->slab.peak());
  }
  std::memcpy(out, v, sizeof(v));
  return LCM_OK;
}

int lcm_serve_busy_ms(lcm_serve* s, double* ms) {
  if (!s || !ms) return fail(LCM_E_ARG, "lcm_serve_busy_ms: null argument");
  std::vector<std::pair<float, float>> iv;
  {
    std::lock_guard<std::mutex> lk(s->st_mu);
    iv = s->intervals;
  }
  std::sort(iv.begin(), iv.end());
  double busy = 0.0, end = -1.0;
  for (auto& p : iv) {
    if (p.first > end) {
      busy += p.second - p.first;
      end = p.second;
    } else if (p.second > end) {
      busy += p.second - end;
      end = p.second;
    }
  }
  *ms = busy;
  return LCM_OK;
}

int lcm_serve_destroy(lcm_serve* s) {
  if (!s) return LCM_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_submit.notify_all();
  s->cv_mem.notify_all();
  s->cv_lane.notify_all();
  s->cv_done.notify_all();
  for (auto& t : s->placers)
    if (t.joinable()) t.join();
  for (Lane* L : s->lanes)
    if (L->th.joinable()) L->th.join();
  std::vector<Req*> dead(s->submitted.begin(), s->submitted.end());
  for (Req* r : dead) unlink(r->path.c_str());
  for (Lane* L : s->lanes) {
    if (L->ctx.stream) hipStreamSynchronize(L->ctx.stream);
    for (Launch& f : L->pending) {
      dead.insert(dead.end(), f.done.begin(), f.done.end());
      L->ev_pool.push_back(f.start);
      L->ev_pool.push_back(f.end);
    }
    dead.insert(dead.end(), L->inflight.begin(), L->inflight.end());
    dead.insert(dead.end(), L->incoming.begin(), L->incoming.end());
    for (hipEvent_t e : L->ev_pool)
      if (e) hipEventDestroy(e);
    for (hipEvent_t e : L->ring_ev)
      if (e) hipEventDestroy(e);
    for (LcmGsTask* p : L->h_ring) hipHostFree(p);
    for (LcmGsTask* p : L->d_ring) hipFree(p);
    if (L->ctx.stream) hipStreamDestroy(L->ctx.stream);
    delete L;
  }
  for (Req* r : s->done_q) dead.push_back(r);
  for (auto& kv : s->finished) dead.push_back(kv.second);
  for (Req* r : dead) delete r;
  if (s->t0) hipEventDestroy(s->t0);
  if (s->slab_dev) hipFree(s->slab_dev);
  if (s->desc_dev) hipFree(s->desc_dev);
  if (s->res_host) hipHostFree(s->res_host);
  delete s;
  return LCM_OK;
}

}  // extern "C"
