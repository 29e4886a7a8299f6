// Serving loop over the batch engine (include/ivqm_serve.h).
//
// One worker thread per server owns an Engine (its stream, device buffers and
// ring of 8 result slots) and a copy stream. Each pass of the loop
//   1. collects every finished batch (ring order) and hands each pair's
//      result to its job; a job with all frames done is published;
//   2. admits queued jobs in arrival order while their staged test frames
//      fit the memory budget (first-fit arena inside one device allocation)
//      and starts their host -> device copies on the copy stream;
//   3. fills free ring slots with batches of frame pairs whose inputs are on
//      the device, oldest job first -- pairs of different jobs, sizes, bit
//      depths and metric sets go into one batch (one launch per phase);
//   4. sleeps briefly when none of this made progress.
// Results of a pair do not depend on what else is in its batch: every kernel
// reads only its own pair's table entry and writes its own outputs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ivqm_serve.h"
#include "ivqm_engine.h"

using namespace ivqm;

#define HIPCHK(expr)                                   \
  do {                                                 \
    hipError_t _e = (expr);                            \
    if (_e != hipSuccess) return hip_error(_e, #expr); \
  } while (0)
#define RCCHK(expr)      \
  do {                   \
    int _rc = (expr);    \
    if (_rc) return _rc; \
  } while (0)

namespace {

using Clock = std::chrono::steady_clock;

double us_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::micro>(b - a).count();
}

// First-fit allocator over one device allocation; offsets in bytes, 256-byte granules.
class Arena {
 public:
  void init(uint64_t cap) {
    cap_ = cap;
    free_.clear();
    free_[0] = cap;
    used_ = peak_ = 0;
  }
  bool alloc(uint64_t bytes, uint64_t* off) {
    bytes = (bytes + 255) & ~uint64_t(255);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const uint64_t rest = it->second - bytes, at = it->first + bytes;
        free_.erase(it);
        if (rest) free_[at] = rest;
        used_ += bytes;
        peak_ = std::max(peak_, used_);
        return true;
      }
    }
    return false;
  }
  void release(uint64_t off, uint64_t bytes) {
    bytes = (bytes + 255) & ~uint64_t(255);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
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
  uint64_t used() const { return used_; }
  uint64_t peak() const { return peak_; }
  uint64_t cap() const { return cap_; }

 private:
  uint64_t cap_ = 0, used_ = 0, peak_ = 0;
  std::map<uint64_t, uint64_t> free_;  // offset -> size
};

struct Format {
  int w = 0, h = 0, bd = 0, cs = 0, bps = 2, csy = 0;  // chroma column / row shifts
  int plane_w(int k) const { return k ? (w >> cs) : w; }
  int plane_h(int k) const { return k ? (h >> csy) : h; }
  // staged copy of one picture: three planes with 64-sample pitches
  uint64_t staged_bytes() const {
    uint64_t b = 0;
    for (int k = 0; k < 3; k++) b += (uint64_t)Engine::scp_pitch(plane_w(k)) * plane_h(k) * bps;
    return b;
  }
};

struct Sequence {
  Format f;
  void* dev = nullptr;
  uint64_t bytes = 0;
  std::vector<PicDev> frames;
  int refs = 0;  // submitted, unfinished jobs using it
};

struct Job {
  ivqm_job d;
  std::vector<const void*> planes;
  std::vector<int32_t> strides;
  std::vector<const void*> mask_planes;  // empty: no mask
  std::vector<int32_t> mask_strides;
  std::vector<const void*> mask_dev;     // per frame, on the device
  std::vector<int32_t> mask_dev_stride;
  std::shared_ptr<Sequence> seq;
  Clock::time_point t_submit, t_admit, t_ready, t_batched;
  bool batched = false;  // a frame of the job went into a batch
  bool staged = false;
  uint64_t arena_off = 0, arena_bytes = 0;
  hipEvent_t uploaded = nullptr;
  bool inputs_ready = false;
  std::vector<PicDev> test;  // per frame, on the device
  int next_frame = 0, done_frames = 0;
  int status = IVQM_OK;
  ivqm_job_result result;
};

struct InFlight {
  bool busy = false;
  std::vector<std::pair<Job*, int>> pairs;  // (job, frame)
};

bool aligned_picture(const PicDev& p, const Format& f) {
  for (int k = 0; k < 3; k++) {
    const bool full = (k == 0) || !f.cs;
    if (full && ((((uintptr_t)p.p[k]) % (size_t)(4 * f.bps)) || (p.stride[k] % 4))) return false;
  }
  return true;
}

}  // namespace

struct ivqm_server {
  ivqm_server_config cfg;
  std::unique_ptr<Engine> eng;
  hipStream_t copy_stream = nullptr;
  void* arena_base = nullptr;
  Arena arena;
  int device = 0;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::unique_ptr<Job>> submitted;  // not yet seen by the worker
  std::deque<ivqm_job_result> done;
  std::map<int32_t, std::shared_ptr<Sequence>> seqs;
  int32_t next_seq = 1;
  uint64_t n_submitted = 0, n_completed = 0, n_failed = 0, n_pairs = 0, n_batches = 0;
  uint64_t resident = 0;
  double max_fill = 0;
  bool stop = false;
  std::thread worker;

  // worker-only state
  std::deque<std::unique_ptr<Job>> pending;   // waiting for memory
  std::deque<std::unique_ptr<Job>> admitted;  // staged or in progress, arrival order
  InFlight slots[kRing];
  std::deque<int> slot_order;  // enqueued slots, oldest first

  void run();
  int step(bool* progress);
  void finish_job(Job* j);
  int admit(Job* j);
};

namespace {

int validate_job(ivqm_server* s, const ivqm_job* j, std::shared_ptr<Sequence>* seq) {
  if (!j) return set_error(IVQM_EINVAL, "job is NULL");
  if (j->num_frames < 1 || j->num_frames > IVQM_MAX_JOB_FRAMES) return set_error(IVQM_EINVAL, "num_frames must be 1..8");
  if (!j->planes) return set_error(IVQM_EINVAL, "planes is NULL");
  const uint32_t m = j->metrics & IVQM_M_ALL;
  if (!m || (j->metrics & ~(uint32_t)IVQM_M_ALL)) return set_error(IVQM_EINVAL, "metrics must be a non-empty IVQM_M_* set");
  auto it = s->seqs.find(j->seq_id);
  if (it == s->seqs.end()) return set_error(IVQM_EINVAL, "unknown seq_id");
  const Sequence& q = *it->second;
  if (j->first_frame < 0 || j->first_frame + j->num_frames > (int)q.frames.size())
    return set_error(IVQM_EINVAL, "frame range outside the sequence");
  for (int i = 0; i < 3 * j->num_frames; i++) {
    if (!j->planes[i]) return set_error(IVQM_EINVAL, "plane pointer is NULL");
    if (j->strides && j->strides[i] < q.f.plane_w(i % 3)) return set_error(IVQM_EINVAL, "stride smaller than plane width");
  }
  if (j->mask_planes) {
    if (m & (IVQM_M_SSIM | IVQM_M_MSSSIM | IVQM_M_IVSSIM | IVQM_M_IVMSSSIM))
      return set_error(IVQM_EINVAL, "structural similarity metrics cannot be combined with a mask");
    if (j->mask_bit_depth < 8 || j->mask_bit_depth > 14) return set_error(IVQM_EINVAL, "mask_bit_depth must be 8..14");
    if (j->mask_bytes_per_sample != 1 && j->mask_bytes_per_sample != 2)
      return set_error(IVQM_EINVAL, "mask_bytes_per_sample must be 1 or 2");
    if (j->mask_bytes_per_sample == 1 && j->mask_bit_depth != 8)
      return set_error(IVQM_EINVAL, "8-bit mask samples require mask_bit_depth 8");
    for (int f = 0; f < j->num_frames; f++) {
      if (!j->mask_planes[f]) return set_error(IVQM_EINVAL, "mask plane pointer is NULL");
      if (j->mask_strides && j->mask_strides[f] < q.f.w) return set_error(IVQM_EINVAL, "mask stride smaller than the width");
    }
  }
  if (j->planes_on_device) {
    for (int f = 0; f < j->num_frames; f++) {
      PicDev p;
      for (int k = 0; k < 3; k++) {
        p.p[k] = j->planes[3 * f + k];
        p.stride[k] = j->strides ? j->strides[3 * f + k] : q.f.plane_w(k);
      }
      if (!aligned_picture(p, q.f))
        return set_error(IVQM_EINVAL, "device planes need 4-sample aligned rows (pointer and stride)");
    }
  }
  *seq = it->second;
  return IVQM_OK;
}

}  // namespace

int ivqm_server::admit(Job* j) {
  const Format& f = j->seq->f;
  j->test.resize(j->d.num_frames);
  if (j->d.planes_on_device) {
    for (int i = 0; i < j->d.num_frames; i++) {
      PicDev& p = j->test[i];
      for (int k = 0; k < 3; k++) {
        p.p[k] = j->planes[3 * i + k];
        p.stride[k] = j->strides.empty() ? f.plane_w(k) : j->strides[3 * i + k];
      }
      p.cs = f.cs;
      p.csy = f.csy;
      p.bps = f.bps;
    }
    j->t_ready = j->t_admit;
    if (!j->mask_planes.empty()) {
      j->mask_dev = j->mask_planes;
      j->mask_dev_stride.resize(j->d.num_frames);
      for (int i = 0; i < j->d.num_frames; i++)
        j->mask_dev_stride[i] = j->mask_strides.empty() ? f.w : j->mask_strides[i];
    }
    j->inputs_ready = true;
    return IVQM_OK;
  }
  // host planes: copy into the arena on the copy stream
  char* base = (char*)arena_base + j->arena_off;
  uint64_t off = 0;
  for (int i = 0; i < j->d.num_frames; i++) {
    PicDev& p = j->test[i];
    for (int k = 0; k < 3; k++) {
      const int wp = f.plane_w(k), hp = f.plane_h(k), pitch = Engine::scp_pitch(wp);
      const int sstride = j->strides.empty() ? wp : j->strides[3 * i + k];
      p.p[k] = base + off;
      p.stride[k] = pitch;
      HIPCHK(hipMemcpy2DAsync(base + off, (size_t)pitch * f.bps, j->planes[3 * i + k], (size_t)sstride * f.bps,
                              (size_t)wp * f.bps, hp, hipMemcpyHostToDevice, copy_stream));
      off += (uint64_t)pitch * hp * f.bps;
    }
    p.cs = f.cs;
    p.csy = f.csy;
    p.bps = f.bps;
  }
  if (!j->mask_planes.empty()) {  // the masks after the test frames, 64-sample pitch
    const int mbps = j->d.mask_bytes_per_sample, pitch = Engine::scp_pitch(f.w);
    j->mask_dev.resize(j->d.num_frames);
    j->mask_dev_stride.assign(j->d.num_frames, pitch);
    for (int i = 0; i < j->d.num_frames; i++) {
      const int sstride = j->mask_strides.empty() ? f.w : j->mask_strides[i];
      j->mask_dev[i] = base + off;
      HIPCHK(hipMemcpy2DAsync(base + off, (size_t)pitch * mbps, j->mask_planes[i], (size_t)sstride * mbps,
                              (size_t)f.w * mbps, f.h, hipMemcpyHostToDevice, copy_stream));
      off += (uint64_t)pitch * f.h * mbps;
    }
  }
  HIPCHK(hipEventCreateWithFlags(&j->uploaded, hipEventDisableTiming));
  HIPCHK(hipEventRecord(j->uploaded, copy_stream));
  return IVQM_OK;
}

void ivqm_server::finish_job(Job* j) {
  ivqm_job_result& r = j->result;
  r.tag = j->d.tag;
  r.status = j->status;
  r.num_frames = j->d.num_frames;
  r.latency_us = us_between(j->t_submit, Clock::now());
  const Clock::time_point now = Clock::now();
  r.queue_us = j->staged || j->d.planes_on_device ? us_between(j->t_submit, j->t_admit) : r.latency_us;
  r.stage_us = r.wait_us = r.device_us = 0;
  if (j->batched) {
    r.stage_us = us_between(j->t_admit, j->t_ready);
    r.wait_us = us_between(j->t_ready, j->t_batched);
    r.device_us = us_between(j->t_batched, now);
  }
  if (j->uploaded) (void)hipEventDestroy(j->uploaded);
  j->uploaded = nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if (j->staged) arena.release(j->arena_off, j->arena_bytes);
  j->staged = false;
  j->seq->refs--;
  done.push_back(r);
  n_completed++;
  if (j->status) n_failed++;
  cv_done.notify_all();
}

// One pass of the loop; *progress tells whether anything moved.
int ivqm_server::step(bool* progress) {
  *progress = false;
  {
    std::lock_guard<std::mutex> lk(mu);
    while (!submitted.empty()) {
      pending.push_back(std::move(submitted.front()));
      submitted.pop_front();
      *progress = true;
    }
  }
  // 1. finished batches, oldest first
  while (!slot_order.empty()) {
    const int sl = slot_order.front();
    const int q = eng->query(sl);
    if (q < 0) return q;
    if (q == 0) break;
    InFlight& F = slots[sl];
    std::vector<ivqm_frame_result> res(F.pairs.size());
    const int rc = eng->collect(sl, res.data());
    for (size_t k = 0; k < F.pairs.size(); k++) {
      Job* j = F.pairs[k].first;
      if (rc) j->status = rc;
      else j->result.frames[F.pairs[k].second] = res[k];
      j->done_frames++;
    }
    F.busy = false;
    F.pairs.clear();
    slot_order.pop_front();
    {
      std::lock_guard<std::mutex> lk(mu);
      n_pairs += res.size();
    }
    *progress = true;
    // publish finished jobs (keep arrival order among the rest)
    for (auto it = admitted.begin(); it != admitted.end();) {
      Job* j = it->get();
      if (j->done_frames == j->d.num_frames) {
        finish_job(j);
        it = admitted.erase(it);
      } else {
        ++it;
      }
    }
  }
  // 2. admission in arrival order
  while (!pending.empty()) {
    Job* j = pending.front().get();
    if (!j->d.planes_on_device) {
      j->arena_bytes = j->seq->f.staged_bytes() * j->d.num_frames;
      if (!j->mask_planes.empty())
        j->arena_bytes += (uint64_t)Engine::scp_pitch(j->seq->f.w) * j->seq->f.h * j->d.mask_bytes_per_sample *
                          j->d.num_frames;
      if (j->arena_bytes > arena.cap()) {
        j->status = set_error(IVQM_ENOMEM, "job larger than the memory budget");
        j->t_admit = Clock::now();
        finish_job(j);
        pending.pop_front();
        continue;
      }
      bool fits;
      {
        std::lock_guard<std::mutex> lk(mu);
        fits = arena.alloc(j->arena_bytes, &j->arena_off);
      }
      if (!fits) break;
      j->staged = true;
    }
    j->t_admit = Clock::now();
    const int rc = admit(j);
    if (rc) {
      j->status = rc;
      finish_job(j);
      pending.pop_front();
      continue;
    }
    admitted.push_back(std::move(pending.front()));
    pending.pop_front();
    *progress = true;
  }
  // 3. batches into free ring slots
  for (;;) {
    if ((int)slot_order.size() >= cfg.max_inflight) break;
    int sl = -1;
    for (int k = 0; k < kRing; k++)
      if (!slots[k].busy) { sl = k; break; }
    if (sl < 0) break;
    std::vector<PairSpec> specs;
    InFlight& F = slots[sl];
    for (auto& up : admitted) {
      Job* j = up.get();
      if (j->next_frame >= j->d.num_frames) continue;
      if (!j->inputs_ready) {
        const hipError_t e = hipEventQuery(j->uploaded);
        if (e == hipErrorNotReady) continue;
        if (e != hipSuccess) return hip_error(e, "hipEventQuery");
        j->inputs_ready = true;
        j->t_ready = Clock::now();
      }
      if (!j->batched) {
        j->batched = true;
        j->t_batched = Clock::now();
      }
      while (j->next_frame < j->d.num_frames && (int)specs.size() < cfg.max_batch_pairs) {
        const int f = j->next_frame++;
        PairSpec s;
        s.p0 = j->test[f];
        s.p1 = j->seq->frames[j->d.first_frame + f];
        s.w = j->seq->f.w;
        s.h = j->seq->f.h;
        s.bd = j->seq->f.bd;
        s.metrics = j->d.metrics & IVQM_M_ALL;
        if (!j->mask_dev.empty()) {
          s.msk = j->mask_dev[f];
          s.msk_stride = j->mask_dev_stride[f];
          s.msk_bps = j->d.mask_bytes_per_sample;
          s.msk_bd = j->d.mask_bit_depth;
        }
        specs.push_back(s);
        F.pairs.emplace_back(j, f);
      }
      if ((int)specs.size() >= cfg.max_batch_pairs) break;
    }
    if (specs.empty()) break;
    const int rc = eng->enqueue(specs.data(), (int)specs.size(), 0, sl, false);
    if (rc) {
      for (auto& pr : F.pairs) {
        pr.first->status = rc;
        pr.first->done_frames++;
      }
      F.pairs.clear();
      for (auto it = admitted.begin(); it != admitted.end();) {
        Job* j = it->get();
        if (j->done_frames == j->d.num_frames) {
          finish_job(j);
          it = admitted.erase(it);
        } else {
          ++it;
        }
      }
      *progress = true;
      continue;
    }
    F.busy = true;
    slot_order.push_back(sl);
    {
      std::lock_guard<std::mutex> lk(mu);
      n_batches++;
      max_fill = std::max(max_fill, (double)specs.size());
    }
    *progress = true;
  }
  return IVQM_OK;
}

void ivqm_server::run() {
  (void)hipSetDevice(device);
  for (;;) {
    bool progress = false;
    int rc = IVQM_OK;
    try {
      rc = step(&progress);
    } catch (const std::exception& e) {
      rc = set_error(IVQM_EINVAL, std::string("internal error: ") + e.what());
    }
    if (rc) {
      // a device error fails every job the worker holds
      for (auto& up : admitted) { up->status = rc; finish_job(up.get()); }
      admitted.clear();
      for (auto& up : pending) { up->status = rc; finish_job(up.get()); }
      pending.clear();
      for (auto& F : slots) { F.busy = false; F.pairs.clear(); }
      slot_order.clear();
    }
    std::unique_lock<std::mutex> lk(mu);
    const bool idle = submitted.empty() && pending.empty() && admitted.empty() && slot_order.empty();
    if (stop && idle) break;
    if (!progress) {
      if (idle) cv_work.wait_for(lk, std::chrono::milliseconds(5));
      else cv_work.wait_for(lk, std::chrono::microseconds(20));
    }
  }
}

extern "C" {

int ivqm_server_config_init(ivqm_server_config* cfg) {
  if (!cfg) return set_error(IVQM_EINVAL, "cfg is NULL");
  std::memset(cfg, 0, sizeof(*cfg));
  RCCHK(ivqm_config_init(&cfg->params));
  cfg->memory_budget = 64ull << 30;
  cfg->max_batch_pairs = 16;
  cfg->max_inflight = 4;
  return IVQM_OK;
}

int ivqm_server_create(const ivqm_server_config* cfg_in, ivqm_server** out) {
  try {
    if (!cfg_in || !out) return set_error(IVQM_EINVAL, "NULL argument");
    *out = nullptr;
    ivqm_server_config cfg = *cfg_in;
    if (!cfg.memory_budget) cfg.memory_budget = 64ull << 30;
    if (!cfg.max_batch_pairs) cfg.max_batch_pairs = 16;
    if (!cfg.max_inflight) cfg.max_inflight = 4;
    if (cfg.max_batch_pairs < 1 || cfg.max_batch_pairs > 256) return set_error(IVQM_EINVAL, "max_batch_pairs must be 1..256");
    if (cfg.max_inflight < 1 || cfg.max_inflight > kRing) return set_error(IVQM_EINVAL, "max_inflight must be 1..8");
    ivqm_config p = cfg.params;
    p.width = p.height = 64;  // per-pair sizes come from the sequences
    p.bit_depth = 10;
    if (p.search_range < 1 || p.search_range > 4) return set_error(IVQM_EINVAL, "search_range must be 1..4");
    RCCHK(validate_ssim_config(p));
    if (p.search_mode_cap < 0 || p.search_mode_cap > 2) return set_error(IVQM_EINVAL, "search_mode_cap must be 0..2");
    std::unique_ptr<ivqm_server> s(new ivqm_server());
    s->cfg = cfg;
    s->cfg.params = p;
    HIPCHK(hipGetDevice(&s->device));
    s->eng.reset(new Engine(p));
    RCCHK(s->eng->init());
    HIPCHK(hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking));
    const hipError_t e = hipMalloc(&s->arena_base, cfg.memory_budget);
    if (e != hipSuccess) {
      (void)hipStreamDestroy(s->copy_stream);
      return e == hipErrorOutOfMemory ? set_error(IVQM_ENOMEM, "memory budget does not fit the device")
                                      : hip_error(e, "hipMalloc");
    }
    s->arena.init(cfg.memory_budget);
    ivqm_server* raw = s.get();
    s->worker = std::thread([raw] { raw->run(); });
    *out = s.release();
    return IVQM_OK;
  } catch (const std::bad_alloc&) {
    return set_error(IVQM_ENOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return set_error(IVQM_EINVAL, std::string("internal error: ") + e.what());
  }
}

int ivqm_server_destroy(ivqm_server* s) {
  if (!s) return set_error(IVQM_EINVAL, "server is NULL");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  (void)s->eng->sync();
  s->eng.reset();
  for (auto& kv : s->seqs)
    if (kv.second->dev) (void)hipFree(kv.second->dev);
  if (s->arena_base) (void)hipFree(s->arena_base);
  if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
  delete s;
  return IVQM_OK;
}

int ivqm_server_add_sequence(ivqm_server* s, const ivqm_sequence_desc* q, int32_t* seq_id) {
  try {
    if (!s || !q || !seq_id) return set_error(IVQM_EINVAL, "NULL argument");
    if (q->width < 8 || q->height < 8) return set_error(IVQM_EINVAL, "picture must be at least 8x8");
    if (q->bit_depth < 8 || q->bit_depth > 16) return set_error(IVQM_EINVAL, "bit_depth must be 8..16");
    if (q->bytes_per_sample != 1 && q->bytes_per_sample != 2) return set_error(IVQM_EINVAL, "bytes_per_sample must be 1 or 2");
    if (q->bytes_per_sample == 1 && q->bit_depth != 8) return set_error(IVQM_EINVAL, "8-bit samples require bit_depth 8");
    if (q->chroma_420 < 0 || q->chroma_420 > 2) return set_error(IVQM_EINVAL, "chroma_420 must be 0, 1 or 2");
    if (q->chroma_420 == 1 && ((q->width & 1) || (q->height & 1))) return set_error(IVQM_EINVAL, "4:2:0 requires even sizes");
    if (q->chroma_420 == 2 && (q->width & 1)) return set_error(IVQM_EINVAL, "4:2:2 requires an even width");
    if (q->num_frames < 1 || !q->planes) return set_error(IVQM_EINVAL, "no frames");
    auto seq = std::make_shared<Sequence>();
    Format& f = seq->f;
    f.w = q->width;
    f.h = q->height;
    f.bd = q->bit_depth;
    f.cs = q->chroma_420 ? 1 : 0;
    f.csy = q->chroma_420 == 1 ? 1 : 0;
    f.bps = q->bytes_per_sample;
    for (int i = 0; i < 3 * q->num_frames; i++) {
      if (!q->planes[i]) return set_error(IVQM_EINVAL, "plane pointer is NULL");
      if (q->strides && q->strides[i] < f.plane_w(i % 3)) return set_error(IVQM_EINVAL, "stride smaller than plane width");
    }
    seq->bytes = f.staged_bytes() * q->num_frames;
    HIPCHK(hipSetDevice(s->device));
    const hipError_t e = hipMalloc(&seq->dev, seq->bytes);
    if (e != hipSuccess)
      return e == hipErrorOutOfMemory ? set_error(IVQM_ENOMEM, "sequence does not fit the device") : hip_error(e, "hipMalloc");
    char* base = (char*)seq->dev;
    uint64_t off = 0;
    for (int i = 0; i < q->num_frames; i++) {
      PicDev p;
      for (int k = 0; k < 3; k++) {
        const int wp = f.plane_w(k), hp = f.plane_h(k), pitch = Engine::scp_pitch(wp);
        const int ss = q->strides ? q->strides[3 * i + k] : wp;
        p.p[k] = base + off;
        p.stride[k] = pitch;
        const hipError_t ce = hipMemcpy2D(base + off, (size_t)pitch * f.bps, q->planes[3 * i + k], (size_t)ss * f.bps,
                                          (size_t)wp * f.bps, hp, hipMemcpyDefault);
        if (ce != hipSuccess) {
          (void)hipFree(seq->dev);
          return hip_error(ce, "hipMemcpy2D");
        }
        off += (uint64_t)pitch * hp * f.bps;
      }
      p.cs = f.cs;
      p.csy = f.csy;
      p.bps = f.bps;
      seq->frames.push_back(p);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    *seq_id = s->next_seq++;
    s->seqs[*seq_id] = seq;
    s->resident += seq->bytes;
    return IVQM_OK;
  } catch (const std::bad_alloc&) {
    return set_error(IVQM_ENOMEM, "host allocation failed");
  }
}

int ivqm_server_remove_sequence(ivqm_server* s, int32_t seq_id) {
  if (!s) return set_error(IVQM_EINVAL, "server is NULL");
  std::shared_ptr<Sequence> seq;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->seqs.find(seq_id);
    if (it == s->seqs.end()) return set_error(IVQM_EINVAL, "unknown seq_id");
    if (it->second->refs > 0) return set_error(IVQM_EBUSY, "sequence still used by submitted jobs");
    seq = it->second;
    s->seqs.erase(it);
    s->resident -= seq->bytes;
  }
  if (seq->dev) HIPCHK(hipFree(seq->dev));
  seq->dev = nullptr;
  return IVQM_OK;
}

int ivqm_server_submit(ivqm_server* s, const ivqm_job* job) {
  try {
    if (!s) return set_error(IVQM_EINVAL, "server is NULL");
    std::unique_ptr<Job> j(new Job());
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->stop) return set_error(IVQM_ESTATE, "server is shutting down");
      std::shared_ptr<Sequence> seq;
      RCCHK(validate_job(s, job, &seq));
      j->seq = seq;
      seq->refs++;
    }
    j->d = *job;
    j->planes.assign(job->planes, job->planes + 3 * job->num_frames);
    if (job->strides) j->strides.assign(job->strides, job->strides + 3 * job->num_frames);
    if (job->mask_planes) {
      j->mask_planes.assign(job->mask_planes, job->mask_planes + job->num_frames);
      if (job->mask_strides) j->mask_strides.assign(job->mask_strides, job->mask_strides + job->num_frames);
    }
    std::memset(&j->result, 0, sizeof(j->result));
    j->t_submit = Clock::now();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->submitted.push_back(std::move(j));
      s->n_submitted++;
    }
    s->cv_work.notify_one();
    return IVQM_OK;
  } catch (const std::bad_alloc&) {
    return set_error(IVQM_ENOMEM, "host allocation failed");
  }
}

int ivqm_server_poll(ivqm_server* s, ivqm_job_result* out, int32_t max, int32_t* n, int32_t wait_us) {
  if (!s || !n || (max > 0 && !out)) return set_error(IVQM_EINVAL, "NULL argument");
  *n = 0;
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->done.empty() && wait_us > 0)
    s->cv_done.wait_for(lk, std::chrono::microseconds(wait_us), [s] { return !s->done.empty(); });
  while (*n < max && !s->done.empty()) {
    out[(*n)++] = s->done.front();
    s->done.pop_front();
  }
  return IVQM_OK;
}

int ivqm_server_drain(ivqm_server* s) {
  if (!s) return set_error(IVQM_EINVAL, "server is NULL");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [s] { return s->n_completed == s->n_submitted; });
  return IVQM_OK;
}

int ivqm_server_get_stats(ivqm_server* s, ivqm_server_stats* o) {
  if (!s || !o) return set_error(IVQM_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(o, 0, sizeof(*o));
  o->jobs_submitted = s->n_submitted;
  o->jobs_completed = s->n_completed;
  o->jobs_failed = s->n_failed;
  o->pairs_completed = s->n_pairs;
  o->batches = s->n_batches;
  o->budget_bytes = s->arena.cap();
  o->budget_in_use = s->arena.used();
  o->budget_peak = s->arena.peak();
  o->resident_bytes = s->resident;
  o->max_batch_fill = s->max_fill;
  return IVQM_OK;
}

}  // extern "C"
