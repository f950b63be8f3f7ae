// serve.cpp — the serving loop behind include/hpt.h (hpt_server_*).
//
// No counterpart in the reference, which renders one scene per process
// (src/bin/hala-pathtracer.rs).  Host code only: the server drives the
// renderer through the same C entry points any host uses (hpt_create,
// hpt_share_scene, hpt_render, hpt_group_render, hpt_read_image,
// hpt_denoise), from threads of its own:
//
//  - a dispatcher admits queued jobs in submission order while their
//    footprints fit the budget (first come, first served: the head of the
//    queue waits for room and holds back the jobs behind it);
//  - streams mode: one host thread per admitted job makes the job's handle,
//    renders its batches on the handle's stream one after the other and reads
//    the image back;
//  - wavefront mode: one host thread per resident scene steps the scene's
//    active jobs together (hpt_group_render; at most two steps queued), and
//    reads back the jobs whose last step was queued one step earlier, while
//    the device runs the next step.
//
// A job's result image reaches host memory before its finish time is taken
// (the latency end point).  Then its handle is kept for a later job of the
// same class (reuse_handles, the default: no handle to make, no frame state
// to allocate, no batch graph to capture), its footprint still charged to
// the budget until a job needs the room, or destroyed.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hpt.h"

namespace hpt {
void set_last_error(const std::string& message);  // hpt_last_error() of this thread (api.cpp)
}

namespace {

// Device bytes per pixel a job holds: per path in flight, its path state (11
// float4 + a float2 of path fields), two queue slots and two shadow rays; with
// shading sort a key byte and a sorted queue entry more; per pixel, the three
// rgba32f images and the rgba8 final image.
constexpr uint64_t kPathBytes = 288;
constexpr uint64_t kSortBytes = 5;
constexpr uint64_t kImageBytes = 52;
constexpr uint32_t kMaxStepJobs = 256;   // jobs per hpt_group_render step
constexpr uint32_t kMaxFramesInFlight = 64;
constexpr uint32_t kServerTag = 0x53545048u;  // "HPTS"

struct Fail : std::runtime_error {
  hpt_status code;
  Fail(hpt_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void need(bool c, const std::string& m, hpt_status code = HPT_ERROR_INVALID_ARGUMENT) {
  if (!c) throw Fail(code, m);
}

// A renderer call that failed: its status and this thread's message.
void ok(hpt_status st, const char* what) {
  if (st != HPT_OK) throw Fail(st, std::string(what) + ": " + hpt_last_error());
}

template <typename F>
hpt_status guarded(F&& fn) {
  try {
    fn();
    hpt::set_last_error("");
    return HPT_OK;
  } catch (const Fail& f) {
    hpt::set_last_error(f.what());
    return f.code;
  } catch (const std::bad_alloc&) {
    hpt::set_last_error("host allocation failed");
    return HPT_ERROR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    hpt::set_last_error(e.what());
    return HPT_ERROR_DEVICE;
  } catch (...) {
    hpt::set_last_error("unknown error");
    return HPT_ERROR_DEVICE;
  }
}

struct Job {
  uint64_t id = 0;
  hpt_job_desc d{};
  uint32_t chunk = 1;      // frames per batch (streams) or per step (wavefront)
  uint64_t fp = 0;         // footprint charged against the budget
  double submitted = 0, admitted = -1, finished = -1;
  hpt_renderer* h = nullptr;
  uint32_t done = 0;       // frames queued so far
  hpt_status status = HPT_OK;
  std::string error;
  bool finished_flag = false;
  std::vector<float> image;  // kept result (keep_images)
};
using JobPtr = std::shared_ptr<Job>;

// What a job handle is made with: a finished job's handle renders a later
// job with the same key without being made again (zero-filled before use, so
// padding compares equal).
struct HandleKey {
  uint32_t scene, width, height, max_depth, rr_depth, tonemap, aces, simple_aces;
  float exposure;
  uint32_t frames, launch, sort, tail, block;
};

struct CachedHandle {
  HandleKey key;
  hpt_renderer* h;
  uint64_t fp;
};

struct SceneSlot {
  hpt_serve_scene_desc desc{};
  // wavefront mode
  std::deque<JobPtr> pending;  // admitted, not yet in the scene's wavefront
  std::thread worker;
  bool worker_started = false;
};

}  // namespace

struct hpt_server {
  uint32_t tag = kServerTag;  // a handle of another kind is refused, not written into
  hpt_server_config cfg{};
  std::chrono::steady_clock::time_point t0;
  std::vector<std::unique_ptr<SceneSlot>> scenes;
  hpt_denoiser* denoiser = nullptr;
  std::mutex net_mu;  // one network, one set of activation buffers

  std::mutex mu;
  std::condition_variable cv;
  std::deque<JobPtr> queue;            // submitted, not admitted
  std::map<uint64_t, JobPtr> jobs;     // every job not yet released
  std::deque<uint64_t> completed;      // finished, not yet returned by next_result
  std::vector<uint64_t> reap;          // streams-mode jobs whose threads have ended
  uint64_t next_id = 1, used = 0, peak = 0;
  uint32_t running = 0;                // admitted, unfinished jobs
  std::list<CachedHandle> cache;       // idle handles of finished jobs, oldest first
  uint64_t cached_bytes = 0, reused = 0;
  uint64_t n_submitted = 0, n_finished = 0, n_failed = 0;
  uint64_t steps = 0, max_step_jobs = 0, step_jobs = 0;
  bool closing = false;
  std::thread dispatcher;
  bool dispatcher_started = false;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  bool dry() const { return cfg.dry_run_ns_per_path != 0; }
};

namespace {

uint32_t resolve(uint32_t v, uint32_t dflt) { return v == HPT_SERVE_DEFAULT ? dflt : v; }

bool ok_server(const hpt_server* s) { return s && s->tag == kServerTag; }

HandleKey handle_key(const hpt_server* s, const Job& j) {
  HandleKey k;
  std::memset(&k, 0, sizeof k);
  k.scene = j.d.scene;
  k.width = j.d.width;
  k.height = j.d.height;
  k.max_depth = j.d.max_depth;
  k.rr_depth = j.d.rr_depth;
  k.tonemap = j.d.enable_tonemap;
  k.aces = j.d.enable_aces;
  k.simple_aces = j.d.use_simple_aces;
  k.exposure = j.d.exposure_value;
  k.frames = s->cfg.mode == HPT_SERVE_WAVEFRONT ? 1 : j.chunk;
  k.launch = resolve(j.d.launch_mode, s->cfg.launch_mode);
  k.sort = resolve(j.d.shade_sort, HPT_SHADE_SORT_OFF);
  k.tail = resolve(j.d.tail_from, HPT_TAIL_FROM_NONE);
  k.block = resolve(j.d.traversal_block, 128);
  return k;
}

// A resident scene's slot (slots are added under the lock and never removed).
SceneSlot& slot(hpt_server* s, uint32_t id) {
  std::lock_guard<std::mutex> g(s->mu);
  return *s->scenes.at(id);
}

// Frames per batch (streams) or per wavefront step, and the footprint.
void plan(const hpt_server* s, const hpt_job_desc& d, uint64_t* fp, uint32_t* chunk) {
  const uint64_t px = (uint64_t)d.width * d.height;
  if (s->cfg.mode == HPT_SERVE_WAVEFRONT) {
    const double want = std::llround((double)s->cfg.step_paths / (double)px);
    const uint64_t k = std::max<uint64_t>(1, std::min<uint64_t>({(uint64_t)d.spp, (uint64_t)s->cfg.max_frames * 4,
                                                                   (uint64_t)std::max(0.0, want)}));
    *chunk = (uint32_t)k;
    *fp = px * ((k + 1) * kPathBytes + kImageBytes);
  } else {
    const uint32_t k = std::max<uint32_t>(1, std::min(resolve(d.frames_in_flight, s->cfg.max_frames), d.spp));
    const bool sorted = resolve(d.shade_sort, HPT_SHADE_SORT_OFF) != HPT_SHADE_SORT_OFF;
    *chunk = k;
    *fp = px * ((uint64_t)k * (kPathBytes + (sorted ? kSortBytes : 0)) + kImageBytes);
  }
}

void validate(const hpt_server* s, const hpt_job_desc* d) {
  need(d != nullptr, "null job");
  need(d->scene < s->scenes.size(), "unknown scene id");
  need(d->width >= 1 && d->height >= 1 && d->width <= 16384 && d->height <= 16384, "image size out of range");
  need(d->spp >= 1, "spp must be at least 1");
  need(d->max_depth >= 1, "max_depth must be at least 1");
  const uint32_t k = d->frames_in_flight;
  need(k == HPT_SERVE_DEFAULT || (k >= 1 && k <= kMaxFramesInFlight), "frames_in_flight must be 1..64");
  need(d->denoise <= 1, "denoise must be 0 or 1");
  const uint32_t tb = d->traversal_block;
  need(tb == HPT_SERVE_DEFAULT || tb == 64 || tb == 128 || tb == 256, "traversal_block must be 64, 128 or 256");
  const uint32_t lm = d->launch_mode;
  need(lm == HPT_SERVE_DEFAULT || lm == HPT_LAUNCH_EAGER || lm == HPT_LAUNCH_GRAPH, "unknown launch mode");
}

// ---------------------------------------------------------------- one job

hpt_renderer* make_handle(hpt_server* s, const Job& j) {
  const SceneSlot& sc = slot(s, j.d.scene);
  if (s->dry()) {
    need(!sc.desc.dry_run_fail, "device lost (dry run)", HPT_ERROR_DEVICE);
    return nullptr;
  }
  hpt_create_info ci{};
  ci.width = j.d.width;
  ci.height = j.d.height;
  ci.max_depth = j.d.max_depth;
  ci.rr_depth = j.d.rr_depth;
  ci.enable_tonemap = j.d.enable_tonemap;
  ci.enable_aces = j.d.enable_aces;
  ci.use_simple_aces = j.d.use_simple_aces;
  ci.features = sc.desc.features;
  ci.max_samples = 0;
  ci.exposure_value = j.d.exposure_value;
  ci.device = s->cfg.device;
  hpt_renderer* h = nullptr;
  ok(hpt_create(&ci, &h), "hpt_create");
  try {
    const bool wave = s->cfg.mode == HPT_SERVE_WAVEFRONT;
    // A wavefront job traces in its scene's group; its own batch stays one frame.
    ok(hpt_set_frames_in_flight(h, wave ? 1 : j.chunk), "hpt_set_frames_in_flight");
    ok(hpt_share_scene(h, sc.desc.scene), "hpt_share_scene");
    ok(hpt_set_launch_mode(h, resolve(j.d.launch_mode, s->cfg.launch_mode)), "hpt_set_launch_mode");
    ok(hpt_set_shade_sort(h, resolve(j.d.shade_sort, HPT_SHADE_SORT_OFF)), "hpt_set_shade_sort");
    ok(hpt_set_tail_from(h, resolve(j.d.tail_from, HPT_TAIL_FROM_NONE)), "hpt_set_tail_from");
    ok(hpt_set_traversal_block(h, resolve(j.d.traversal_block, 128)), "hpt_set_traversal_block");
  } catch (...) {
    hpt_destroy(h);
    throw;
  }
  return h;
}

void dry_sleep(const hpt_server* s, const Job& j, uint32_t frames) {
  const uint64_t ns = (uint64_t)frames * j.d.width * j.d.height * s->cfg.dry_run_ns_per_path;
  std::this_thread::sleep_for(std::chrono::nanoseconds(ns));
}

// The job's image in host memory: accumulated radiance, or the denoised
// colour computed on the device from the handle's resident images.
void read_result(hpt_server* s, Job& j, std::vector<float>& img) {
  const size_t n = (size_t)j.d.width * j.d.height * 4;
  img.resize(n);
  if (s->dry()) {
    std::fill(img.begin(), img.end(), (float)j.id);
    return;
  }
  if (j.d.denoise) {
    std::lock_guard<std::mutex> g(s->net_mu);
    need(s->denoiser != nullptr, "the job asks for denoising and the server has no denoiser", HPT_ERROR_STATE);
    ok(hpt_denoise(s->denoiser, j.h, img.data(), n * sizeof(float), 0), "hpt_denoise");
  } else {
    ok(hpt_read_image(j.h, HPT_IMAGE_ACCUM_RGBA32F, img.data(), n * sizeof(float)), "hpt_read_image");
  }
}

// Ends a job: its image (unless it already failed), its finish time, its
// handle, its budget share.
void finish(hpt_server* s, Job& j, hpt_status st, const std::string& err) {
  std::vector<float> img;
  if (st == HPT_OK) {
    try {
      read_result(s, j, img);
    } catch (const Fail& f) {
      st = f.code;
      img.clear();
      j.error = f.what();
    } catch (const std::exception& e) {
      st = HPT_ERROR_DEVICE;
      img.clear();
      j.error = e.what();
    }
  } else {
    j.error = err;
  }
  const double t = s->now();
  // A handle that ended well is kept for the next job of its class (its
  // footprint stays charged to the budget while it waits); any other is freed.
  const bool keep = j.h && st == HPT_OK && s->cfg.reuse_handles;
  if (j.h && !keep) {
    hpt_destroy(j.h);
    j.h = nullptr;
  }
  std::lock_guard<std::mutex> g(s->mu);
  if (keep) {
    s->cache.push_back(CachedHandle{handle_key(s, j), j.h, j.fp});
    s->cached_bytes += j.fp;
    j.h = nullptr;
  }
  j.finished = t;
  j.status = st;
  if (s->cfg.keep_images && st == HPT_OK) j.image = std::move(img);
  j.finished_flag = true;
  s->used -= j.fp;
  s->running -= 1;
  s->n_finished += 1;
  if (st != HPT_OK) s->n_failed += 1;
  s->completed.push_back(j.id);
  s->cv.notify_all();
}

// Streams mode: the whole job on its own handle and stream.
void run_job(hpt_server* s, JobPtr j) {
  hpt_status st = HPT_OK;
  std::string err;
  try {
    if (j->h)  // a handle of the same class, kept from a finished job
      ok(hpt_reset_accumulation(j->h), "hpt_reset_accumulation");
    else
      j->h = make_handle(s, *j);
    // One batch in flight per job: the streams of the jobs share the
    // process's few hardware queues, each in order, so a job that queued all
    // of its batches at once would hold the jobs behind it on its queue.
    while (j->done < j->d.spp) {
      const uint32_t n = std::min(j->chunk, j->d.spp - j->done);
      if (s->dry()) {
        dry_sleep(s, *j, n);
      } else {
        ok(hpt_render(j->h, n, j->d.camera_index), "hpt_render");
        ok(hpt_wait_idle(j->h), "hpt_wait_idle");
      }
      j->done += n;
    }
  } catch (const Fail& f) {
    st = f.code;
    err = f.what();
  } catch (const std::exception& e) {
    st = HPT_ERROR_DEVICE;
    err = e.what();
  }
  finish(s, *j, st, err);
  std::lock_guard<std::mutex> g(s->mu);
  s->reap.push_back(j->id);
  s->cv.notify_all();
}

// ---------------------------------------------------------- wavefront mode

struct GroupHolder {
  hpt_group* g = nullptr;
  ~GroupHolder() {
    if (g) hpt_group_destroy(g);
  }
};

hpt_group* new_group(hpt_server* s, const SceneSlot& sc) {
  hpt_group* g = nullptr;
  ok(hpt_group_create(s->cfg.device, &g), "hpt_group_create");
  try {
    ok(hpt_group_set_launch_mode(g, resolve(sc.desc.step_launch_mode, s->cfg.launch_mode)), "group launch mode");
    ok(hpt_group_set_shade_sort(g, resolve(sc.desc.step_shade_sort, HPT_SHADE_SORT_OFF)), "group shade sort");
    ok(hpt_group_set_tail_from(g, resolve(sc.desc.step_tail_from, HPT_TAIL_FROM_NONE)), "group tail");
    ok(hpt_group_set_traversal_block(g, resolve(sc.desc.step_traversal_block, 128)), "group block");
  } catch (...) {
    hpt_group_destroy(g);
    throw;
  }
  return g;
}

// The wavefront of one scene: step after step, the next chunk of every
// active job together; finished jobs leave, newly admitted ones join.
void scene_worker(hpt_server* s, uint32_t scene) {
  SceneSlot& sc = slot(s, scene);
  GroupHolder gh;
  std::vector<JobPtr> active, finishing;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [&] { return !sc.pending.empty() || !active.empty() || !finishing.empty() || s->closing; });
      while (!sc.pending.empty()) {
        active.push_back(sc.pending.front());
        sc.pending.pop_front();
      }
      if (active.empty() && finishing.empty()) return;  // closing
    }
    std::vector<JobPtr> ended;
    if (!active.empty()) {
      const size_t m = std::min<size_t>(active.size(), kMaxStepJobs);
      std::vector<JobPtr> step(active.begin(), active.begin() + m);
      std::vector<uint32_t> frames(m), cams(m);
      std::vector<hpt_renderer*> hs(m);
      for (size_t i = 0; i < m; ++i) {
        frames[i] = std::min(step[i]->chunk, step[i]->d.spp - step[i]->done);
        cams[i] = step[i]->d.camera_index;
        hs[i] = step[i]->h;
      }
      hpt_status st = HPT_OK;
      std::string err;
      try {
        if (s->dry()) {
          for (size_t i = 0; i < m; ++i) dry_sleep(s, *step[i], frames[i]);
        } else {
          if (!gh.g) gh.g = new_group(s, sc);
          ok(hpt_group_render(gh.g, hs.data(), frames.data(), cams.data(), (uint32_t)m), "hpt_group_render");
        }
      } catch (const Fail& f) {
        st = f.code;
        err = f.what();
      } catch (const std::exception& e) {
        st = HPT_ERROR_DEVICE;
        err = e.what();
      }
      if (st != HPT_OK) {  // the step's jobs fail; the scene goes on
        for (const JobPtr& j : step) {
          active.erase(std::find(active.begin(), active.end(), j));
          finish(s, *j, st, err);
        }
        continue;
      }
      {
        std::lock_guard<std::mutex> g(s->mu);
        s->steps += 1;
        s->max_step_jobs = std::max<uint64_t>(s->max_step_jobs, m);
        s->step_jobs += m;
      }
      for (size_t i = 0; i < m; ++i) {
        step[i]->done += frames[i];
        if (step[i]->done >= step[i]->d.spp) {
          active.erase(std::find(active.begin(), active.end(), step[i]));
          ended.push_back(step[i]);
        }
      }
    }
    // Jobs whose last step was queued before this one: their read-back waits
    // for that step while the device runs the next.
    for (const JobPtr& j : finishing) finish(s, *j, HPT_OK, "");
    finishing = std::move(ended);
  }
}

// ------------------------------------------------------------- admission

void maybe_release(hpt_server* s, uint64_t id);

// Admission of the queue's head, under the lock: a kept handle of its class
// if there is one (its footprint is already charged), else room in the
// budget, made by evicting kept handles oldest first when needed (returned
// in `evicted` for the caller to free before the job allocates).
bool try_admit(hpt_server* s, Job& j, uint32_t cap, std::vector<hpt_renderer*>& evicted) {
  if (cap != 0 && s->running >= cap) return false;
  if (s->cfg.reuse_handles && !s->cache.empty()) {
    const HandleKey k = handle_key(s, j);
    for (auto it = s->cache.end(); it != s->cache.begin();) {
      --it;  // the most recently kept first
      if (std::memcmp(&it->key, &k, sizeof k) != 0) continue;
      j.h = it->h;
      s->cached_bytes -= it->fp;
      s->used += j.fp;
      s->reused += 1;
      s->cache.erase(it);
      return true;
    }
  }
  while (s->used + s->cached_bytes + j.fp > s->cfg.budget_bytes && !s->cache.empty()) {
    evicted.push_back(s->cache.front().h);
    s->cached_bytes -= s->cache.front().fp;
    s->cache.pop_front();
  }
  if (s->used + s->cached_bytes + j.fp > s->cfg.budget_bytes) return false;
  s->used += j.fp;
  return true;
}

// Admits jobs; in streams mode it also owns the jobs' threads and joins each
// one after it ends (and all of them before it returns).
void dispatcher(hpt_server* s) {
  std::map<uint64_t, std::thread> workers;
  auto join = [&](const std::vector<uint64_t>& ids) {
    for (uint64_t id : ids) {
      auto it = workers.find(id);
      if (it == workers.end()) continue;
      it->second.join();
      workers.erase(it);
    }
  };
  for (;;) {
    JobPtr j;
    std::vector<uint64_t> reaped;
    std::vector<hpt_renderer*> evicted;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [&] { return !s->queue.empty() || s->closing; });
      if (s->queue.empty()) {  // closing with nothing queued
        lk.unlock();
        for (auto& kv : workers) kv.second.join();
        return;
      }
      JobPtr head = s->queue.front();
      const uint32_t cap = s->cfg.mode == HPT_SERVE_STREAMS ? s->cfg.max_running : 0;
      while (!try_admit(s, *head, cap, evicted)) s->cv.wait(lk);
      s->queue.pop_front();
      s->running += 1;
      s->peak = std::max(s->peak, s->used + s->cached_bytes);
      head->admitted = s->now();
      j = head;
      reaped.swap(s->reap);
    }
    // Idle handles evicted to make room are freed before the job allocates.
    for (hpt_renderer* h : evicted) hpt_destroy(h);
    join(reaped);
    if (s->cfg.mode == HPT_SERVE_STREAMS) {
      workers.emplace(j->id, std::thread(run_job, s, j));
      continue;
    }
    // Wavefront mode: the job's one-frame handle, then its scene's wavefront.
    try {
      if (j->h)
        ok(hpt_reset_accumulation(j->h), "hpt_reset_accumulation");
      else
        j->h = make_handle(s, *j);
    } catch (const Fail& f) {
      finish(s, *j, f.code, f.what());
      continue;
    } catch (const std::exception& e) {
      finish(s, *j, HPT_ERROR_DEVICE, e.what());
      continue;
    }
    std::lock_guard<std::mutex> g(s->mu);
    SceneSlot& sc = *s->scenes.at(j->d.scene);
    sc.pending.push_back(j);
    if (!sc.worker_started) {
      sc.worker = std::thread(scene_worker, s, j->d.scene);
      sc.worker_started = true;
    }
    s->cv.notify_all();
  }
}

void fill_result(const Job& j, hpt_job_result* out) {
  std::memset(out, 0, sizeof(*out));
  out->job_id = j.id;
  out->status = j.status;
  out->width = j.d.width;
  out->height = j.d.height;
  out->frames = j.status == HPT_OK ? j.d.spp : j.done;
  out->has_image = j.image.empty() ? 0 : 1;
  out->footprint_bytes = j.fp;
  out->submitted_s = j.submitted;
  out->admitted_s = j.admitted;
  out->finished_s = j.finished;
  std::strncpy(out->error, j.error.c_str(), sizeof(out->error) - 1);
}

// Forgets a finished job once its result was handed out and its image read.
void maybe_release(hpt_server* s, uint64_t id) {
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return;
  const Job& j = *it->second;
  const bool listed = std::find(s->completed.begin(), s->completed.end(), id) != s->completed.end();
  if (j.finished_flag && !listed && j.image.empty()) s->jobs.erase(it);
}

}  // namespace

extern "C" {

void hpt_server_config_init(hpt_server_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->device = 0;
  c->mode = HPT_SERVE_STREAMS;
  c->budget_bytes = 192ull << 30;
  c->max_frames = 32;
  c->keep_images = 0;
  c->step_paths = 8ull << 20;
  c->launch_mode = HPT_LAUNCH_GRAPH;
  c->max_running = 8;
  c->dry_run_ns_per_path = 0;
  c->reuse_handles = 1;
}

void hpt_job_desc_init(hpt_job_desc* d) {
  if (!d) return;
  std::memset(d, 0, sizeof(*d));
  d->max_depth = 8;
  d->rr_depth = 3;
  d->spp = 1;
  d->enable_tonemap = 1;
  d->enable_aces = 1;
  d->frames_in_flight = HPT_SERVE_DEFAULT;
  d->launch_mode = HPT_SERVE_DEFAULT;
  d->shade_sort = HPT_SERVE_DEFAULT;
  d->tail_from = HPT_SERVE_DEFAULT;
  d->traversal_block = HPT_SERVE_DEFAULT;
}

hpt_status hpt_server_create(const hpt_server_config* config, hpt_server** out) {
  return guarded([&] {
    need(config && out, "null argument");
    need(config->mode == HPT_SERVE_STREAMS || config->mode == HPT_SERVE_WAVEFRONT, "unknown serving mode");
    need(config->max_frames >= 1 && config->max_frames <= kMaxFramesInFlight, "max_frames must be 1..64");
    need(config->mode != HPT_SERVE_WAVEFRONT || config->step_paths >= 1, "step_paths must be at least 1");
    need(config->launch_mode == HPT_LAUNCH_EAGER || config->launch_mode == HPT_LAUNCH_GRAPH, "unknown launch mode");
    need(config->keep_images <= 1, "keep_images must be 0 or 1");
    need(config->reuse_handles <= 1, "reuse_handles must be 0 or 1");
    auto s = std::make_unique<hpt_server>();
    s->cfg = *config;
    s->t0 = std::chrono::steady_clock::now();
    s->dispatcher = std::thread(dispatcher, s.get());
    s->dispatcher_started = true;
    *out = s.release();
  });
}

hpt_status hpt_server_add_scene(hpt_server* s, const hpt_serve_scene_desc* scene, uint32_t* scene_id) {
  return guarded([&] {
    need(ok_server(s) && scene && scene_id, "null argument");
    need(s->dry() || scene->scene != nullptr, "null scene handle");
    need(scene->dry_run_fail == 0 || s->dry(), "dry_run_fail applies to dry runs only");
    need(scene->features < 8, "unknown feature bits");
    std::lock_guard<std::mutex> g(s->mu);
    auto slot = std::make_unique<SceneSlot>();
    slot->desc = *scene;
    s->scenes.push_back(std::move(slot));
    *scene_id = (uint32_t)(s->scenes.size() - 1);
  });
}

hpt_status hpt_server_set_denoiser(hpt_server* s, hpt_denoiser* d) {
  return guarded([&] {
    need(ok_server(s), "null or invalid server handle");
    std::lock_guard<std::mutex> g(s->net_mu);
    s->denoiser = d;
  });
}

hpt_status hpt_server_plan_job(hpt_server* s, const hpt_job_desc* job, uint64_t* footprint_bytes,
                               uint32_t* chunk_frames) {
  return guarded([&] {
    need(ok_server(s) && footprint_bytes && chunk_frames, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    validate(s, job);
    plan(s, *job, footprint_bytes, chunk_frames);
  });
}

hpt_status hpt_server_submit(hpt_server* s, const hpt_job_desc* job, uint64_t* job_id) {
  return guarded([&] {
    need(ok_server(s) && job_id, "null argument");
    auto j = std::make_shared<Job>();
    std::lock_guard<std::mutex> g(s->mu);
    need(!s->closing, "the server is closing", HPT_ERROR_STATE);
    validate(s, job);
    if (job->denoise && !s->dry()) need(s->denoiser != nullptr, "denoise asked and no denoiser set", HPT_ERROR_STATE);
    j->d = *job;
    plan(s, j->d, &j->fp, &j->chunk);
    need(j->fp <= s->cfg.budget_bytes,
         "job needs " + std::to_string(j->fp) + " B, above the budget of " + std::to_string(s->cfg.budget_bytes) + " B");
    j->id = s->next_id++;
    j->submitted = s->now();
    s->jobs[j->id] = j;
    s->queue.push_back(j);
    s->n_submitted += 1;
    *job_id = j->id;
    s->cv.notify_all();
  });
}

hpt_status hpt_server_wait(hpt_server* s, uint64_t job_id, hpt_job_result* out) {
  return guarded([&] {
    need(ok_server(s) && out, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    need(it != s->jobs.end(), "unknown or released job id");
    JobPtr j = it->second;
    s->cv.wait(lk, [&] { return j->finished_flag; });
    fill_result(*j, out);
  });
}

hpt_status hpt_server_next_result(hpt_server* s, double timeout_s, hpt_job_result* out, int* got) {
  return guarded([&] {
    need(ok_server(s) && out && got, "null argument");
    *got = 0;
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] { return !s->completed.empty(); };
    if (timeout_s < 0)
      s->cv.wait(lk, ready);
    else
      s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), ready);
    if (s->completed.empty()) return;
    const uint64_t id = s->completed.front();
    s->completed.pop_front();
    fill_result(*s->jobs.at(id), out);
    *got = 1;
    maybe_release(s, id);
  });
}

hpt_status hpt_server_read_image(hpt_server* s, uint64_t job_id, float* dst, size_t bytes) {
  return guarded([&] {
    need(ok_server(s) && dst, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->jobs.find(job_id);
    need(it != s->jobs.end(), "unknown or released job id");
    Job& j = *it->second;
    need(j.finished_flag, "the job has not finished", HPT_ERROR_STATE);
    need(!j.image.empty(), "the job holds no image (failed, already read, or keep_images is off)", HPT_ERROR_STATE);
    need(bytes == j.image.size() * sizeof(float), "destination size is not width*height*16");
    std::memcpy(dst, j.image.data(), bytes);
    std::vector<float>().swap(j.image);
    maybe_release(s, job_id);
  });
}

hpt_status hpt_server_drain(hpt_server* s) {
  return guarded([&] {
    need(ok_server(s), "null or invalid server handle");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv.wait(lk, [&] { return s->n_finished == s->n_submitted; });
  });
}

hpt_status hpt_server_clock(hpt_server* s, double* seconds) {
  return guarded([&] {
    need(ok_server(s) && seconds, "null argument");
    *seconds = s->now();
  });
}

hpt_status hpt_server_get_stats(hpt_server* s, hpt_server_stats* out) {
  return guarded([&] {
    need(ok_server(s) && out, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    out->submitted = s->n_submitted;
    out->finished = s->n_finished;
    out->failed = s->n_failed;
    out->bytes_in_use = s->used;
    out->peak_bytes = s->peak;
    out->wavefront_steps = s->steps;
    out->max_jobs_per_step = s->max_step_jobs;
    out->jobs_in_steps = s->step_jobs;
    out->handles_reused = s->reused;
    out->cached_bytes = s->cached_bytes;
  });
}

void hpt_server_destroy(hpt_server* s) {
  if (!ok_server(s)) return;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv.wait(lk, [&] { return s->n_finished == s->n_submitted; });
    s->closing = true;
    s->cv.notify_all();
  }
  if (s->dispatcher_started) s->dispatcher.join();
  for (auto& sc : s->scenes)
    if (sc->worker_started) sc->worker.join();
  for (const CachedHandle& c : s->cache) hpt_destroy(c.h);
  s->cache.clear();
  s->tag = 0;
  delete s;
}

}  // extern "C"
