// Batched relaxation of many samples: see batch.h.
#include "batch.h"

#include <algorithm>
#include <cstring>
#include <thread>

namespace zxt {

namespace {
[[noreturn]] void fail(const std::string& what) { throw EngineError{what}; }
}  // namespace

Batch::Batch(bool use_graphs, uint32_t streams) : use_graphs_(use_graphs) {
  CaptureLock lock(false);
  Engine::check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "stream");
  prof.set_stream(stream);
  if (streams > 0) {
    Engine::check(hipEventCreateWithFlags(&fork_, hipEventDisableTiming), "event");
    for (uint32_t k = 0; k < std::min<uint32_t>(streams, 8); ++k) {
      hipStream_t s = nullptr;
      Engine::check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream");
      aux_.push_back(s);
      hipEvent_t ev = nullptr;
      Engine::check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "event");
      join_.push_back(ev);
    }
  }
}

Batch::~Batch() {
  (void)hipStreamSynchronize(stream);
  jobs_.clear();
  CaptureLock lock(false);
  if (graph_) (void)hipGraphExecDestroy(graph_);
  if (span_) (void)hipGraphExecDestroy(span_);
  for (const Retired& r : retired_) {
    (void)hipGraphExecDestroy(r.exec);
    (void)hipEventDestroy(r.done);
  }
  for (Table* t : {&ap_table_, &step_table_, &part_step_table_[0], &part_step_table_[1]}) {
    if (t->host_desc) (void)hipHostFree(t->host_desc);
    if (t->host_start) (void)hipHostFree(t->host_start);
  }
  for (NearTable& t : near_table_) {
    if (t.host_desc) (void)hipHostFree(t.host_desc);
    if (t.host_start) (void)hipHostFree(t.host_start);
  }
  for (hipEvent_t e : inflight_) (void)hipEventDestroy(e);
  for (hipEvent_t e : spare_) (void)hipEventDestroy(e);
  for (hipStream_t s : aux_) (void)hipStreamDestroy(s);
  for (hipEvent_t e : join_) (void)hipEventDestroy(e);
  if (fork_) (void)hipEventDestroy(fork_);
  (void)hipStreamDestroy(stream);
}

Batch::Job* Batch::find(uint64_t id) const {
  for (const auto& j : jobs_)
    if (j->id == id) return j.get();
  fail("UnknownJob");
}

uint64_t Batch::add(uint32_t group, const float* q, uint32_t n, uint32_t seed, double delta, JobKind kind,
                    uint32_t steps) {
  auto job = std::make_unique<Job>();
  job->engine = std::make_unique<Engine>(stream);
  Engine& e = *job->engine;
  if (q)
    e.set(group, q, size_t(n) * 4);
  else
    e.generate_random(group, n, seed);
  e.set_neighbourhood(delta);
  job->kind = kind;
  job->remaining = kind == JobKind::Relax ? std::min<uint32_t>(steps, 1000) : 0;
  job->id = next_id_++;
  const uint64_t id = job->id;
  jobs_.push_back(std::move(job));
  return id;
}

// Every device pointer a multi-job kernel dereferences must be set before launch.
static void check_desc(const JobDesc& d, bool allpairs) {
  bool ok = d.q && d.ops && d.n && d.rows.force && d.rows.energy && d.rows.nearest;
  if (allpairs)
    ok = ok && d.ap_partial && d.ap_pnear && d.plan.splits();
  else
    ok = ok && d.prev && d.state && d.bb_partials;
  if (!ok) fail("IncompleteJobDescriptor");
}

// Uploads a job table (only when it changed) and returns its JobSet.
JobSet Batch::upload(Table& t, const std::vector<JobDesc>& descs) {
  const size_t count = descs.size();
  JobSet s;
  if (count == 0) {  // nothing to launch, nothing to upload (also under capture)
    t.last.clear();
    t.set = s;
    return s;
  }
  for (const JobDesc& d : descs) check_desc(d, &t == &ap_table_);
  bool same = t.last.size() == count;
  for (size_t k = 0; same && k < count; ++k) same = std::memcmp(&t.last[k], &descs[k], sizeof(JobDesc)) == 0;
  if (same) return t.set;
  // A captured step must replay with the tables it was captured with.
  if (capturing_) throw GraphTableChanged{};
  s.count = uint32_t(count);
  if (count > t.capacity) {
    Engine::check(hipStreamSynchronize(stream), "wait");
    CaptureLock lock(false);
    if (t.host_desc) (void)hipHostFree(t.host_desc);
    if (t.host_start) (void)hipHostFree(t.host_start);
    t.capacity = std::max<size_t>(count, 2 * t.capacity);
    Engine::check(hipHostMalloc(reinterpret_cast<void**>(&t.host_desc), t.capacity * sizeof(JobDesc)), "pinned");
    Engine::check(hipHostMalloc(reinterpret_cast<void**>(&t.host_start), 3 * (t.capacity + 1) * sizeof(uint32_t)),
                  "pinned");
    t.desc.ensure(t.capacity * sizeof(JobDesc));
    t.start.ensure(3 * (t.capacity + 1) * sizeof(uint32_t));
  } else {
    // The previous upload of these pinned buffers may still be in flight.
    Engine::check(hipStreamSynchronize(stream), "wait");
  }
  uint32_t* point = t.host_start;
  uint32_t* ap = point + (count + 1);
  uint32_t* median = ap + (count + 1);
  point[0] = ap[0] = median[0] = 0;
  for (size_t k = 0; k < count; ++k) {
    t.host_desc[k] = descs[k];
    point[k + 1] = point[k] + descs[k].point_blocks;
    ap[k + 1] = ap[k] + descs[k].point_blocks * descs[k].plan.splits();
    median[k + 1] = median[k] + descs[k].median_blocks;
  }
  Engine::check(hipMemcpyAsync(t.desc.p, t.host_desc, count * sizeof(JobDesc), hipMemcpyHostToDevice, stream),
                "table");
  Engine::check(hipMemcpyAsync(t.start.p, t.host_start, 3 * (count + 1) * sizeof(uint32_t), hipMemcpyHostToDevice,
                               stream),
                "table");
  const uint32_t* dstart = t.start.as<uint32_t>();
  s.table = t.desc.as<JobDesc>();
  s.point_start = dstart;
  s.ap_start = dstart + (count + 1);
  s.median_start = dstart + 2 * (count + 1);
  s.point_total = point[count];
  s.ap_total = ap[count];
  s.median_total = median[count];
  t.last = descs;
  t.set = s;
  return s;
}

// The near-field job table (sorted by force variant) and its stage prefix sums, uploaded
// only when it changed, as upload does for the other tables.
NearSet Batch::upload_near(NearTable& t, const std::vector<NearJob>& jobs) {
  const size_t count = jobs.size();
  if (count == 0) {
    t.last.clear();
    t.set = NearSet{};
    return t.set;
  }
  bool same = t.last.size() == count;
  for (size_t k = 0; same && k < count; ++k) same = std::memcmp(&t.last[k], &jobs[k], sizeof(NearJob)) == 0;
  if (same) return t.set;
  if (capturing_) throw GraphTableChanged{};
  const size_t words = size_t(kNearStages) * (count + 1);
  Engine::check(hipStreamSynchronize(stream), "wait");  // the last upload may still be in flight
  if (count > t.capacity) {
    CaptureLock lock(false);
    if (t.host_desc) (void)hipHostFree(t.host_desc);
    if (t.host_start) (void)hipHostFree(t.host_start);
    t.host_desc = nullptr;
    t.host_start = nullptr;
    t.capacity = std::max<size_t>(count, 2 * t.capacity);
    const size_t cap_words = size_t(kNearStages) * (t.capacity + 1);
    Engine::check(hipHostMalloc(reinterpret_cast<void**>(&t.host_desc), t.capacity * sizeof(NearJob)), "pinned");
    Engine::check(hipHostMalloc(reinterpret_cast<void**>(&t.host_start), cap_words * sizeof(uint32_t)), "pinned");
    t.desc.ensure(t.capacity * sizeof(NearJob));
    t.start.ensure(cap_words * sizeof(uint32_t));
  }
  NearSet s;
  s.count = uint32_t(count);
  for (uint32_t st = 0; st < kNearStages; ++st) {
    uint32_t* row = t.host_start + size_t(st) * (count + 1);
    row[0] = 0;
    for (size_t k = 0; k < count; ++k) row[k + 1] = row[k] + near_stage_blocks(jobs[k], st);
    s.total[st] = row[count];
  }
  const uint32_t* force = t.host_start + size_t(kNsForce) * (count + 1);
  for (size_t k = 0; k < count; ++k) {
    if (s.keys == 0 || s.key[s.keys - 1] != jobs[k].key) {
      if (s.keys == kNearKeys) fail("TooManyNearVariants");
      s.key[s.keys] = jobs[k].key;
      s.key_first[s.keys] = force[k];
      ++s.keys;
    }
    s.key_blocks[s.keys - 1] = force[k + 1] - s.key_first[s.keys - 1];
  }
  std::memcpy(t.host_desc, jobs.data(), count * sizeof(NearJob));
  Engine::check(hipMemcpyAsync(t.desc.p, t.host_desc, count * sizeof(NearJob), hipMemcpyHostToDevice, stream),
                "table");
  Engine::check(hipMemcpyAsync(t.start.p, t.host_start, words * sizeof(uint32_t), hipMemcpyHostToDevice, stream),
                "table");
  s.table = t.desc.as<NearJob>();
  s.start = t.start.as<uint32_t>();
  t.last = jobs;
  t.set = s;
  return s;
}

void Batch::finish(Job& j, bool recorded) {
  try {
    j.engine->finish_relax(recorded);
    j.state = 1;
  } catch (const EngineError& e) {
    j.state = 2;
    j.error = e.message;
  }
}

uint32_t Batch::step(uint32_t most, uint32_t* taken_out) {
  if (taken_out) *taken_out = 0;
  reap_graphs();
  uint32_t completed = 0;
  // Measure and spacing jobs run as they come (each ends with a wait).
  for (auto& jp : jobs_) {
    Job& j = *jp;
    if (j.state != 0 || j.kind == JobKind::Relax) continue;
    try {
      if (j.kind == JobKind::Measure)
        j.engine->measure();
      else
        j.engine->spacing();
      j.state = 1;
    } catch (const EngineError& e) {
      j.state = 2;
      j.error = e.message;
    }
    ++completed;
  }
  std::vector<Job*> active;
  std::vector<uint64_t> key;
  bool fresh = false;
  for (auto& jp : jobs_) {
    Job& j = *jp;
    if (j.state != 0 || j.kind != JobKind::Relax) continue;
    if (!j.started) {
      try {
        j.engine->begin_relax();
        j.started = true;
        fresh = true;
      } catch (const EngineError& e) {
        j.state = 2;
        j.error = e.message;
        ++completed;
        continue;
      }
    }
    if (j.remaining == 0) {  // a relax of 0 steps: only the measure that ends it
      finish(j, false);
      ++completed;
      continue;
    }
    active.push_back(&j);
    key.push_back(j.id);
  }
  if (active.empty()) return completed;
  // A step is the same launch sequence as the last one when the jobs are the same and none
  // starts now (a first step sizes buffers and may wait): then a captured graph replays it.
  const bool steady = use_graphs_ && !prof.on() && !fresh && key == last_key_;
  last_key_ = key;
  // Whether the extra streams are forked from the batch stream and not yet joined back.
  bool forked = false;
  auto join_aux = [&](uint32_t lanes) {
    for (uint32_t k = 0; k < lanes; ++k) {
      Engine::check(hipEventRecord(join_[k], aux_[k]), "join");
      Engine::check(hipStreamWaitEvent(stream, join_[k], 0), "join");
    }
    forked = false;
  };
  // Records one step; `first` and `last`: whether it starts and ends a recorded run of
  // steps (a single step is both).
  auto record = [&](bool first, bool last) {
    // Tables first (a captured step cannot upload), then the near-field launches, forked
    // onto extra streams after the last step's moves, beside the all-pairs launch.
    std::vector<JobDesc> ap;
    std::vector<std::pair<NearJob, Job*>> near;
    for (Job* j : active) {
      Engine& e = *j->engine;
      e.prepare_step();
      if (e.relax_state_.has_near) {
        NearJob nj;
        if (e.near_job(nj)) near.push_back({nj, j});
      } else {
        JobDesc d;
        if (e.step_forces(d)) ap.push_back(d);
      }
    }
    // Jobs of one force variant share a launch.  With two extra streams the near-field
    // jobs are dealt alternately (in that order) to two sets of similar work, each built and
    // run on its own stream: one set's build overlaps the other's force launches.
    std::stable_sort(near.begin(), near.end(), [](const auto& a, const auto& b) { return a.first.key < b.first.key; });
    // (Profiled steps stay on one stream: the profiler's events are on the batch stream.)
    const uint32_t lanes = prof.on() ? 0u : std::min<uint32_t>(uint32_t(aux_.size()), near.size() >= 2 ? 2u : 1u);
    std::vector<NearJob> part[2];
    std::vector<Job*> owner[2];
    for (size_t k = 0; k < near.size(); ++k) {
      part[lanes == 2 ? k % 2 : 0].push_back(near[k].first);
      owner[lanes == 2 ? k % 2 : 0].push_back(near[k].second);
    }
    // With two sets, each set's BB step follows its force launches on its own stream, so a
    // run of steps joins the streams only at its end: a set goes on to its next build while
    // the other set's force launches still run.  The sets' jobs are disjoint and a job's
    // step reads only its own buffers.  Between runs the streams are joined, so the next
    // upload (which waits for the batch stream) cannot overwrite a table still in use.
    const bool independent = lanes == 2;
    std::vector<JobDesc> steps, part_steps[2];
    for (Job* j : active) {
      int in = -1;
      for (int p = 0; independent && p < 2 && in < 0; ++p)
        if (std::find(owner[p].begin(), owner[p].end(), j) != owner[p].end()) in = p;
      (in < 0 ? steps : part_steps[in]).push_back(j->engine->step_job());
    }
    const JobSet ap_set = upload(ap_table_, ap);
    const JobSet step_set = upload(step_table_, steps);
    const JobSet part_set[2] = {upload(part_step_table_[0], part_steps[0]), upload(part_step_table_[1], part_steps[1])};
    const NearSet near_set[2] = {upload_near(near_table_[0], part[0]), upload_near(near_table_[1], part[1])};
    const bool split = lanes > 0 && !near.empty() && (lanes == 2 || !ap.empty());
    if (split && (first || !independent)) {
      Engine::check(hipEventRecord(fork_, stream), "fork");
      for (uint32_t k = 0; k < lanes; ++k) Engine::check(hipStreamWaitEvent(aux_[k], fork_, 0), "fork");
      forked = true;
    }
    for (uint32_t k = 0; k < 2; ++k) {
      const hipStream_t on = split ? aux_[k < lanes ? k : 0] : stream;
      launch_near_jobs(prof, on, near_set[k]);
      if (independent) launch_bb_step(prof, on, part_set[k]);  // move and fold included
    }
    if (!ap.empty()) launch_allpairs(prof, stream, ap_set);
    if (split && (last || !independent)) join_aux(lanes);
    launch_bb_step(prof, stream, step_set);  // the all-pairs jobs' (or every job's)
  };
  // Captures `steps` recorded steps into one executable graph; null when a job table
  // changed during the capture (nothing captured has run: the caller records eagerly).
  auto capture = [&](uint32_t steps) -> hipGraphExec_t {
    hipGraph_t g = nullptr;
    bool changed = false;
    hipError_t ended;
    {
      CaptureLock lock(true);
      Engine::check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "capture");
      capturing_ = true;
      try {
        for (uint32_t k = 0; k < steps; ++k) record(k == 0, k + 1 == steps);
      } catch (const GraphTableChanged&) {
        // A later step of a run found a table changed: the streams forked by its first step
        // join back, so that the capture ends whole (and is then dropped).
        if (forked) join_aux(std::min<uint32_t>(uint32_t(aux_.size()), 2u));
        changed = true;
      } catch (...) {
        capturing_ = false;
        (void)hipStreamEndCapture(stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      capturing_ = false;
      ended = hipStreamEndCapture(stream, &g);
    }
    if (changed) {
      if (g) (void)hipGraphDestroy(g);
      Engine::check(ended, "capture");
      return nullptr;
    }
    Engine::check(ended, "capture");
    hipGraphExec_t exec = nullptr;
    hipError_t made;
    {
      // Instantiation allocates the graph's argument memory and the streams of its
      // parallel branches: like any allocation, not while another thread captures.
      CaptureLock lock(false);
      made = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
    }
    Engine::check(made, "graph");
    ++captures_;
    return exec;
  };
  uint32_t taken = 1;
  try {
    if (steady && graph_ && graph_key_ == key) {
      // Steps that end no job (each job's last step finishes it on the host) can run as one
      // launch of a graph of `span` steps: the same launches, recorded back to back.
      uint32_t span = std::min(most, kMaxSpan);
      for (Job* j : active) span = std::min(span, j->remaining - 1);
      if (span >= 2) {
        if (!span_ || span_key_ != key || span_steps_ != span) {
          retire_graph(span_);
          span_key_.clear();
          span_ = capture(span);
          if (span_) {
            span_key_ = key;
            span_steps_ = span;
          }
        }
        if (span_) {
          Engine::check(hipGraphLaunch(span_, stream), "graph");
          replays_ += span;
          taken = span;
        } else {
          record(true, true);
        }
      } else {
        Engine::check(hipGraphLaunch(graph_, stream), "graph");
        ++replays_;
      }
    } else if (steady) {
      retire_graph(graph_);
      retire_graph(span_);
      graph_key_.clear();
      span_key_.clear();
      graph_ = capture(1);
      if (graph_) {
        graph_key_ = key;
        Engine::check(hipGraphLaunch(graph_, stream), "graph");
      } else {
        record(true, true);  // capture the next step
      }
    } else {
      record(true, true);
    }
  } catch (const EngineError& e) {
    for (Job* j : active) {
      j->state = 2;
      j->error = e.message;
    }
    err = e.message;
    return completed + uint32_t(active.size());
  }
  mark_step();
  if (taken_out) *taken_out = taken;
  std::vector<Job*> ending;
  for (Job* j : active) {
    for (uint32_t k = 0; k < taken; ++k) j->engine->stepped();
    j->remaining -= taken;
    if (j->remaining == 0) ending.push_back(j);
  }
  finish_all(ending);
  return completed + uint32_t(ending.size());
}

// The jobs that reached their steps end as a relax call does (finish_relax), but their
// final force passes run as multi-sample launches (one per near-field build stage and
// force variant, one for all pairs) and one wait covers them all.  A job whose near field
// overflowed redoes its pass alone, as finish_relax would.
void Batch::finish_all(const std::vector<Job*>& ending) {
  if (ending.empty()) return;
  std::vector<Job*> ready;
  std::vector<NearJob> near;
  std::vector<JobDesc> ap;
  for (Job* j : ending) {
    Engine& e = *j->engine;
    try {
      e.finish_prepare();
      if (e.relax_state_.has_near) {
        NearJob nj;
        e.near_job(nj);
        near.push_back(nj);
      } else {
        ap.push_back(e.relax_state_.allpairs_desc(e, e.count(), e.order()));
      }
      ready.push_back(j);
    } catch (const EngineError& x) {
      j->state = 2;
      j->error = x.message;
    }
  }
  try {
    std::stable_sort(near.begin(), near.end(), [](const NearJob& a, const NearJob& b) { return a.key < b.key; });
    const NearSet near_set = upload_near(near_table_[0], near);
    const JobSet ap_set = upload(ap_table_, ap);
    launch_near_jobs(prof, stream, near_set);
    if (!ap.empty()) launch_allpairs(prof, stream, ap_set);
    for (Job* j : ready) j->engine->finish_reads(true);
    Engine::check(hipStreamSynchronize(stream), "wait");
  } catch (const EngineError& x) {
    for (Job* j : ready) {
      j->state = 2;
      j->error = x.message;
    }
    return;
  }
  // The host's share (the last step's result, per-point separations and measures) is the
  // jobs' own: it runs on several host threads.  A pass to redo waits for the others.
  std::vector<int> outcome(ready.size(), 0);  // 0 done, 1 redo, 2 failed
  std::vector<std::string> message(ready.size());
  const size_t threads = std::min<size_t>(ready.size(), std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
  auto work = [&](size_t first) {
    for (size_t k = first; k < ready.size(); k += threads) {
      try {
        outcome[k] = ready[k]->engine->finish_check(true) ? 0 : 1;
      } catch (const EngineError& x) {
        outcome[k] = 2;
        message[k] = x.message;
      }
    }
  };
  std::vector<std::thread> pool;
  for (size_t t = 1; t < threads; ++t) pool.emplace_back(work, t);
  work(0);
  for (std::thread& t : pool) t.join();
  for (size_t k = 0; k < ready.size(); ++k) {
    Job& j = *ready[k];
    try {
      if (outcome[k] == 2) throw EngineError{message[k]};
      if (outcome[k] == 1) j.engine->finish_relax(false);
      j.state = 1;
    } catch (const EngineError& x) {
      j.state = 2;
      j.error = x.message;
    }
  }
}

int Batch::status(uint64_t id) const { return find(id)->state; }
std::string Batch::stats(uint64_t id) { return find(id)->engine->stats(); }
std::string Batch::error(uint64_t id) const { return find(id)->error; }
const float* Batch::quaternions(uint64_t id) const { return find(id)->engine->quaternions(); }
uint32_t Batch::count(uint64_t id) const { return find(id)->engine->count(); }
size_t Batch::bytes(uint64_t id) const { return find(id)->engine->device_bytes(); }

void Batch::remove(uint64_t id) {
  Engine::check(hipStreamSynchronize(stream), "wait");
  for (auto it = jobs_.begin(); it != jobs_.end(); ++it) {
    if ((*it)->id == id) {
      jobs_.erase(it);
      return;
    }
  }
  fail("UnknownJob");
}

hipEvent_t Batch::take_event() {
  hipEvent_t ev = nullptr;
  if (!spare_.empty()) {
    ev = spare_.back();
    spare_.pop_back();
  } else {
    Engine::check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "event");
  }
  return ev;
}

// Steps in flight: an event recorded after each step's launches.
void Batch::mark_step() {
  hipEvent_t ev = take_event();
  Engine::check(hipEventRecord(ev, stream), "event");
  inflight_.push_back(ev);
}

// Called outside capture only: the event marks the end of everything the stream holds,
// the graph's last launch included.
void Batch::retire_graph(hipGraphExec_t& g) {
  if (!g) return;
  hipEvent_t ev = take_event();
  Engine::check(hipEventRecord(ev, stream), "event");
  retired_.push_back({g, ev});
  g = nullptr;
}

void Batch::reap_graphs() {
  size_t kept = 0;
  for (const Retired& r : retired_) {
    if (hipEventQuery(r.done) == hipSuccess) {
      CaptureLock lock(false);  // frees what instantiation allocated
      (void)hipGraphExecDestroy(r.exec);
      spare_.push_back(r.done);
    } else {
      retired_[kept++] = r;
    }
  }
  retired_.resize(kept);
}

uint32_t Batch::in_flight() {
  reap_graphs();
  while (!inflight_.empty() && hipEventQuery(inflight_.front()) == hipSuccess) {
    spare_.push_back(inflight_.front());
    inflight_.pop_front();
  }
  return uint32_t(inflight_.size());
}

void Batch::wait(uint32_t most) {
  while (in_flight() > most) {
    Engine::check(hipEventSynchronize(inflight_.front()), "wait");
    spare_.push_back(inflight_.front());
    inflight_.pop_front();
  }
}

uint32_t Batch::running() const {
  uint32_t c = 0;
  for (const auto& j : jobs_) c += j->state == 0;
  return c;
}

}  // namespace zxt
