// Many samples relaxed together (no counterpart in the reference, which runs one sample
// per engine): every job is an Engine on the batch's stream; per step, the all-pairs
// force passes of all jobs run as one launch, the near-field passes of all jobs as one
// launch per build stage and per force-kernel variant, and the BB step and fold of all
// jobs as one launch per stage.  A job's arithmetic is the same as alone (the
// multi-sample kernels share the single-sample kernels' code), so its result is
// bit-identical alone and in the mix.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <deque>
#include <memory>
#include <string>
#include <vector>

#include "engine.h"

namespace zxt {

enum class JobKind : uint32_t { Relax = 0, Measure = 1, Spacing = 2 };

class Batch {
 public:
  // use_graphs: replay steady-state steps from a captured graph.  streams > 0: the
  // near-field launches of a step run on an extra stream, beside the all-pairs launch, and
  // with 2 or more on two (half the near-field jobs each); 0: all on the batch stream.
  Batch(bool use_graphs, uint32_t streams);
  ~Batch();

  // Adds a job: positions q (n rows) or, when q is null, n seeded Haar orientations.
  // delta: relaxation neighbourhood (0 all pairs).  Returns its id.  Throws EngineError.
  uint64_t add(uint32_t group, const float* q, uint32_t n, uint32_t seed, double delta, JobKind kind,
               uint32_t steps);
  // One BB step for every running relax job, then the jobs that reached their steps are
  // finished (one wait each) and pending measure/spacing jobs run.  Returns how many jobs
  // completed.  With most > 1, steps that end no job may go as one launch of a graph of up
  // to `most` steps (at most kMaxSpan); *taken receives how many steps were taken.
  uint32_t step(uint32_t most = 1, uint32_t* taken = nullptr);
  static constexpr uint32_t kMaxSpan = 64;
  // 0 running, 1 done, 2 failed (message in error(id)); throws for an unknown id.
  int status(uint64_t id) const;
  std::string stats(uint64_t id);
  std::string error(uint64_t id) const;
  const float* quaternions(uint64_t id) const;
  uint32_t count(uint64_t id) const;
  size_t bytes(uint64_t id) const;
  void remove(uint64_t id);
  uint32_t running() const;
  Engine& engine(uint64_t id) const { return *find(id)->engine; }
  // Steps whose launches have not finished on the device; wait() blocks until at most
  // `most` are left (a serving loop keeps each batch at most a step or two ahead).
  uint32_t in_flight();
  void wait(uint32_t most);
  // Graphs captured, and steps replayed from one.
  uint64_t captures() const { return captures_; }
  uint64_t replays() const { return replays_; }

  hipStream_t stream = nullptr;
  Prof prof;
  std::string err;

 private:
  struct Job {
    uint64_t id = 0;
    std::unique_ptr<Engine> engine;
    JobKind kind = JobKind::Relax;
    uint32_t remaining = 0;
    bool started = false;
    int state = 0;  // 0 running, 1 done, 2 failed
    std::string error;
  };
  struct Table {
    DevBuf desc, start;
    JobDesc* host_desc = nullptr;  // pinned
    uint32_t* host_start = nullptr;
    size_t capacity = 0;
    std::vector<JobDesc> last;
    JobSet set;
  };
  Job* find(uint64_t id) const;
  void mark_step();
  hipEvent_t take_event();
  // A replaced graph is destroyed only after the stream has passed its last launch: HIP
  // does not promise to keep an executable graph alive while a launch of it still runs.
  void retire_graph(hipGraphExec_t& g);
  void reap_graphs();
  struct Retired {
    hipGraphExec_t exec;
    hipEvent_t done;
  };
  std::vector<Retired> retired_;
  std::deque<hipEvent_t> inflight_;
  std::vector<hipEvent_t> spare_;
  JobSet upload(Table& t, const std::vector<JobDesc>& descs);
  struct NearTable {
    DevBuf desc, start;
    NearJob* host_desc = nullptr;  // pinned
    uint32_t* host_start = nullptr;
    size_t capacity = 0;
    std::vector<NearJob> last;
    NearSet set;
  };
  // Near-field jobs go to up to two sets, each on its own extra stream, so that one set's
  // build overlaps the other's force launches.
  NearTable near_table_[2];
  NearSet upload_near(NearTable& t, const std::vector<NearJob>& jobs);
  void finish(Job& j, bool recorded);
  void finish_all(const std::vector<Job*>& ending);

  std::vector<std::unique_ptr<Job>> jobs_;
  uint64_t next_id_ = 1;
  Table ap_table_, step_table_;
  // With two near-field sets on their own streams, each set's BB step runs on its stream
  // (its own table), so within a recorded run of steps the sets never wait for each other.
  Table part_step_table_[2];
  bool use_graphs_ = false;
  bool capturing_ = false;
  struct GraphTableChanged {};
  uint64_t captures_ = 0, replays_ = 0;
  std::vector<hipStream_t> aux_;
  std::vector<hipEvent_t> join_;
  hipEvent_t fork_ = nullptr;
  hipGraphExec_t graph_ = nullptr;  // one step
  std::vector<uint64_t> graph_key_;
  hipGraphExec_t span_ = nullptr;  // span_steps_ steps, for steps that end no job
  std::vector<uint64_t> span_key_;
  uint32_t span_steps_ = 0;
  std::vector<uint64_t> last_key_;
};

}  // namespace zxt
