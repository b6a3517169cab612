// Batched search over many ligands and the continuous-batching server.
#pragma once
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.hpp"

namespace cdk {

// Many jobs (ligands x conformers x quaternions) searched together; the
// fused engine mixes their poses in shared launches.
struct Batch {
  Context* ctx = nullptr;
  Protein* prot = nullptr;
  struct Item {
    const Ligand* lig;
    int nconf, nquat;
    bool shared;
    std::vector<float> coords, quats;
  };
  std::vector<Item> items;
  Prepared prep;
  bool prepared = false, dirty = true;
  void prepare();
  void run(Engine engine);
  const std::vector<PoseResults>& results() const { return prep.results; }
};

struct Request {
  int64_t id = 0;
  int prot = 0;
  std::unique_ptr<Ligand> lig;
  std::vector<float> coords, quats;
  int nconf = 0, nquat = 0;
  bool shared = false;
  size_t npose = 0, next = 0, done = 0;
  size_t bytes = 0;
  double served_s = 0;  // GPU seconds attributed so far (estimated at launch, measured at completion)
  bool collected = false;  // returned by result() while a launch set in flight still held it
  PoseResults res;
  std::chrono::steady_clock::time_point t_submit, t_done;
  enum State { kQueued, kActive, kDone, kFailed } state = kQueued;
  int err_code = 0;
  std::string err;
};

// Server statistics (cdk_server_stats): index -> meaning.
enum ServerStat {
  kStatDone = 0,        // requests finished
  kStatPoses,           // poses finished
  kStatLaunchSets,      // launch sets run
  kStatMeanSetPoses,    // mean poses per launch set
  kStatBusy,            // GPU seconds of the launch sets (HIP events)
  kStatAdmitted,        // bytes admitted now (workspace + active requests)
  kStatQueued,          // requests waiting for admission
  kStatPeak,            // peak admitted bytes
  kStatTracked,         // requests the server holds (not yet collected)
  kStatInFlight,        // launch sets submitted and not yet finished (0, 1 or 2)
  kStatCount
};

class Server {
 public:
  Server(Context& ctx, std::vector<Protein*> prots, size_t budget, int max_batch);
  ~Server();
  int64_t submit(int prot, int natom, const float* charge, const float* eps, const float* radius, int nconf,
                 const float* coords, int nquat, const float* quats, bool shared);
  // Returns kOk, kErrPending, or the request's error code.  A finished
  // request is released by the call that returns its final state.
  int result(int64_t id, int wait_ms, float* energy, int* index, float* mincoor, double* latency_s,
             std::string* err);
  int stats(int n, double* out);

 private:
  // One launch set: its jobs (pose ranges of requests of one protein) and
  // device tables.
  struct Flight {
    Prepared prep;
    std::vector<Job> jobs;
    std::vector<Request*> owners;
    std::vector<double> est_s;  // GPU seconds charged to each owner at launch
    int prot = 0;
    bool live = false;
  };
  void loop();
  bool pick(Flight& f);           // next launch set (lock held); false if no work
  void complete(Flight& f, int code, const std::string& msg);  // results back (lock taken inside)
  void admit();                   // admission by footprint, arrival order (lock held)
  Context& ctx_;
  std::vector<Protein*> prots_;
  std::vector<double> s_per_pose_;  // running GPU seconds per pose, per protein
  size_t budget_;
  int max_batch_;
  std::mutex m_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<int64_t> queue_;
  std::vector<Request*> active_;
  std::unordered_map<int64_t, std::unique_ptr<Request>> reqs_;
  int64_t next_id_ = 1;
  size_t reserved_ = 0, admitted_ = 0, peak_ = 0;
  Flight fl_[2];
  int in_flight_ = 0;
  bool stop_ = false;
  double busy_s_ = 0;
  int64_t n_done_ = 0, poses_done_ = 0, launches_ = 0;
  std::thread th_;
};

}  // namespace cdk
