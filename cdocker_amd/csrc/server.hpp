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
  PoseResults res;
  std::chrono::steady_clock::time_point t_submit, t_done;
  enum State { kQueued, kActive, kDone, kFailed } state = kQueued;
  int err_code = 0;
  std::string err;
};

class Server {
 public:
  Server(Context& ctx, std::vector<Protein*> prots, size_t budget, int max_batch);
  ~Server();
  int64_t submit(int prot, int natom, const float* charge, const float* eps, const float* radius, int nconf,
                 const float* coords, int nquat, const float* quats, bool shared);
  // Returns kOk, kErrPending, or the request's error code.
  int result(int64_t id, int wait_ms, float* energy, int* index, float* mincoor, double* latency_s,
             std::string* err);
  void stats(double* out);

 private:
  void loop();
  Context& ctx_;
  std::vector<Protein*> prots_;
  size_t budget_;
  int max_batch_;
  std::mutex m_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<int64_t> queue_;
  std::vector<Request*> active_;
  std::unordered_map<int64_t, std::unique_ptr<Request>> reqs_;
  int64_t next_id_ = 1;
  size_t reserved_ = 0, admitted_ = 0, peak_ = 0;
  Prepared prep_;  // device tables reused by every launch set
  bool stop_ = false;
  double busy_s_ = 0;
  int64_t n_done_ = 0, poses_done_ = 0, launches_ = 0;
  std::thread th_;
};

}  // namespace cdk
