// Batched search over many ligands and the continuous-batching server.
#include "server.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace cdk {

void Batch::run(Engine engine) {
  if (!prepared || dirty) {
    prepare();
  }
  execute_search(*ctx, prep, engine, ctx->use_graph);
}

void Batch::prepare() {
  std::vector<Job> jobs;
  jobs.reserve(items.size());
  for (const Item& it : items) {
    Job j;
    j.lig = it.lig;
    j.nconf = it.nconf;
    j.coords = it.coords.data();
    j.nquat = it.nquat;
    j.quats = it.quats.data();
    j.shared_quats = it.shared;
    jobs.push_back(j);
  }
  prepare_search(*ctx, *prot, jobs, prep);
  prepared = true;
  dirty = false;
}

Server::Server(Context& ctx, std::vector<Protein*> prots, size_t budget, int max_batch)
    : ctx_(ctx), prots_(std::move(prots)), budget_(budget), max_batch_(std::max(1, max_batch)) {
  // The fused engine's working arrays never exceed the context's workspace
  // budget; that much of the HBM budget is reserved up front.
  reserved_ = ctx_.ws_budget;
  if (budget_ <= reserved_)
    throw Error(kErrBudget, "server HBM budget must exceed the context workspace budget");
  th_ = std::thread([this] { loop(); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(m_);
    stop_ = true;
  }
  cv_work_.notify_all();
  th_.join();
}

int64_t Server::submit(int prot, int natom, const float* charge, const float* eps, const float* radius, int nconf,
                       const float* coords, int nquat, const float* quats, bool shared) {
  if (prot < 0 || prot >= (int)prots_.size()) throw Error(kErrInvalid, "protein index out of range");
  if (nconf <= 0 || nquat <= 0) throw Error(kErrInvalid, "request needs conformers and quaternions");
  const Protein& P = *prots_[prot];
  auto r = std::make_unique<Request>();
  r->prot = prot;
  r->lig = make_ligand(P, natom, charge, eps, radius);
  r->coords.assign(coords, coords + (size_t)nconf * natom * 3);
  r->quats.assign(quats, quats + (size_t)(shared ? 1 : nconf) * nquat * 4);
  r->nconf = nconf;
  r->nquat = nquat;
  r->shared = shared;
  r->npose = (size_t)nconf * nquat;
  // Device footprint of the request while admitted: its pose slots,
  // conformer coordinates, per-pose results and quaternions.  The launch-set
  // working arrays are the context's fixed workspace, reserved at creation.
  r->bytes = r->coords.size() * 4 + r->quats.size() * 4 +
             r->npose * (sizeof(PoseSlot) + sizeof(RotMeta) + sizeof(unsigned long long));
  if (r->bytes > budget_ - reserved_)
    throw Error(kErrBudget, "request footprint exceeds the server's HBM budget");
  r->res.energy.resize(r->npose);
  r->res.index.resize(r->npose * 3);
  r->res.mincoor.resize(r->npose * 3);
  r->t_submit = std::chrono::steady_clock::now();
  int64_t id;
  {
    std::lock_guard<std::mutex> lk(m_);
    id = next_id_++;
    r->id = id;
    queue_.push_back(id);
    reqs_[id] = std::move(r);
  }
  cv_work_.notify_all();
  return id;
}

int Server::result(int64_t id, int wait_ms, float* energy, int* index, float* mincoor, double* latency_s,
                   std::string* err) {
  std::unique_lock<std::mutex> lk(m_);
  auto it = reqs_.find(id);
  if (it == reqs_.end()) throw Error(kErrInvalid, "unknown request id");
  Request* r = it->second.get();
  auto finished = [&] { return r->state == Request::kDone || r->state == Request::kFailed; };
  if (wait_ms < 0)
    cv_done_.wait(lk, finished);
  else if (wait_ms > 0)
    cv_done_.wait_for(lk, std::chrono::milliseconds(wait_ms), finished);
  if (!finished()) return kErrPending;
  if (r->state == Request::kFailed) {
    if (err) *err = r->err;
    return r->err_code;
  }
  if (energy) std::memcpy(energy, r->res.energy.data(), r->npose * sizeof(float));
  if (index) std::memcpy(index, r->res.index.data(), r->npose * 3 * sizeof(int));
  if (mincoor) std::memcpy(mincoor, r->res.mincoor.data(), r->npose * 3 * sizeof(float));
  if (latency_s) *latency_s = std::chrono::duration<double>(r->t_done - r->t_submit).count();
  return kOk;
}

void Server::stats(double* out) {
  std::lock_guard<std::mutex> lk(m_);
  out[0] = (double)n_done_;
  out[1] = (double)poses_done_;
  out[2] = (double)launches_;
  out[3] = launches_ ? (double)poses_done_ / (double)launches_ : 0.0;
  out[4] = busy_s_;
  out[5] = (double)(reserved_ + admitted_);
  out[6] = (double)queue_.size();
  out[7] = (double)(reserved_ + peak_);
}

void Server::loop() {
  for (;;) {
    std::vector<Job> jobs;
    std::vector<Request*> owners;
    int prot = 0;
    {
      std::unique_lock<std::mutex> lk(m_);
      for (;;) {
        // Admission by device footprint, in arrival order.
        while (!queue_.empty()) {
          Request* r = reqs_[queue_.front()].get();
          if (!active_.empty() && reserved_ + admitted_ + r->bytes > budget_) break;
          admitted_ += r->bytes;
          peak_ = std::max(peak_, admitted_);
          r->state = Request::kActive;
          active_.push_back(r);
          queue_.pop_front();
        }
        if (!active_.empty()) break;
        if (stop_) return;
        cv_work_.wait(lk);
      }
      // One launch set per protein: the protein of the oldest active request.
      prot = active_.front()->prot;
      std::vector<Request*> mine;
      for (Request* r : active_)
        if (r->prot == prot && r->next < r->npose) mine.push_back(r);
      size_t left = (size_t)max_batch_;
      const size_t share = std::max<size_t>(1, left / std::max<size_t>(1, mine.size()));
      std::vector<size_t> take(mine.size(), 0);
      for (int pass = 0; pass < 2 && left > 0; ++pass)
        for (size_t i = 0; i < mine.size() && left > 0; ++i) {
          Request* r = mine[i];
          const size_t avail = r->npose - r->next - take[i];
          const size_t k = std::min({avail, left, pass == 0 ? share : left});
          take[i] += k;
          left -= k;
        }
      for (size_t i = 0; i < mine.size(); ++i) {
        if (take[i] == 0) continue;
        Request* r = mine[i];
        Job j;
        j.lig = r->lig.get();
        j.nconf = r->nconf;
        j.coords = r->coords.data();
        j.nquat = r->nquat;
        j.quats = r->quats.data();
        j.shared_quats = r->shared;
        j.pose_begin = r->next;
        j.pose_end = r->next + take[i];
        r->next += take[i];
        jobs.push_back(j);
        owners.push_back(r);
      }
    }
    const auto t0 = std::chrono::steady_clock::now();
    int code = kOk;
    std::string msg;
    try {
      prepare_search(ctx_, *prots_[prot], jobs, prep_);
      execute_search(ctx_, prep_, kEngineFused, false);
    } catch (const Error& e) {
      code = e.code;
      msg = e.what();
    } catch (const std::exception& e) {
      code = kErrHip;
      msg = e.what();
    }
    const auto t1 = std::chrono::steady_clock::now();
    {
      std::lock_guard<std::mutex> lk(m_);
      busy_s_ += std::chrono::duration<double>(t1 - t0).count();
      ++launches_;
      for (size_t i = 0; i < jobs.size(); ++i) {
        Request* r = owners[i];
        const size_t b = jobs[i].pose_begin, n = jobs[i].npose();
        if (code == kOk) {
          const PoseResults& o = prep_.results[i];
          std::copy(o.energy.begin(), o.energy.end(), r->res.energy.begin() + b);
          std::copy(o.index.begin(), o.index.end(), r->res.index.begin() + b * 3);
          std::copy(o.mincoor.begin(), o.mincoor.end(), r->res.mincoor.begin() + b * 3);
          poses_done_ += (int64_t)n;
        }
        r->done += n;
        const bool failed = code != kOk;
        if (failed || r->done >= r->npose) {
          if (r->state == Request::kActive) {
            r->state = failed ? Request::kFailed : Request::kDone;
            r->err_code = code;
            r->err = msg;
            r->t_done = t1;
            admitted_ -= r->bytes;
            active_.erase(std::find(active_.begin(), active_.end(), r));
            if (!failed) ++n_done_;
          }
        }
      }
    }
    cv_done_.notify_all();
  }
}

}  // namespace cdk
