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
  // Starting guess of GPU seconds per pose (about 1.2 us at 64^3, scaled by
  // grid points); replaced by measurements as launch sets finish.
  for (const Protein* p : prots_) s_per_pose_.push_back(1.2e-6 * (double)p->G() / 262144.0);
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
  // The request is looked up again after every wake-up: another caller may
  // have collected (and released) it meanwhile.
  auto it = reqs_.find(id);
  if (it == reqs_.end() || it->second->collected)
    throw Error(kErrInvalid, "unknown request id (never submitted, or already collected)");
  Request* r = it->second.get();
  auto finished = [&] {
    it = reqs_.find(id);
    r = it == reqs_.end() ? nullptr : it->second.get();
    return !r || r->collected || r->state == Request::kDone || r->state == Request::kFailed;
  };
  if (wait_ms < 0)
    cv_done_.wait(lk, finished);
  else if (wait_ms > 0)
    cv_done_.wait_for(lk, std::chrono::milliseconds(wait_ms), finished);
  if (!finished()) return kErrPending;
  if (!r || r->collected) throw Error(kErrInvalid, "request collected by another caller while this one waited");
  int code = kOk;
  if (r->state == Request::kFailed) {
    if (err) *err = r->err;
    code = r->err_code;
  } else {
    if (energy) std::memcpy(energy, r->res.energy.data(), r->npose * sizeof(float));
    if (index) std::memcpy(index, r->res.index.data(), r->npose * 3 * sizeof(int));
    if (mincoor) std::memcpy(mincoor, r->res.mincoor.data(), r->npose * 3 * sizeof(float));
    if (latency_s) *latency_s = std::chrono::duration<double>(r->t_done - r->t_submit).count();
  }
  // Released now.  Only a failed request can still be in a launch set in
  // flight (its later poses); it is marked collected and released when that
  // set completes.
  bool pinned = false;
  for (const Flight& f : fl_)
    if (f.live && std::find(f.owners.begin(), f.owners.end(), r) != f.owners.end()) pinned = true;
  if (pinned)
    r->collected = true;
  else
    reqs_.erase(it);
  return code;
}

int Server::stats(int n, double* out) {
  std::lock_guard<std::mutex> lk(m_);
  double v[kStatCount];
  v[kStatDone] = (double)n_done_;
  v[kStatPoses] = (double)poses_done_;
  v[kStatLaunchSets] = (double)launches_;
  v[kStatMeanSetPoses] = launches_ ? (double)poses_done_ / (double)launches_ : 0.0;
  v[kStatBusy] = busy_s_;
  v[kStatAdmitted] = (double)(reserved_ + admitted_);
  v[kStatQueued] = (double)queue_.size();
  v[kStatPeak] = (double)(reserved_ + peak_);
  v[kStatTracked] = (double)reqs_.size();
  v[kStatInFlight] = (double)in_flight_;
  const int k = std::min(n, (int)kStatCount);
  for (int i = 0; i < k; ++i) out[i] = v[i];
  return k;
}

void Server::admit() {
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
}

// The next launch set: least attained service first.  The active request
// with remaining poses and the fewest GPU seconds served so far picks the
// protein; every active request of that protein with remaining poses shares
// the set (equal shares of max_batch_, what is left to the least served
// first).  Each owner is charged its estimated GPU seconds now, corrected by
// the measured time when the set finishes, so a long request cannot hold the
// device against requests that arrive after it, on either protein.
bool Server::pick(Flight& f) {
  Request* lead = nullptr;
  for (Request* r : active_)
    if (r->next < r->npose && (!lead || r->served_s < lead->served_s ||
                               (r->served_s == lead->served_s && r->id < lead->id)))
      lead = r;
  if (!lead) return false;
  const int prot = lead->prot;
  std::vector<Request*> mine;
  for (Request* r : active_)
    if (r->prot == prot && r->next < r->npose) mine.push_back(r);
  std::sort(mine.begin(), mine.end(), [](const Request* a, const Request* b) {
    return a->served_s < b->served_s || (a->served_s == b->served_s && a->id < b->id);
  });
  size_t left = (size_t)max_batch_;
  const size_t share = std::max<size_t>(1, left / mine.size());
  std::vector<size_t> take(mine.size(), 0);
  for (int pass = 0; pass < 2 && left > 0; ++pass)
    for (size_t i = 0; i < mine.size() && left > 0; ++i) {
      const size_t avail = mine[i]->npose - mine[i]->next - take[i];
      const size_t k = std::min({avail, left, pass == 0 ? share : left});
      take[i] += k;
      left -= k;
    }
  f.jobs.clear();
  f.owners.clear();
  f.est_s.clear();
  f.prot = prot;
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
    const double est = (double)take[i] * s_per_pose_[prot];
    r->served_s += est;
    f.jobs.push_back(j);
    f.owners.push_back(r);
    f.est_s.push_back(est);
  }
  return true;
}

void Server::complete(Flight& f, int code, const std::string& msg) {
  const auto t1 = std::chrono::steady_clock::now();
  double gpu_s = 0;
  if (code == kOk) {
    try {
      gpu_s = 1e-3 * (double)search_gpu_ms(f.prep);
    } catch (const std::exception&) {
      gpu_s = 0;
    }
  }
  std::lock_guard<std::mutex> lk(m_);
  size_t poses = 0;
  for (const Job& j : f.jobs) poses += j.npose();
  busy_s_ += gpu_s;
  ++launches_;
  if (code == kOk && poses > 0 && gpu_s > 0) {
    // running estimate of GPU seconds per pose on this protein
    double& s = s_per_pose_[f.prot];
    s = 0.8 * s + 0.2 * gpu_s / (double)poses;
  }
  for (size_t i = 0; i < f.jobs.size(); ++i) {
    Request* r = f.owners[i];
    const size_t b = f.jobs[i].pose_begin, n = f.jobs[i].npose();
    if (r->collected) continue;  // collected while this set was in flight
    if (code == kOk && poses > 0) r->served_s += gpu_s * (double)n / (double)poses - f.est_s[i];
    if (code == kOk && r->state == Request::kActive) {
      const PoseResults& o = f.prep.results[i];
      std::copy(o.energy.begin(), o.energy.end(), r->res.energy.begin() + b);
      std::copy(o.index.begin(), o.index.end(), r->res.index.begin() + b * 3);
      std::copy(o.mincoor.begin(), o.mincoor.end(), r->res.mincoor.begin() + b * 3);
      poses_done_ += (int64_t)n;
    }
    r->done += n;
    const bool failed = code != kOk;
    if ((failed || r->done >= r->npose) && r->state == Request::kActive) {
      r->state = failed ? Request::kFailed : Request::kDone;
      r->err_code = code;
      r->err = msg;
      r->t_done = t1;
      admitted_ -= r->bytes;
      active_.erase(std::find(active_.begin(), active_.end(), r));
      if (!failed) ++n_done_;
    }
  }
  f.live = false;
  --in_flight_;
  // requests collected while this set was in flight, and now in no other set
  for (Request* r : f.owners) {
    if (!r->collected) continue;
    bool other = false;
    for (const Flight& g : fl_)
      if (g.live && std::find(g.owners.begin(), g.owners.end(), r) != g.owners.end()) other = true;
    if (!other) reqs_.erase(r->id);
  }
  f.owners.clear();
  f.jobs.clear();
}

// Two launch sets in flight: while set i runs on the device, the host picks
// set i + 1, builds its tables and enqueues it behind set i on the context
// stream, then waits for set i and hands its results back.  The device always
// has the next set queued, and the host work of one set overlaps the device
// work of the other.
void Server::loop() {
  int older = 0;  // fl_[older]: the live set submitted first (if any is live)
  for (;;) {
    Flight* nf = nullptr;
    {
      std::unique_lock<std::mutex> lk(m_);
      for (;;) {
        if (!stop_) admit();
        Flight& free_slot = fl_[older].live ? fl_[older ^ 1] : fl_[older];
        if (!stop_ && !free_slot.live && pick(free_slot)) {
          nf = &free_slot;
          nf->live = true;
          ++in_flight_;
          break;
        }
        if (fl_[older].live) break;  // nothing new: finish the live set
        if (stop_) return;
        cv_work_.wait(lk);
      }
    }
    if (nf) {
      int code = kOk;
      std::string msg;
      try {
        build_search(*prots_[nf->prot], nf->jobs, nf->prep);
        submit_search(ctx_, nf->prep);
      } catch (const Error& e) {
        code = e.code;
        msg = e.what();
      } catch (const std::exception& e) {
        code = kErrHip;
        msg = e.what();
      }
      if (code != kOk) {
        complete(*nf, code, msg);
        cv_done_.notify_all();
        if (nf == &fl_[older]) continue;  // it was the only live set
        nf = nullptr;
      }
    }
    Flight& f = fl_[older];
    if (!f.live) continue;
    if (nf == &f) continue;  // the only live set is the one just submitted: pick another behind it first
    int code = kOk;
    std::string msg;
    try {
      finish_search(ctx_, f.prep);
    } catch (const Error& e) {
      code = e.code;
      msg = e.what();
    } catch (const std::exception& e) {
      code = kErrHip;
      msg = e.what();
    }
    complete(f, code, msg);
    cv_done_.notify_all();
    older ^= 1;
  }
}

}  // namespace cdk
