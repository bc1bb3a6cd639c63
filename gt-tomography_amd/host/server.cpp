// Reconstruction server: a queue of FDK / OS-SART requests served by worker
// lanes on one GPU (include/gtt_c.h, gtt_server_*).
//
// The reference runs one reconstruction per process (xray/FDK_reconstruct_3d.cpp,
// xray/cuCBOS_reconstruct.cpp); this is the batching layer over the same
// programs.  Each lane is a host thread with its own non-blocking HIP stream
// (default priority); every kernel, copy and reduction a request issues goes
// to that stream, and every operator of a request owns its own libcbct
// context, so requests on different lanes overlap on the GPU (small scans do
// not fill 256 CUs on their own) and share no state.  No kernel on the path
// uses atomics or cross-request reductions, so a request's result does not
// depend on what runs beside it.
//
// Admission: a request's HBM footprint is estimated from its sizes at submit;
// the head of the FIFO starts when a lane is free and the footprints of the
// running requests, the memory idle lanes keep, and its own fit the budget
// (no overtaking, so the order of starts is the order of arrival).
//
// A lane keeps its device memory between requests (its pooled projector
// contexts, their shared scratch, the freed blocks of its arrays): the next
// request of similar size reuses it with no allocation and no device-wide
// synchronisation.  What a lane keeps is counted against the budget while it
// is idle; when the head of the queue does not fit, idle lanes free theirs.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "gtt/fdk.h"
#include "gtt/operators.h"
#include "gtt/solvers.h"
#include "gtt_c.h"
#include "scan.h"

using namespace gtt;

namespace {

typedef std::chrono::steady_clock Clock;

struct Job {
  long long id = 0;
  gtt_request req{};
  size_t footprint = 0;
  Clock::time_point t_submit, t_start, t_end;
  bool done = false;
  int status = 0;
  double scale = 0.0;
  bool autotune = false;   // tune the forward projector for the scan's geometry first
  bool graphs = false;     // OS-SART iterations 2.. replayed as captured HIP graphs
  std::string message;
};

double ms(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// Device bytes a request may hold at once, in FP32 elements (BF16 working
// copies need less), with slack:
//  OS-SART: the iterate, the update image and S column normalisers
//    (S + 2 volumes); the uploaded projections and their subset-ordered copy,
//    the row normalisers and the residual (4 sinograms); the scratch the
//    subset operators share: the padded volume copy (up to 1.3 volumes for
//    small grids) and the padded copy of one subset's projections (up to 4x
//    a subset sinogram as quad texels).
//  FDK: the image and the padded volume copy (2.3 volumes); the projections,
//    the filtered projections, A f, the filter staging and the padded copy of
//    all views (up to 4x as quad texels): 8 sinograms.
size_t footprint_bytes(const gtt_request &r) {
  const size_t nvox = (size_t)r.geom.nx * r.geom.ny * r.geom.nz;
  const size_t nproj = (size_t)(r.geom.nu + 3) * (r.geom.nv + 3) * (size_t)r.nviews;
  const size_t slack = size_t(64) << 20;
  // the LDS-staged backprojector's rectangle table: 8 bytes per 16 x 16 x 8
  // (or x 16) voxel block and view of one backprojection call
  const size_t blocks = (size_t)((r.geom.nx + 15) / 16) * ((r.geom.ny + 15) / 16) * ((r.geom.nz + 7) / 8);
  if (r.kind == GTT_REQ_FDK) return 4 * (nvox * 23 / 10 + 8 * nproj) + 8 * blocks * (size_t)r.nviews + slack;
  const size_t S = r.subsets > 0 ? (size_t)r.subsets : 1;
  return 4 * (nvox * (10 * S + 33) / 10 + 4 * nproj + 4 * ((nproj + S - 1) / S)) +
         8 * blocks * (((size_t)r.nviews + S - 1) / S) + slack;
}

// The request's storage type is the lane's default while its operators are
// built (they keep it); restored afterwards.
struct StorageScope {
  int saved;
  explicit StorageScope(int dtype) : saved(default_storage()) { set_default_storage(dtype); }
  ~StorageScope() { set_default_storage(saved); }
};

void run_request(Job &j) {
  const gtt_request &r = j.req;
  StorageScope storage(r.storage == CBCT_BF16 ? CBCT_BF16 : CBCT_F32);
  // the caller's projections are only read (uploaded once), never copied
  Scan s = make_scan(&r.geom, r.angles_deg, r.offsets, r.nviews, nullptr, nullptr, 0, r.projections, true);
  if (r.kind == GTT_REQ_FDK) {
    std::shared_ptr<std::vector<float>> F;
    if (r.filter) F = std::make_shared<std::vector<float>>(r.filter, r.filter + r.geom.nu + 1);
    const float spp = r.samples_per_pixel > 0 ? r.samples_per_pixel : 1.5f;
    FDKResult out = FDK_reconstruct_3d(s.acq, s.binning, s.vol_dims, s.dims_mm, F, spp);
    std::memcpy(r.volume_out, out.image->get_data_ptr(), out.image->get_number_of_bytes());
    j.scale = out.scale;
  } else {
    if (r.subsets <= 0) throw std::runtime_error("number of subsets must be positive");
    auto E = std::make_shared<CBSubsetOperator>(r.subsets);
    E->setup(s.acq, s.binning, s.dims_mm, false);
    E->set_domain_dimensions(&s.vol_dims);
    if (j.autotune) E->autotune();
    cuNDArray<float> b = upload_in_subset_order(*s.acq->get_projections(), E->get_permutation());
    osSARTSolver<cuNDArray<float>> solver;
    solver.set_encoding_operator(E);
    solver.set_max_iterations(r.iterations);
    solver.set_beta(r.beta);
    solver.set_gamma(r.gamma);
    solver.set_non_negativity_constraint(r.nonneg != 0);
    solver.set_graphs(j.graphs);
    auto x = solver.solve(&b);
    copy_d2h(r.volume_out, x->get_data_ptr(), x->get_number_of_bytes());
  }
}

}  // namespace

struct gtt_server_s {
  int device = 0;
  bool autotune = false;
  bool graphs = false;
  size_t budget = 0;
  size_t in_use = 0;
  size_t kept = 0;                 // memory idle lanes keep (sum over lanes)
  int running = 0;
  bool stop = false;
  long long next_id = 1;
  std::mutex mu;
  std::condition_variable cv;  // queue changes, completions, stop
  std::deque<std::shared_ptr<Job>> queue;
  std::map<long long, std::shared_ptr<Job>> jobs;
  std::vector<std::thread> lanes;
  std::string err;

  bool head_fits() const {
    return !queue.empty() && in_use + kept + queue.front()->footprint <= budget;
  }

  void lane_main() {
    hipStream_t st = nullptr;
    std::string init_err;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
      init_err = "gtt_server: cannot create the lane's HIP stream";
    set_current_stream(st);
    size_t my_kept = 0;   // this lane's share of `kept`
    for (;;) {
      std::shared_ptr<Job> j;
      bool trim_first = false;
      {
        std::unique_lock<std::mutex> lk(mu);
        // queued work drains before a stop takes effect; a lane that keeps
        // memory wakes to free it when the head does not fit otherwise
        cv.wait(lk, [&] {
          return (stop && queue.empty()) || head_fits() || (!queue.empty() && my_kept > 0);
        });
        if (!head_fits()) {
          if (queue.empty()) break;
          lk.unlock();
          trim_thread_memory();
          lk.lock();
          kept -= my_kept;
          my_kept = 0;
          cv.notify_all();
          continue;
        }
        j = queue.front();
        queue.pop_front();
        // the memory this lane keeps becomes part of the request's own; more
        // than the request's footprint is freed first, so that the request
        // holds no more than its footprint says
        trim_first = my_kept > j->footprint;
        kept -= my_kept;
        my_kept = 0;
        in_use += j->footprint;
        ++running;
        j->t_start = Clock::now();
      }
      if (trim_first) trim_thread_memory();
      int status = 0;
      std::string msg;
      if (!init_err.empty()) {
        status = -1;
        msg = init_err;
      } else {
        try {
          run_request(*j);
          if (hipStreamSynchronize(st) != hipSuccess) throw std::runtime_error("gtt_server: stream error");
        } catch (const std::exception &e) {
          status = -1;
          msg = e.what();
        }
        if (status) trim_thread_memory();   // after a failure, start clean
      }
      const size_t keep = init_err.empty() ? retained_device_bytes() : 0;
      {
        std::lock_guard<std::mutex> lk(mu);
        j->t_end = Clock::now();
        j->status = status;
        j->message = msg;
        j->done = true;
        in_use -= j->footprint;
        my_kept = keep;
        kept += keep;
        --running;
      }
      cv.notify_all();
    }
    release_thread_resources();
    {
      std::lock_guard<std::mutex> lk(mu);
      kept -= my_kept;
    }
    if (st) (void)hipStreamDestroy(st);
  }
};

extern "C" {

int gtt_server_create(int device, int lanes, size_t hbm_budget_bytes, gtt_server *out) {
  if (!out) return -1;
  *out = nullptr;
  if (lanes <= 0 || lanes > 64) return -1;
  auto *s = new gtt_server_s();
  s->device = device;
  size_t budget = hbm_budget_bytes;
  if (budget == 0) {
    size_t free_b = 0, total_b = 0;
    if (hipSetDevice(device) != hipSuccess || hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
      delete s;
      return -1;
    }
    budget = free_b / 4 * 3;
  }
  s->budget = budget;
  for (int i = 0; i < lanes; ++i) s->lanes.emplace_back([s] { s->lane_main(); });
  *out = s;
  return 0;
}

int gtt_server_destroy(gtt_server s) {
  if (!s) return -1;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  for (auto &t : s->lanes) t.join();
  delete s;
  return 0;
}

const char *gtt_server_last_error(gtt_server s) { return s ? s->err.c_str() : "null server"; }

size_t gtt_request_footprint(const gtt_request *req) { return req ? footprint_bytes(*req) : 0; }

int gtt_server_set_graphs(gtt_server s, int on) {
  if (!s) return -1;
  std::lock_guard<std::mutex> lk(s->mu);
  s->graphs = on != 0;
  return 0;
}

int gtt_server_set_autotune(gtt_server s, int on) {
  if (!s) return -1;
  std::lock_guard<std::mutex> lk(s->mu);
  s->autotune = on != 0;
  return 0;
}

int gtt_server_submit(gtt_server s, const gtt_request *req, long long *id) {
  if (!s) return -1;
  std::lock_guard<std::mutex> lk(s->mu);
  if (!req || !id || !req->projections || !req->volume_out || !req->angles_deg || req->nviews <= 0) {
    s->err = "gtt_server_submit: illegal request (null array or no views)";
    return -1;
  }
  if (req->kind != GTT_REQ_FDK && req->kind != GTT_REQ_OS_SART) {
    s->err = "gtt_server_submit: unknown request kind";
    return -1;
  }
  if (req->storage != CBCT_F32 && req->storage != CBCT_BF16) {
    s->err = "gtt_server_submit: storage must be CBCT_F32 or CBCT_BF16";
    return -1;
  }
  if (s->stop) {
    s->err = "gtt_server_submit: server is stopping";
    return -1;
  }
  auto j = std::make_shared<Job>();
  j->req = *req;
  j->autotune = s->autotune;
  j->graphs = s->graphs;
  j->footprint = footprint_bytes(*req);
  if (j->footprint > s->budget) {
    s->err = "gtt_server_submit: request needs " + std::to_string(j->footprint) +
             " bytes of device memory, more than the budget of " + std::to_string(s->budget);
    return -1;
  }
  j->id = s->next_id++;
  j->t_submit = Clock::now();
  s->jobs[j->id] = j;
  s->queue.push_back(j);
  *id = j->id;
  s->cv.notify_all();
  return 0;
}

namespace {
// Wait for request id (without bound when timeout_ms < 0); GTT_WAIT_TIMEOUT
// leaves it registered so it can be waited for again.
int wait_request(gtt_server s, long long id, long long timeout_ms, gtt_result *res, const char *who) {
  if (!s) return -1;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) {
    s->err = std::string(who) + ": unknown request id " + std::to_string(id);
    return -1;
  }
  std::shared_ptr<Job> j = it->second;
  if (timeout_ms < 0) {
    s->cv.wait(lk, [&] { return j->done; });
  } else if (!s->cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return j->done; })) {
    s->err = std::string(who) + ": request " + std::to_string(id) + " not finished after " +
             std::to_string(timeout_ms) + " ms";
    return GTT_WAIT_TIMEOUT;
  }
  s->jobs.erase(id);
  if (res) {
    res->status = j->status;
    res->queue_ms = ms(j->t_submit, j->t_start);
    res->run_ms = ms(j->t_start, j->t_end);
    res->latency_ms = ms(j->t_submit, j->t_end);
    res->scale = j->scale;
    std::snprintf(res->message, sizeof(res->message), "%s", j->message.c_str());
  }
  if (j->status) s->err = j->message;
  return j->status;
}
}  // namespace

int gtt_server_wait(gtt_server s, long long id, gtt_result *res) {
  return wait_request(s, id, -1, res, "gtt_server_wait");
}

int gtt_server_wait_for(gtt_server s, long long id, long long timeout_ms, gtt_result *res) {
  if (timeout_ms < 0) {
    if (s) s->err = "gtt_server_wait_for: negative time limit";
    return -1;
  }
  return wait_request(s, id, timeout_ms, res, "gtt_server_wait_for");
}

}  // extern "C"
