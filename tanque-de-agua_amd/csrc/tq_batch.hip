// tq_batch.hip -- independent tanks stepped together (the batched configuration).
//
// Every tank is a tq_sim with its own HIP stream.  A step hands the tanks to a fixed pool of
// host threads (tank k to thread k mod T, so one handle is only ever used by one thread); each
// thread enqueues its tanks' frames (avancar, fisica-cpu.js:155-159) one after another without
// waiting, then waits for their streams.  The device overlaps the tanks' kernels on its hardware
// queues: a small tank's kernels each fill only a few of the 256 CUs, so tanks run side by side
// instead of one after another.  With a fixed PCG iteration count (the reference's mode) a frame
// needs no host round trip; a residual-terminated solve polls its own stream, which only delays
// the tanks of the same thread.
#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>

#include "tq_internal.hpp"

struct tq_batch {
  std::vector<tq_sim*> tanks;
  int nthreads = 1;
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_go, cv_done;
  unsigned gen = 0;
  int pending = 0;
  bool quit = false;
  double dt = 0;
  std::vector<int> subs;
  int err = 0;
  std::string err_msg;
};

namespace {

void worker_main(tq_batch* b, int w) {
  unsigned seen = 0;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(b->mu);
      b->cv_go.wait(lk, [&] { return b->quit || b->gen != seen; });
      if (b->quit) return;
      seen = b->gen;
    }
    int e = 0;
    std::string msg;
    const size_t n = b->tanks.size();
    for (size_t k = (size_t)w; k < n && !e; k += (size_t)b->nthreads) {
      e = tq::step_enqueue(b->tanks[k], b->dt, b->subs[k]);
      if (e) msg = "tank " + std::to_string(k) + ": " + tq_last_error();
    }
    for (size_t k = (size_t)w; k < n; k += (size_t)b->nthreads) {
      const int e2 = tq::stream_wait(b->tanks[k]);
      if (e2 && !e) {
        e = e2;
        msg = "tank " + std::to_string(k) + ": " + tq_last_error();
      }
    }
    std::lock_guard<std::mutex> lk(b->mu);
    if (e && !b->err) {
      b->err = e;
      b->err_msg = msg;
    }
    if (--b->pending == 0) b->cv_done.notify_one();
  }
}

void stop_workers(tq_batch* b) {
  {
    std::lock_guard<std::mutex> lk(b->mu);
    b->quit = true;
  }
  b->cv_go.notify_all();
  for (auto& t : b->workers)
    if (t.joinable()) t.join();
  b->workers.clear();
}

}  // namespace

extern "C" {

int tq_batch_create(const tq_desc* descs, int n, int threads, tq_batch** out) {
  if (!out || !descs || n <= 0) {
    tq::set_error("tq_batch_create: bad arguments");
    return TQ_E_ARG;
  }
  *out = nullptr;
  tq_batch* b = new tq_batch();
  for (int k = 0; k < n; k++) {
    tq_sim* s = nullptr;
    if (int e = tq_create(&descs[k], &s)) {
      const std::string msg = "tank " + std::to_string(k) + ": " + tq_last_error();
      for (tq_sim* t : b->tanks) tq_destroy(t);
      delete b;
      tq::set_error(msg);
      return e;
    }
    tq::drop_second_stream(s);
    b->tanks.push_back(s);
  }
  b->nthreads = std::max(1, std::min(threads, std::min(n, 64)));
  b->subs.assign((size_t)n, 1);
  for (int w = 0; w < b->nthreads; w++) b->workers.emplace_back(worker_main, b, w);
  *out = b;
  return 0;
}

int tq_batch_size(tq_batch* b, int* n) {
  if (!b || !n) {
    tq::set_error("tq_batch_size: null argument");
    return TQ_E_ARG;
  }
  *n = (int)b->tanks.size();
  return 0;
}

tq_sim* tq_batch_tank(tq_batch* b, int k) {
  if (!b || k < 0 || k >= (int)b->tanks.size()) return nullptr;
  return b->tanks[(size_t)k];
}

int tq_batch_step(tq_batch* b, double dt, const int* substeps) {
  if (!b || !substeps || !(dt > 0)) {
    tq::set_error("tq_batch_step: bad arguments");
    return TQ_E_ARG;
  }
  for (size_t k = 0; k < b->tanks.size(); k++) {
    if (substeps[k] < 0) {
      tq::set_error("tq_batch_step: negative substep count");
      return TQ_E_ARG;
    }
    if (b->tanks[k]->tape) {
      // a recorded tank's frames go through tq_step, which logs them
      tq::set_error("tq_batch_step: a tank is recording frames (tq_tape_begin)");
      return TQ_E_STATE;
    }
  }
  std::unique_lock<std::mutex> lk(b->mu);
  b->dt = dt;
  b->subs.assign(substeps, substeps + b->tanks.size());
  b->err = 0;
  b->err_msg.clear();
  b->pending = b->nthreads;
  b->gen++;
  b->cv_go.notify_all();
  b->cv_done.wait(lk, [&] { return b->pending == 0; });
  if (b->err) {
    tq::set_error(b->err_msg);
    return b->err;
  }
  return 0;
}

int tq_batch_destroy(tq_batch* b) {
  if (!b) return 0;
  stop_workers(b);
  for (tq_sim* s : b->tanks) tq_destroy(s);
  delete b;
  return 0;
}

}  // extern "C"
