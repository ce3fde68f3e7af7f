// bem_prof.hip -- kernel-time registry: when enabled, a pair of HIP events
// brackets every launch of the tracked kernels on the stream they run on,
// and each launch carries its algorithmic work (flops, element pairs or
// points, as the caller states).  Read back through bh_prof_read.
#include <hip/hip_runtime.h>
#include <mutex>
#include <vector>

#include "../../include/bem3_hip.h"

namespace {
struct Rec {
  int kid;
  double work, bytes;
  hipEvent_t a, b;
  bool closed;
};
struct Prof {
  bool on = false;
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
};
Prof P;
std::mutex M;
// the records this thread opened and has not closed (threads driving
// different streams time their kernels independently)
thread_local std::vector<size_t> t_open;

hipEvent_t take()
{
  if (P.used < P.pool.size()) return P.pool[P.used++];
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  P.pool.push_back(e);
  P.used++;
  return e;
}
}  // namespace

extern "C" {

// whether the registry records launches (its events cannot go into a
// captured graph)
int bem_prof_on(void) { return P.on ? 1 : 0; }

// destroys the event pool (library shutdown)
void bem_prof_release(void)
{
  std::lock_guard<std::mutex> lk(M);
  for (hipEvent_t e : P.pool) (void)hipEventDestroy(e);
  P.pool.clear();
  P.recs.clear();
  P.used = 0;
  P.on = false;
}

// bytes: the launch's algorithmic HBM traffic (operands read once, results
// written once), where the caller states it (0 otherwise)
void bem_prof_begin_b(int kid, double work, double bytes, hipStream_t st)
{
  if (!P.on || kid < 0 || kid >= BH_K_COUNT) return;
  std::lock_guard<std::mutex> lk(M);
  Rec r{kid, work, bytes, take(), take(), false};
  if (!r.a || !r.b) return;
  (void)hipEventRecord(r.a, st);
  t_open.push_back(P.recs.size());
  P.recs.push_back(r);
}

void bem_prof_begin(int kid, double work, hipStream_t st) { bem_prof_begin_b(kid, work, 0.0, st); }

void bem_prof_end(int kid, hipStream_t st)
{
  if (!P.on) return;
  std::lock_guard<std::mutex> lk(M);
  if (t_open.empty()) return;
  const size_t i = t_open.back();
  t_open.pop_back();
  if (i >= P.recs.size() || P.recs[i].kid != kid) return;
  (void)hipEventRecord(P.recs[i].b, st);
  P.recs[i].closed = true;
}

int bh_prof_enable(int on)
{
  std::lock_guard<std::mutex> lk(M);
  P.on = on != 0;
  P.recs.clear();
  P.used = 0;
  t_open.clear();
  return 0;
}

int bh_prof_read_bytes(int kid, double *ms, double *work, double *bytes, int *launches)
{
  if (kid < 0 || kid >= BH_K_COUNT) return BH_EARG;
  std::lock_guard<std::mutex> lk(M);
  double t = 0.0, w = 0.0, by = 0.0;
  int n = 0;
  for (const Rec &r : P.recs) {
    if (r.kid != kid || !r.closed) continue;
    if (hipEventSynchronize(r.b) != hipSuccess) return BH_EHIP;
    float e = 0.0f;
    if (hipEventElapsedTime(&e, r.a, r.b) != hipSuccess) return BH_EHIP;
    t += e;
    w += r.work;
    by += r.bytes;
    n++;
  }
  if (ms) *ms = t;
  if (work) *work = w;
  if (bytes) *bytes = by;
  if (launches) *launches = n;
  return 0;
}

int bh_prof_read(int kid, double *ms, double *work, int *launches)
{
  return bh_prof_read_bytes(kid, ms, work, nullptr, launches);
}
}
