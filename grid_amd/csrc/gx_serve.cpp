// Solve serving: a trace of Wilson solve requests replayed against resident gauge fields.
// Admission is FIFO, by idle slot and by device-memory footprint against a budget; every slot
// owns a stream (gx_ctx) and a host thread, so admitted solves overlap on the device while new
// requests keep arriving (continuous batching at request granularity; the sources of one request
// are solved together as one multi-right-hand-side solve).  No Grid counterpart: Grid drives one
// solve at a time from one host thread (ConjugateGradient.h:68-257).
//
// Memory: the operators and fields of a request stay cached on the slot that ran it, keyed by
// (lattice, nsrc), so a later request of the same shape allocates nothing (hipFree synchronises
// the whole device and would stall the other slots).  The budget covers every cached entry;
// entries of idle slots are evicted, least recently used first, when an admission needs room.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/gx.h"

namespace gx {
void set_error(const char *fmt, ...);  // gx_last_error() text (gx_api.hip)
// bytes of the multi-RHS CG's per-tile slice sums over H4 half sites of Ls slices (gx_internal.h)
size_t hop_norm_bytes(int64_t H4, int Ls);
}

namespace {

using Clock = std::chrono::steady_clock;

struct Lattice {
  int d[4];
  bool operator<(const Lattice &o) const { return std::memcmp(d, o.d, sizeof(d)) < 0; }
  bool operator==(const Lattice &o) const { return std::memcmp(d, o.d, sizeof(d)) == 0; }
};

// Half-field count of a multi-RHS red-black solve on a cloned Wilson operator (FP64 halves):
// operator temporaries 0-3, CG vectors 4-6, the red-black halves 10-14, the residual check's
// full field (2), source and solution (4), and with mixed precision the FP64 outer vectors 7-8
// plus nine inner-precision halves: FP32 (mixed = 1, 4.5 FP64 halves) or FP16 storage (mixed =
// 2, 52 of every 192 bytes: 12 binary16 pairs and the scale per entry), the operator's per-slice
// Schur coefficients, and (nsrc >= 2) the multi-RHS CG's per-tile slice sums.
double footprint(const int d[4], int nsrc, int mixed) {
  const int64_t H4 = (int64_t)d[0] * d[1] * d[2] * d[3] / 2;
  const double hb = (double)H4 * nsrc * 12 * 16;
  const double coef = 48.0 * nsrc;  // the operator's per-slice coefficients (4 doubles and 4 floats)
  const double tiles = nsrc >= 2 ? (double)gx::hop_norm_bytes(H4, nsrc) : 0.0;
  if (mixed == 2) return hb * (20.0 + 9.0 * 52.0 / 192.0) + coef + tiles;
  return hb * (mixed ? 24.5 : 18.0) + coef + tiles;
}

struct Entry {
  Lattice lat{};
  int nsrc = 0;             // right-hand sides of the cached operator and fields
  gx_op op = nullptr;
  gx_field src = nullptr, sol = nullptr;
  gx_op base = nullptr;     // links the clone currently reads
  std::vector<double> masses;  // mass of every slice the operator was last set to
  double bytes = 0;
  uint64_t last_use = 0;
};

struct Slot {
  gx_ctx ctx = nullptr;
  std::thread th;
  std::vector<Entry> cache;
  int req = -1;             // first request of the assigned solve, -1 when idle
  std::vector<int> batch;   // the requests of the assigned solve (req first), slices in this order
  int nsrc = 0;             // right-hand sides of the assigned solve
  double reserve = 0;       // bytes reserved for a new cache entry of the assigned solve
};

void destroy_entry(Entry &e) {
  gx_field_destroy(e.src);
  gx_field_destroy(e.sol);
  gx_op_destroy(e.op);
  e = Entry();
}

struct Server {
  const gx_serve_params &p;
  const gx_request *req;
  int n;
  gx_response *resp;
  double **solutions;
  std::map<std::pair<Lattice, int>, gx_op> bases;
  std::vector<Slot> slots;
  std::mutex mu;
  std::condition_variable cv_disp, cv_slot;
  double in_use = 0;
  int done = 0;
  bool stop = false;
  uint64_t tick = 0;
  Clock::time_point t0;

  Server(const gx_serve_params &p_, const gx_request *r, int n_, gx_response *o, double **sol)
      : p(p_), req(r), n(n_), resp(o), solutions(sol), slots(p_.slots) {}

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

  // The assigned solve of slot k: its requests' sources side by side in the slices of one
  // multi-RHS field (request b at slices off_b .. off_b + nsrc_b - 1, mass m_b on those slices),
  // one red-black solve, then each request's statistics, digest and solution.
  int run(int k) {
    Slot &s = slots[k];
    const gx_request &q = req[s.batch[0]];
    Lattice lat;
    std::memcpy(lat.d, q.dims, sizeof(lat.d));
    gx_op base = bases.at({lat, q.gauge});
    const int N = s.nsrc;
    std::vector<double> masses;
    for (int r : s.batch)
      for (int j = 0; j < req[r].nsrc; j++) masses.push_back(req[r].mass);
    Entry *e = nullptr;
    for (auto &c : s.cache)
      if (c.lat == lat && c.nsrc == N) e = &c;
    int rc = GX_OK;
    if (!e) {
      Entry ne;
      ne.lat = lat;
      ne.nsrc = N;
      ne.bytes = s.reserve;
      rc = gx_wilson_clone(base, s.ctx, N, q.mass, &ne.op);
      if (rc == GX_OK) rc = gx_field_create(s.ctx, q.dims, N, GX_F64, GX_FULL, &ne.src);
      if (rc == GX_OK) rc = gx_field_create(s.ctx, q.dims, N, GX_F64, GX_FULL, &ne.sol);
      if (rc != GX_OK) {
        destroy_entry(ne);
        return rc;
      }
      ne.base = base;
      ne.masses.assign(N, q.mass);
      s.cache.push_back(ne);
      e = &s.cache.back();
    }
    if (e->base != base) {
      if ((rc = gx_op_share_links(e->op, base))) return rc;
      e->base = base;
    }
    if (e->masses != masses) {
      bool uniform = true;
      for (double m : masses) uniform = uniform && m == masses[0];
      rc = uniform ? gx_op_set_mass(e->op, masses[0]) : gx_op_set_slice_masses(e->op, N, masses.data());
      if (rc) return rc;
      e->masses = masses;
    }
    int off = 0;
    for (int r : s.batch) {
      if ((rc = gx_field_gaussian_slices(e->src, req[r].seed, off, req[r].nsrc))) return rc;
      off += req[r].nsrc;
    }
    gx_solver_stats st{};
    double res = 0;
    const int src = gx_schur_solve(e->op, e->src, e->sol, p.tol, p.maxit, p.mixed, &st, &res);
    if (src != GX_OK && src != GX_ENOCONV) return src;
    std::vector<double> all;  // the whole solution on the host, when a fused request wants its own
    const size_t V = (size_t)q.dims[0] * q.dims[1] * q.dims[2] * q.dims[3];
    off = 0;
    for (int r : s.batch) {
      const int n = req[r].nsrc;
      gx_response &o = resp[r];
      o.iterations = st.iterations;
      o.outer_iterations = st.outer_iterations;
      o.residual = res;
      o.solver_residual = st.true_residual;
      o.batch = (int)s.batch.size();
      if ((rc = gx_field_digest_slices(e->sol, off, n, &o.digest))) return rc;
      if (solutions && solutions[r]) {
        if (s.batch.size() == 1) {
          if ((rc = gx_field_export(e->sol, solutions[r]))) return rc;
        } else {
          if (all.empty()) {
            all.resize(V * N * 24);
            if ((rc = gx_field_export(e->sol, all.data()))) return rc;
          }
          for (size_t x = 0; x < V; x++)
            std::memcpy(solutions[r] + x * n * 24, all.data() + (x * N + off) * 24, sizeof(double) * 24 * n);
        }
      }
      off += n;
    }
    return src;
  }

  void slot_loop(int k) {
    (void)hipSetDevice(p.device);
    Slot &s = slots[k];
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      cv_slot.wait(lk, [&] { return stop || s.req >= 0; });
      if (s.req < 0) return;
      const int r = s.req;
      const size_t before = s.cache.size();
      lk.unlock();
      const int rc = run(k);
      lk.lock();
      const double t = now();
      for (int b : s.batch) {
        resp[b].status = rc;
        resp[b].slot = k;
        resp[b].t_done = t;
        resp[b].batch = (int)s.batch.size();
      }
      if (s.cache.size() == before) in_use -= s.reserve;  // nothing new was cached
      for (auto &c : s.cache)
        if (c.lat == Lattice{{req[r].dims[0], req[r].dims[1], req[r].dims[2], req[r].dims[3]}} && c.nsrc == s.nsrc)
          c.last_use = ++tick;
      s.reserve = 0;
      s.req = -1;
      for (size_t i = 0; i < s.batch.size(); i++) note_done();
      s.batch.clear();
      cv_disp.notify_all();
    }
  }

  // One more request finished (called with mu held); with p.progress > 0, report every N of them.
  void note_done() {
    done++;
    if (p.progress > 0 && (done % p.progress == 0 || done == n)) {
      std::fprintf(stderr, "gx_serve: %d of %d requests done, %.1f s\n", done, n, now());
      std::fflush(stderr);
    }
  }

  // Evict cached entries of idle slots (least recently used first) until `need` bytes fit,
  // sparing entry `keep` of slot `keep_slot`.  Called with mu held.
  void evict_for(double need, int keep_slot, const Entry *keep) {
    while (in_use + need > p.budget) {
      int bs = -1;
      size_t bi = 0;
      uint64_t best = UINT64_MAX;
      for (int k = 0; k < (int)slots.size(); k++) {
        if (slots[k].req >= 0) continue;
        for (size_t i = 0; i < slots[k].cache.size(); i++) {
          const Entry &c = slots[k].cache[i];
          if (k == keep_slot && &c == keep) continue;
          if (c.last_use < best) {
            best = c.last_use;
            bs = k;
            bi = i;
          }
        }
      }
      if (bs < 0) return;
      in_use -= slots[bs].cache[bi].bytes;
      destroy_entry(slots[bs].cache[bi]);
      slots[bs].cache.erase(slots[bs].cache.begin() + bi);
    }
  }

  int serve(double *wall) {
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return req[a].arrival < req[b].arrival; });
    std::deque<int> pending;
    int next = 0;
    t0 = Clock::now();
    std::unique_lock<std::mutex> lk(mu);
    while (done < n) {
      const double t = now();
      while (next < n && (!p.realtime || req[order[next]].arrival <= t)) pending.push_back(order[next++]);
      while (!pending.empty()) {
        const int r = pending.front();
        const gx_request &q = req[r];
        const Lattice lat{{q.dims[0], q.dims[1], q.dims[2], q.dims[3]}};
        // with fusion, waiting requests of the same lattice and gauge field join the head's solve
        // (FIFO order among them) while the right-hand sides stay within fuse_rhs
        std::vector<int> batch{r};
        int N = q.nsrc;
        if (p.fuse_rhs > 0)
          for (size_t i = 1; i < pending.size(); i++) {
            const gx_request &o = req[pending[i]];
            if (std::memcmp(o.dims, q.dims, sizeof(q.dims)) == 0 && o.gauge == q.gauge && N + o.nsrc <= p.fuse_rhs &&
                footprint(q.dims, N + o.nsrc, p.mixed) <= p.budget) {
              batch.push_back(pending[i]);
              N += o.nsrc;
            }
          }
        // an idle slot that has this shape cached needs no new memory; otherwise the idle slot
        // with the least cached memory takes it
        int k = -1;
        const Entry *hit = nullptr;
        double least = 0;
        for (int j = 0; j < (int)slots.size() && !hit; j++) {
          if (slots[j].req >= 0) continue;
          double mine = 0;
          for (auto &c : slots[j].cache) {
            mine += c.bytes;
            if (c.lat == lat && c.nsrc == N) hit = &c;
          }
          if (hit || k < 0 || mine < least) {
            k = j;
            least = mine;
          }
        }
        if (k < 0) break;  // no idle slot
        const double need = hit ? 0 : footprint(q.dims, N, p.mixed);
        if (need > p.budget) {  // the head alone can never fit (members join only within the budget): refuse it
          resp[r].status = GX_ENOMEM;
          resp[r].slot = -1;
          resp[r].footprint = need;
          resp[r].t_admit = resp[r].t_done = t;
          pending.pop_front();
          note_done();
          continue;
        }
        if (in_use + need > p.budget) evict_for(need, k, hit);
        if (in_use + need > p.budget) break;  // wait for running solves to finish
        in_use += need;
        Slot &sl = slots[k];
        sl.reserve = need;
        sl.batch = batch;
        sl.nsrc = N;
        sl.req = r;
        for (int b : batch) {
          resp[b].t_admit = t;
          resp[b].footprint = (hit ? hit->bytes : need) * req[b].nsrc / N;
        }
        for (size_t i = batch.size(); i-- > 1;)
          pending.erase(std::find(pending.begin(), pending.end(), batch[i]));
        pending.pop_front();
        cv_slot.notify_all();
      }
      if (done >= n) break;
      if (next < n && p.realtime && pending.empty()) {
        const double dt = req[order[next]].arrival - now();
        if (dt > 0) cv_disp.wait_for(lk, std::chrono::duration<double>(dt));
      } else if (next < n && p.realtime) {
        const double dt = req[order[next]].arrival - now();
        cv_disp.wait_for(lk, std::chrono::duration<double>(dt > 0 ? dt : 0));
      } else {
        cv_disp.wait(lk);
      }
    }
    *wall = now();
    return GX_OK;
  }
};

}  // namespace

extern "C" {

int gx_serve_footprint(const int dims[4], int nsrc, int mixed, double *bytes) {
  if (!dims || !bytes || nsrc < 1) {
    gx::set_error("serve_footprint: bad arguments");
    return GX_EINVAL;
  }
  *bytes = footprint(dims, nsrc, mixed);
  return GX_OK;
}

int gx_serve(const gx_serve_params *p, const gx_request *req, int n, gx_response *resp, double **solutions,
             double *wall, double *resident) {
  if (!p || !resp || !wall || n < 0 || (n > 0 && !req) || p->slots < 1 || p->n_gauge < 1 || p->budget <= 0 ||
      p->fuse_rhs < 0 || p->fuse_rhs > 64) {
    gx::set_error("serve: bad parameters (slots, n_gauge and budget must be positive, fuse_rhs 0..64)");
    return GX_EINVAL;
  }
  for (int i = 0; i < n; i++) {
    const gx_request &q = req[i];
    bool ok = q.nsrc >= 1 && q.nsrc <= 16 && q.gauge >= 0 && q.gauge < p->n_gauge;
    for (int d = 0; d < 4; d++) ok = ok && q.dims[d] >= 2 && !(q.dims[d] & 1);
    if (!ok) {
      gx::set_error("serve: request %d invalid (even dims >= 2, nsrc 1..16, gauge < n_gauge)", i);
      return GX_EINVAL;
    }
  }
  std::memset(resp, 0, sizeof(gx_response) * (size_t)n);
  if (hipSetDevice(p->device) != hipSuccess) return GX_EHIP;
  Server sv(*p, req, n, resp, solutions);
  gx_ctx loader = nullptr;
  int rc = gx_ctx_create(p->device, &loader);
  // resident gauge fields of every lattice in the trace, each with FP32 and FP16 copies of the links
  double res_bytes = 0;
  for (int i = 0; i < n && rc == GX_OK; i++) {
    Lattice lat{{req[i].dims[0], req[i].dims[1], req[i].dims[2], req[i].dims[3]}};
    if (sv.bases.count({lat, 0})) continue;
    const size_t V = (size_t)lat.d[0] * lat.d[1] * lat.d[2] * lat.d[3];
    std::vector<double> U(V * 4 * 18);
    for (int g = 0; g < p->n_gauge && rc == GX_OK; g++) {
      if (p->gauge_dir) {  // a real ensemble: NERSC file, checked against its header
        char path[4096];
        std::snprintf(path, sizeof(path), "%s/ckpoint_lat.%dx%dx%dx%d.%d", p->gauge_dir, lat.d[0], lat.d[1],
                      lat.d[2], lat.d[3], g);
        rc = gx_nersc_read(loader, path, lat.d, U.data(), nullptr);
      } else {
        rc = gx_random_gauge(lat.d, p->gauge_seed + (uint64_t)g, p->gauge_eps, U.data());
      }
      gx_op base = nullptr, warm = nullptr;
      if (rc == GX_OK) rc = gx_wilson_create(loader, lat.d, 1, 0.1, U.data(), nullptr, &base);
      if (rc != GX_OK) break;
      sv.bases[{lat, g}] = base;
      rc = gx_wilson_clone(base, loader, 1, 0.1, &warm);  // builds the FP32 and FP16 links
      gx_op_destroy(warm);
      size_t b = 0;
      if (rc == GX_OK) rc = gx_op_footprint(base, &b);
      res_bytes += (double)b;
    }
  }
  for (int k = 0; k < p->slots && rc == GX_OK; k++) rc = gx_ctx_create(p->device, &sv.slots[k].ctx);
  if (rc == GX_OK) {
    for (int k = 0; k < p->slots; k++) sv.slots[k].th = std::thread([&sv, k] { sv.slot_loop(k); });
    rc = sv.serve(wall);
    {
      std::lock_guard<std::mutex> lk(sv.mu);
      sv.stop = true;
    }
    sv.cv_slot.notify_all();
    for (auto &s : sv.slots) s.th.join();
  }
  for (auto &s : sv.slots) {
    for (auto &e : s.cache) destroy_entry(e);
    gx_ctx_destroy(s.ctx);
  }
  for (auto &b : sv.bases) gx_op_destroy(b.second);
  gx_ctx_destroy(loader);
  if (resident) *resident = res_bytes;
  return rc;
}

}  // extern "C"
