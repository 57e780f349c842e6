// Test-only driver of the host-stub build: runs the library's real host code (bemfmm.cpp, server.cpp, the
// plan) against the stub HIP runtime and CPU kernels, with several contexts at once, under AddressSanitizer
// or ThreadSanitizer. Problems come from a file written by tests/test_host_stub.py:
//   int32 count; per problem: int32 n_nodes, n_faces, n_dof, n_stages; double X[n_stages][3 n_nodes];
//   int32 tri[3 n_faces], face_dof[3 n_faces], kind[n_dof], point[n_dof]; double phi[n_stages][n_dof],
//   phin[n_stages][n_dof]
// Every problem is a session of n_stages BVPs (moved nodes between stages). Each session is solved alone
// first (one context, one thread); then all sessions run at once, either on one thread per session with a
// context each (mode "threads") or through the request loop (mode "server"). Every result in the mix must
// equal its result alone to 1e-10 relative, and the stub must record no API-usage violation.
// Usage: driver <problems file> <threads|server> <fp64|fp32> <slots> [p]   |   driver selftest
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bemfmm.h"
#include "../../include/bemfmm_server.h"
#include "hipstub.hpp"

namespace {

struct Problem {
  int n_nodes = 0, n_faces = 0, n_dof = 0, n_stages = 0;
  std::vector<std::vector<double>> X, phi, phin;
  std::vector<int32_t> tri, face_dof, kind, point;
};

template <typename T>
void rd(FILE* f, T* p, size_t n) {
  if (std::fread(p, sizeof(T), n, f) != n) {
    std::fprintf(stderr, "driver: short read\n");
    std::exit(2);
  }
}

std::vector<Problem> load(const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    std::fprintf(stderr, "driver: cannot open %s\n", path);
    std::exit(2);
  }
  int32_t n = 0;
  rd(f, &n, 1);
  std::vector<Problem> ps(n);
  for (auto& p : ps) {
    int32_t h[4];
    rd(f, h, 4);
    p.n_nodes = h[0];
    p.n_faces = h[1];
    p.n_dof = h[2];
    p.n_stages = h[3];
    p.X.assign(p.n_stages, std::vector<double>(3 * (size_t)p.n_nodes));
    for (auto& x : p.X) rd(f, x.data(), x.size());
    p.tri.resize(3 * (size_t)p.n_faces);
    p.face_dof.resize(3 * (size_t)p.n_faces);
    p.kind.resize(p.n_dof);
    p.point.resize(p.n_dof);
    rd(f, p.tri.data(), p.tri.size());
    rd(f, p.face_dof.data(), p.face_dof.size());
    rd(f, p.kind.data(), p.kind.size());
    rd(f, p.point.data(), p.point.size());
    p.phi.assign(p.n_stages, std::vector<double>(p.n_dof));
    p.phin.assign(p.n_stages, std::vector<double>(p.n_dof));
    for (auto& x : p.phi) rd(f, x.data(), x.size());
    for (auto& x : p.phin) rd(f, x.data(), x.size());
  }
  std::fclose(f);
  return ps;
}

bemfmm_opts options(bool fp32, int p) {
  bemfmm_opts o;
  bemfmm_default_opts(&o);
  o.p = p;
  o.leaf_max = 32;
  o.precision = fp32 ? BEMFMM_PRECISION_FP32_REFINE : BEMFMM_PRECISION_FP64;
  o.preconditioner = BEMFMM_PRECOND_SCHWARZ;
  o.use_graphs = 1;  // the capture path (and its gate against other threads' frees) is the one to exercise
  return o;
}

// one session on its own context: every stage re-prepares (moved nodes) and solves
int run_session(const Problem& p, const bemfmm_opts& o, std::vector<std::vector<double>>& out) {
  bemfmm_ctx* c = bemfmm_ctx_create(0, &o);
  if (!c) return -100;
  out.assign(p.n_stages, std::vector<double>(p.n_dof, 0.));
  std::vector<double> x0(p.n_dof, 0.);
  int rc = 0;
  for (int s = 0; s < p.n_stages && rc == 0; ++s) {
    rc = bemfmm_set_geometry(c, p.n_nodes, p.X[s].data(), p.n_faces, p.tri.data(), p.face_dof.data(), p.n_dof,
                             p.kind.data(), p.point.data());
    if (!rc) rc = bemfmm_set_bc(c, p.phi[s].data(), p.phin[s].data());
    if (!rc) rc = bemfmm_prepare(c);
    bemfmm_result r{};
    if (!rc) rc = bemfmm_solve(c, x0.data(), out[s].data(), 1e-10, 600, 600, &r);
    if (std::getenv("DRIVER_TRACE_BYTES")) std::printf("stage %d live device bytes %lld\n", s, hipstub_live_device_bytes());
    if (rc) std::fprintf(stderr, "driver: stage %d rc %d: %s\n", s, rc, bemfmm_last_error(c));
  }
  bemfmm_ctx_destroy(c);
  return rc;
}

double rel_diff(const std::vector<double>& a, const std::vector<double>& b) {
  double d = 0., n = 0.;
  for (size_t i = 0; i < a.size(); ++i) {
    d += (a[i] - b[i]) * (a[i] - b[i]);
    n += b[i] * b[i];
  }
  return std::sqrt(d / (n > 0 ? n : 1.));
}

}  // namespace

// Check of the checker: one thread captures a stream while another frees device memory (an implicit
// device synchronisation); the stub must record it. Returns 0 when it does.
int selftest() {
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return 2;
  void* p = nullptr;
  if (hipMalloc(&p, 64) != hipSuccess) return 2;
  std::mutex mu;
  std::condition_variable cv;
  int phase = 0;
  std::thread cap([&] {
    (void)hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    {
      std::unique_lock<std::mutex> lk(mu);
      phase = 1;
      cv.notify_all();
      cv.wait(lk, [&] { return phase == 2; });
    }
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(s, &g);
    (void)hipGraphDestroy(g);
  });
  {
    std::unique_lock<std::mutex> lk(mu);
    cv.wait(lk, [&] { return phase == 1; });
  }
  (void)hipFree(p);  // while the other thread captures
  {
    std::lock_guard<std::mutex> lk(mu);
    phase = 2;
    cv.notify_all();
  }
  cap.join();
  (void)hipStreamDestroy(s);
  const int nv = hipstub_violation_count();
  std::printf("selftest: violations %d\n", nv);
  return nv >= 1 ? 0 : 1;
}

int main(int argc, char** argv) {
  if (argc == 2 && std::string(argv[1]) == "selftest") return selftest();
  if (argc < 5) {
    std::fprintf(stderr, "usage: driver <problems> <threads|server> <fp64|fp32> <slots> [p]\n");
    return 2;
  }
  const std::vector<Problem> ps = load(argv[1]);
  const std::string mode = argv[2];
  const bool fp32 = std::string(argv[3]) == "fp32";
  const int slots = std::atoi(argv[4]);
  const int P = argc > 5 ? std::atoi(argv[5]) : 8;
  const bemfmm_opts o = options(fp32, P);
  const int np = (int)ps.size();
  // alone
  std::vector<std::vector<std::vector<double>>> alone(np), mix(np);
  for (int i = 0; i < np; ++i) {
    const int rc = run_session(ps[i], o, alone[i]);
    std::printf("alone %d: rc %d\n", i, rc);
    if (rc) return 1;
  }
  std::fflush(stdout);
  int failures = 0;
  if (mode == "threads") {
    std::vector<std::thread> th;
    std::vector<int> rcs(np, 0);
    for (int i = 0; i < np; ++i) th.emplace_back([&, i] { rcs[i] = run_session(ps[i], o, mix[i]); });
    for (auto& t : th) t.join();
    for (int i = 0; i < np; ++i)
      if (rcs[i]) {
        std::printf("mix %d: rc %d\n", i, rcs[i]);
        ++failures;
      }
  } else {
    bemfmm_server_opts so;
    bemfmm_server_default_opts(&so);
    so.slots = slots;
    bemfmm_server* srv = bemfmm_server_create(&so);
    if (!srv) return 2;
    std::vector<std::thread> th;
    std::vector<int> rcs(np, 0);
    for (int i = 0; i < np; ++i)
      th.emplace_back([&, i] {
        const Problem& p = ps[i];
        mix[i].assign(p.n_stages, std::vector<double>(p.n_dof, 0.));
        for (int s = 0; s < p.n_stages; ++s) {
          bemfmm_problem q{};
          q.n_nodes = p.n_nodes;
          q.X = p.X[s].data();
          q.n_faces = p.n_faces;
          q.tri = p.tri.data();
          q.face_dof = p.face_dof.data();
          q.n_dof = p.n_dof;
          q.dof_kind = p.kind.data();
          q.dof_point = p.point.data();
          q.phi = p.phi[s].data();
          q.phin = p.phin[s].data();
          q.x0 = nullptr;
          q.opts = o;
          q.tol = 1e-10;
          q.max_iter = 600;
          q.restart = 600;
          const int64_t t = bemfmm_server_submit(srv, &q, i);
          if (t <= 0) {
            rcs[i] = (int)t;
            return;
          }
          bemfmm_result r{};
          const int rc = bemfmm_server_wait(srv, t, mix[i][s].data(), &r, nullptr);
          if (rc) {
            rcs[i] = rc;
            return;
          }
        }
        bemfmm_server_close_session(srv, i);
      });
    for (auto& t : th) t.join();
    bemfmm_server_destroy(srv);
    for (int i = 0; i < np; ++i)
      if (rcs[i]) {
        std::printf("mix %d: rc %d\n", i, rcs[i]);
        ++failures;
      }
  }
  double worst = 0.;
  for (int i = 0; i < np && !failures; ++i)
    for (int s = 0; s < ps[i].n_stages; ++s) worst = std::fmax(worst, rel_diff(mix[i][s], alone[i][s]));
  const int nv = hipstub_violation_count();
  std::printf("live device bytes after teardown: %lld\n", hipstub_live_device_bytes());
  std::printf("mix vs alone: max relative difference %.3e; violations %d; captures active at once (max) %d\n", worst,
              nv, hipstub_max_concurrent_captures());
  for (int i = 0; i < nv && i < 20; ++i) std::printf("violation: %s\n", hipstub_violation(i));
  return (failures || nv || !(worst <= 1e-10)) ? 1 : 0;
}
