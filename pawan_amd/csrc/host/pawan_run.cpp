// pawan_run.cpp -- a driver written against the class API exactly as the
// reference's validation drivers are (src/pawan.cpp:140-177): wakes ->
// __parallel -> __rk4(T, n) -> integrate(S, IO, diagnose).  The wakes are
// read from a binary file, or built as a ring by __vring/__resolve.
//
//   pawan_run <input> <name> <T> <n> [euler|rk3|rk4] [fp64|fp64_ref|fp32c|learned]
//             [nu] [diagnose 0|1]
//
// learned: the tree path.  Its network is read from $PAWAN_FF_WEIGHTS (36,748
// float32 values) when set, else trained on samples of the initial state
// ($PAWAN_FF_SAMPLES, default 200,000; $PAWAN_FF_STEPS Adam steps, default
// 5,000).  $PAWAN_FF_THETA (0.5) and $PAWAN_FF_TOL (0: opening angle only)
// set the traversal.
//
// input: int64 nwakes, then per wake int64 n followed by n x 3 positions,
// n x 3 strengths, n radii, n volumes (doubles, native endianness); or
// "vring:R,core,nLayers,nAzi,sig": a __vring resolved by __resolve::rebuild
// and copied into a plain wake, as src/pawan.cpp:57-84 builds its rings.
// Writes $PAWAN_DATA_DIR/<name>Wake.wake and .diagnosis in the reference's
// formats (readable by its scripts/readWake.py and readDiagnostics.py).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "pawan_gpu.h"
#include "pawan/integration.h"
#include "pawan/interaction.h"
#include "pawan/resolve.h"
#include "pawan/wake.h"

static void read_exact(FILE* f, void* p, size_t bytes) {
  if (bytes && std::fread(p, 1, bytes, f) != bytes) throw std::runtime_error("short read");
}

static double env_or(const char* name, double dflt) {
  const char* v = std::getenv(name);
  return v ? std::atof(v) : dflt;
}

static void check(int rc, const char* what) {
  if (rc != PW_OK) throw std::runtime_error(std::string(what) + ": " + pw_last_error());
}

// network and traversal of the tree path on the interaction's context
static void setup_learned(pawan::__interaction* S) {
  pw_ctx* ctx = S->context();
  check(pw_ff_set_params(ctx, env_or("PAWAN_FF_THETA", 0.5), env_or("PAWAN_FF_TOL", 0.0), 1), "pw_ff_set_params");
  if (const char* path = std::getenv("PAWAN_FF_WEIGHTS")) {
    std::vector<float> w(36748);
    FILE* f = std::fopen(path, "rb");
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    read_exact(f, w.data(), w.size() * sizeof(float));
    std::fclose(f);
    check(pw_ff_set_weights(ctx, w.data(), static_cast<int64_t>(w.size())), "pw_ff_set_weights");
    return;
  }
  check(pw_ff_sample(ctx, static_cast<int64_t>(env_or("PAWAN_FF_SAMPLES", 200000)), 1), "pw_ff_sample");
  double rep[10] = {0};
  check(pw_ff_train(ctx, static_cast<int64_t>(env_or("PAWAN_FF_STEPS", 5000)), 1024, 2e-3, 1, rep, 10),
        "pw_ff_train");
  std::fprintf(stderr, "pawan_run: network trained: held-out loss %.4f (expansion alone %.4f), %.0f ms\n", rep[1],
               rep[2], rep[9]);
}

int main(int argc, char** argv) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s input name T n [scheme] [precision] [nu] [diagnose]\n", argv[0]);
    return 2;
  }
  try {
    const std::string scheme = argc > 5 ? argv[5] : "rk4";
    const std::string prec = argc > 6 ? argv[6] : "fp64";
    const double T = std::atof(argv[3]);
    const size_t n = static_cast<size_t>(std::atoll(argv[4]));
    std::vector<std::unique_ptr<pawan::__wake>> wakes;
    if (std::strncmp(argv[1], "vring:", 6) == 0) {
      double R = 0, core = 0, sig = 0;
      int nl = 0, nazi = 0;
      if (std::sscanf(argv[1] + 6, "%lf,%lf,%d,%d,%lf", &R, &core, &nl, &nazi, &sig) != 5)
        throw std::runtime_error("vring:R,core,nLayers,nAzi,sig expected");
      pawan::__vring ring(R, core, nl, nazi, sig);
      pawan::__interaction resolver(&ring);
      pawan::__io io(std::string(argv[2]) + "_vring");
      pawan::__resolve().rebuild(&resolver, &io);
      wakes.emplace_back(new pawan::__wake(&ring));
    }
    FILE* in = wakes.empty() ? std::fopen(argv[1], "rb") : nullptr;
    if (wakes.empty() && !in) throw std::runtime_error(std::string("cannot open ") + argv[1]);
    int64_t nw = 0;
    if (in) read_exact(in, &nw, sizeof nw);
    for (int64_t w = 0; w < nw; ++w) {
      int64_t np = 0;
      read_exact(in, &np, sizeof np);
      std::vector<double> pos(3 * np), alpha(3 * np), rad(np), vol(np);
      read_exact(in, pos.data(), pos.size() * 8);
      read_exact(in, alpha.data(), alpha.size() * 8);
      read_exact(in, rad.data(), rad.size() * 8);
      read_exact(in, vol.data(), vol.size() * 8);
      wakes.emplace_back(new pawan::__wake(static_cast<size_t>(np), pos.data(), alpha.data(),
                                           rad.data(), vol.data()));
    }
    if (in) std::fclose(in);

    pawan::__interaction* S = new pawan::__parallel(wakes[0].get());
    for (size_t w = 1; w < wakes.size(); ++w) S->addWake(wakes[w].get());
    if (argc > 7) S->setNu(std::atof(argv[7]));
    S->setPrecision(prec == "fp32c"      ? PW_FP32C
                    : prec == "fp64_ref" ? PW_FP64_REF
                    : prec == "learned"  ? PW_LEARNED
                                         : PW_FP64);
    if (prec == "learned") setup_learned(S);
    const bool diagnose = argc > 8 ? std::atoi(argv[8]) != 0 : true;

    pawan::__io* IO = new pawan::__io(argv[2]);
    pawan::__integration* IN = scheme == "rk3"     ? new pawan::__rk3(T, n)
                               : scheme == "euler" ? new pawan::__integration(T, n)
                                                   : new pawan::__rk4(T, n);
    IN->integrate(S, IO, diagnose);
    delete IN;
    delete IO;
    delete S;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "pawan_run: %s\n", e.what());
    return 1;
  }
  return 0;
}
