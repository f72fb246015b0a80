// pawan_run.cpp -- a driver written against the class API exactly as the
// reference's validation drivers are (src/pawan.cpp:140-177): wakes ->
// __parallel -> __rk4(T, n) -> integrate(S, IO, diagnose).  The wakes are
// read from a binary file instead of being built by __vring/__resolve.
//
//   pawan_run <input> <name> <T> <n> [euler|rk3|rk4] [fp64|fp64_ref|fp32c]
//             [nu] [diagnose 0|1]
//
// input: int64 nwakes, then per wake int64 n followed by n x 3 positions,
// n x 3 strengths, n radii, n volumes (doubles, native endianness).
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
#include "pawan/wake.h"

static void read_exact(FILE* f, void* p, size_t bytes) {
  if (bytes && std::fread(p, 1, bytes, f) != bytes) throw std::runtime_error("short read");
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
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) throw std::runtime_error(std::string("cannot open ") + argv[1]);
    int64_t nw = 0;
    read_exact(in, &nw, sizeof nw);
    std::vector<std::unique_ptr<pawan::__wake>> wakes;
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
    std::fclose(in);

    pawan::__interaction* S = new pawan::__parallel(wakes[0].get());
    for (size_t w = 1; w < wakes.size(); ++w) S->addWake(wakes[w].get());
    if (argc > 7) S->setNu(std::atof(argv[7]));
    S->setPrecision(prec == "fp32c" ? PW_FP32C : prec == "fp64_ref" ? PW_FP64_REF : PW_FP64);
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
