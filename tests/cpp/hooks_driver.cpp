// hooks_driver.cpp -- test driver: a subclass of pawan::__parallel that
// replaces some of the reference's overridable hooks, the way the
// reference's own __parallel replaces __interaction's
// (src/interaction/parallel.h:48-70):
//   * interact(__wake *W), the pairs within one wake, runs on the host with
//     the oracle's restatement of INTERACT (interaction_utils.h:456-501) in
//     the reference's loop order (interaction.cpp:262-281);
//   * calculateKineticEnergy(__wake *W) returns twice the base value.
// The cross-wake blocks stay with the base class (the device).
//
//   hooks_driver <input> <name> <dt> <nu>
//
// input: the wake file of pawan_run.  Writes <name>.rates (per wake N x 3
// velocity then N x 3 vorticity rate, doubles), <name>.diag (15 doubles,
// diagnose() in writediagnosis order) and, through __rk4::integrate with one
// step, $PAWAN_DATA_DIR/<name>Wake.wake; prints the hook call counts.
#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "pawan/integration.h"
#include "pawan/interaction.h"
#include "pawan/wake.h"

extern "C" void or_interact(double nu, double s_src, double s_trg, const double *r_src, const double *r_trg,
                            const double *a_src, const double *a_trg, double v_src, double v_trg, double *dr_src,
                            double *dr_trg, double *da_src, double *da_trg);

namespace {

class HostSelfPairs : public pawan::__parallel {
 public:
  int self_calls = 0, ke_calls = 0;
  HostSelfPairs(pawan::__wake *W1, pawan::__wake *W2) : pawan::__parallel(W1, W2) {}

 protected:
  double calculateKineticEnergy(pawan::__wake *W) override {
    ++ke_calls;
    return 2.0 * pawan::__interaction::calculateKineticEnergy(W);
  }

 private:
  void interact(pawan::__wake *W) override {
    ++self_calls;
    for (size_t s = 0; s < W->_numParticles; ++s)
      for (size_t t = s + 1; t < W->_numParticles; ++t)
        or_interact(_nu, W->_radius->get(s), W->_radius->get(t), W->_position->row(s), W->_position->row(t),
                    W->_vorticity->row(s), W->_vorticity->row(t), W->_volume->get(s), W->_volume->get(t),
                    W->_velocity->row(s), W->_velocity->row(t), W->_retvorcity->row(s), W->_retvorcity->row(t));
  }
};

void read_exact(FILE *f, void *p, size_t bytes) {
  if (bytes && std::fread(p, 1, bytes, f) != bytes) throw std::runtime_error("short read");
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: hooks_driver <input> <name> <dt> <nu>\n");
    return 2;
  }
  try {
    FILE *in = std::fopen(argv[1], "rb");
    if (!in) throw std::runtime_error("cannot open input");
    int64_t nw = 0;
    read_exact(in, &nw, sizeof nw);
    if (nw != 2) throw std::runtime_error("two wakes expected");
    std::vector<std::unique_ptr<pawan::__wake>> wakes;
    for (int64_t w = 0; w < nw; ++w) {
      int64_t np = 0;
      read_exact(in, &np, sizeof np);
      std::vector<double> pos(3 * np), alpha(3 * np), rad(np), vol(np);
      read_exact(in, pos.data(), pos.size() * 8);
      read_exact(in, alpha.data(), alpha.size() * 8);
      read_exact(in, rad.data(), rad.size() * 8);
      read_exact(in, vol.data(), vol.size() * 8);
      wakes.emplace_back(new pawan::__wake(static_cast<size_t>(np), pos.data(), alpha.data(), rad.data(), vol.data()));
    }
    std::fclose(in);
    const std::string name = argv[2];
    const double dt = std::atof(argv[3]);
    HostSelfPairs S(wakes[0].get(), wakes[1].get());
    S.setNu(std::atof(argv[4]));

    S.solve();
    FILE *f = std::fopen((name + ".rates").c_str(), "wb");
    for (auto &w : wakes) {
      std::fwrite(w->_velocity->ptr(), sizeof(double), 3 * w->_numParticles, f);
      std::fwrite(w->_retvorcity->ptr(), sizeof(double), 3 * w->_numParticles, f);
    }
    std::fclose(f);

    S.diagnose();
    f = std::fopen((name + ".diag").c_str(), "wb");
    S.writediagnosis(f);
    std::fclose(f);

    pawan::__rk4 rk(dt, 1);
    pawan::__io io(name);
    rk.integrate(&S, &io, false);
    std::printf("self_calls %d ke_calls %d\n", S.self_calls, S.ke_calls);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "hooks_driver: %s\n", e.what());
    return 1;
  }
  return 0;
}
