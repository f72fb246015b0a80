// relax_driver.cpp -- test driver for particle management through the class
// API: a wake copied into MAXNUMPARTICLES rows (wake.cpp:558-562, as the
// reference's validation drivers build theirs), one __interaction over it,
// a right-hand side, relax(stepnum) (interaction.cpp:367-373 ->
// __wake::relax, wake.cpp:161-165), and a right-hand side of the relaxed
// wake.
//
//   relax_driver <input> <output>
//
// input: N (int64), stepnum (double), then N x 3 positions, N x 3
// strengths, N radii, N volumes, N birth strengths, N birth steps (doubles).
// output: N after relax (int64), then for the N particles: positions,
// strengths, volumes, birth strengths, velocities, vorticity rates.
#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <vector>

#include "pawan/interaction.h"
#include "pawan/wake.h"

namespace {

void read_exact(FILE *f, void *p, size_t bytes) {
  if (bytes && std::fread(p, 1, bytes, f) != bytes) throw std::runtime_error("short read");
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: relax_driver <input> <output>\n");
    return 2;
  }
  try {
    FILE *in = std::fopen(argv[1], "rb");
    if (!in) throw std::runtime_error("cannot open input");
    int64_t n = 0;
    double stepnum = 0.0;
    read_exact(in, &n, 8);
    read_exact(in, &stepnum, 8);
    std::vector<double> pos(3 * n), alpha(3 * n), rad(n), vol(n), birth(n), act(n);
    read_exact(in, pos.data(), pos.size() * 8);
    read_exact(in, alpha.data(), alpha.size() * 8);
    read_exact(in, rad.data(), rad.size() * 8);
    read_exact(in, vol.data(), vol.size() * 8);
    read_exact(in, birth.data(), birth.size() * 8);
    read_exact(in, act.data(), act.size() * 8);
    std::fclose(in);
    pawan::__wake w0(static_cast<size_t>(n), pos.data(), alpha.data(), rad.data(), vol.data());
    for (int64_t i = 0; i < n; ++i) w0._birthstrength->set(i, birth[i]);
    pawan::__wake W(&w0);  // MAXNUMPARTICLES rows; _active is not copied (wake.cpp:542-556)
    for (int64_t i = 0; i < n; ++i) W._active->set(i, act[i]);
    pawan::__interaction IN(&W);
    IN.setNu(1e-4);
    IN.solve();  // the device copy exists before relax
    size_t step = static_cast<size_t>(stepnum);
    IN.relax(step);
    IN.solve();
    FILE *out = std::fopen(argv[2], "wb");
    if (!out) throw std::runtime_error("cannot open output");
    const int64_t m = static_cast<int64_t>(W._numParticles);
    std::fwrite(&m, 8, 1, out);
    std::fwrite(W._position->ptr(), 8, 3 * m, out);
    std::fwrite(W._vorticity->ptr(), 8, 3 * m, out);
    std::fwrite(W._volume->ptr(), 8, m, out);
    std::fwrite(W._birthstrength->ptr(), 8, m, out);
    std::fwrite(W._velocity->ptr(), 8, 3 * m, out);
    std::fwrite(W._retvorcity->ptr(), 8, 3 * m, out);
    std::fclose(out);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "relax_driver: %s\n", e.what());
    return 1;
  }
  return 0;
}
