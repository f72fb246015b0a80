// pawan_host.cpp -- the C++ class API (pawan::__wake, __interaction,
// __integration, __rk4, __rk3, __io) over the C-ABI of include/pawan_gpu.h.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "pawan_gpu.h"
#include "pawan/integration.h"
#include "pawan/interaction.h"
#include "pawan/resolve.h"
#include "pawan/wake.h"

namespace pawan {

// ---------------------------------------------------------------------------
// __wake (src/wake/wake.cpp)
// ---------------------------------------------------------------------------

void __wake::create_particles(const int &n) {
  _numParticles = n;
  _npidx = _numParticles;
  _maxmem = false;
  _size = 2 * _numParticles * 3;
  _maxnumParticles = _numParticles;
  _maxsize = 2 * _maxnumParticles * 3;
  _position = new mat(_numParticles, 3);
  _velocity = new mat(_numParticles, 3);
  _vorticity = new mat(_numParticles, 3);
  _retvorcity = new mat(_numParticles, 3);
  _radius = new vec(_numParticles);
  _volume = new vec(_numParticles);
  _birthstrength = new vec(_numParticles);
  _vorticityfield = new mat(_numParticles, 3);
  _active = new vec(_numParticles);
}

void __wake::initialise_memory() {
  _numParticles = 0;
  _npidx = 0;
  _maxmem = false;
  _size = 0;
  _maxnumParticles = MAXNUMPARTICLES;
  _maxsize = 2 * _maxnumParticles * 3;
  _position = new mat(_maxnumParticles, 3);
  _velocity = new mat(_maxnumParticles, 3);
  _vorticity = new mat(_maxnumParticles, 3);
  _retvorcity = new mat(_maxnumParticles, 3);
  _radius = new vec(_maxnumParticles);
  _volume = new vec(_maxnumParticles);
  _birthstrength = new vec(_maxnumParticles);
  _vorticityfield = new mat(_maxnumParticles, 3);
  _active = new vec(_maxnumParticles);
}

__wake::__wake() {
  create_particles(1);
  _radius->set(0, 1.0);
  _volume->set(0, 1.0);
  _birthstrength->set(0, 1.0);
}

__wake::__wake(const int &n) { create_particles(n); }

__wake::__wake(size_t n, const double *pos, const double *alpha, const double *radius,
               const double *volume) {
  create_particles(static_cast<int>(n));
  for (size_t i = 0; i < n; ++i) {
    for (size_t j = 0; j < 3; ++j) {
      _position->set(i, j, pos[3 * i + j]);
      _vorticity->set(i, j, alpha[3 * i + j]);
    }
    _radius->set(i, radius[i]);
    _volume->set(i, volume[i]);
    _birthstrength->set(i, std::sqrt(alpha[3 * i] * alpha[3 * i] + alpha[3 * i + 1] * alpha[3 * i + 1] +
                                     alpha[3 * i + 2] * alpha[3 * i + 2]));
    _active->set(i, 1.0);
  }
}

void __wake::addParticles(__wake *W) {
  for (size_t i = 0; i < W->_numParticles; ++i) {
    for (size_t j = 0; j < 3; ++j) {
      _position->set(i, j, W->_position->get(i, j));
      _vorticity->set(i, j, W->_vorticity->get(i, j));
      _vorticityfield->set(i, j, W->_vorticityfield->get(i, j));
    }
    _radius->set(i, W->_radius->get(i));
    _volume->set(i, W->_volume->get(i));
    _birthstrength->set(i, W->_birthstrength->get(i));
  }
  _numParticles = W->_numParticles;
  _npidx = _numParticles;
  _size = 2 * _numParticles * 3;
}

// wake.cpp:61-110, on the device
void __wake::split(size_t &stepnum) {
  int64_t n = static_cast<int64_t>(_numParticles), k = static_cast<int64_t>(_npidx), ns = 0;
  if (pw_wake_split(0, static_cast<double>(stepnum), static_cast<int64_t>(_maxnumParticles), &n, &k,
                    _position->ptr(), _vorticity->ptr(), _radius->ptr(), _volume->ptr(), _birthstrength->ptr(),
                    _active->ptr(), &ns) != PW_OK)
    throw std::runtime_error(std::string("__wake::split: ") + pw_last_error());
  _numParticles = static_cast<size_t>(n);
  _npidx = static_cast<size_t>(k);
  _size = 2 * _numParticles * 3;
}

// wake.cpp:112-159, on the device
void __wake::merge(size_t &stepnum) {
  int64_t nm = 0;
  if (pw_wake_merge(0, static_cast<double>(stepnum), static_cast<int64_t>(_numParticles), _position->ptr(),
                    _vorticity->ptr(), _radius->ptr(), _volume->ptr(), _active->ptr(), &nm) != PW_OK)
    throw std::runtime_error(std::string("__wake::merge: ") + pw_last_error());
}

// wake.cpp:161-165
void __wake::relax(size_t &stepnum) {
  split(stepnum);
  merge(stepnum);
}

__wake::__wake(__wake *W) {
  if (W->_numParticles > MAXNUMPARTICLES)
    throw std::length_error("__wake copy: more than MAXNUMPARTICLES particles");
  initialise_memory();
  addParticles(W);
}

__wake::~__wake() {
  delete _position;
  delete _velocity;
  delete _vorticity;
  delete _retvorcity;
  delete _radius;
  delete _volume;
  delete _birthstrength;
  delete _vorticityfield;
  delete _active;
}

void __wake::print() {
  std::printf("_numParticles %zu\n", _numParticles);
  for (size_t i = 0; i < _numParticles; ++i)
    std::printf("%zu: x=(%g %g %g) a=(%g %g %g) s=%g V=%g\n", i, _position->get(i, 0),
                _position->get(i, 1), _position->get(i, 2), _vorticity->get(i, 0),
                _vorticity->get(i, 1), _vorticity->get(i, 2), _radius->get(i), _volume->get(i));
}

static void fwrite_all(const double *p, size_t n, FILE *f) {
  if (n && std::fwrite(p, sizeof(double), n, f) != n) throw std::runtime_error("fwrite failed");
}

void __wake::save(FILE *f) {
  std::fwrite(&_numParticles, sizeof(size_t), 1, f);
  fwrite_all(_position->ptr(), _position->data.size(), f);
  fwrite_all(_vorticity->ptr(), _vorticity->data.size(), f);
  fwrite_all(_radius->ptr(), _radius->size, f);
  fwrite_all(_volume->ptr(), _volume->size, f);
  fwrite_all(_birthstrength->ptr(), _birthstrength->size, f);
}

void __wake::write(FILE *f) {
  std::fwrite(&_numParticles, sizeof(size_t), 1, f);
  fwrite_all(_position->ptr(), _position->data.size(), f);
  fwrite_all(_vorticity->ptr(), _vorticity->data.size(), f);
  fwrite_all(_radius->ptr(), _radius->size, f);
  fwrite_all(_active->ptr(), _active->size, f);
  fwrite_all(_volume->ptr(), _volume->size, f);
  fwrite_all(_birthstrength->ptr(), _birthstrength->size, f);
}

void __wake::setStates(const vec *state) {
  const size_t off = _maxnumParticles * 3;
  for (size_t i = 0; i < _maxnumParticles; ++i)
    for (size_t j = 0; j < 3; ++j) {
      _position->set(i, j, state->get(3 * i + j));
      _vorticity->set(i, j, state->get(off + 3 * i + j));
    }
}

void __wake::getRates(vec *rate) {
  for (size_t i = 0; i < _numParticles; ++i)
    for (size_t j = 0; j < 3; ++j) {
      size_t ind = i * 3 + j;
      rate->set(ind, _velocity->get(i, j));
      ind += _maxsize / 2;
      rate->set(ind, _retvorcity->get(i, j));
    }
}

void __wake::getStates(vec *state) {
  for (size_t i = 0; i < _numParticles; ++i)
    for (size_t j = 0; j < 3; ++j) {
      size_t ind = i * 3 + j;
      state->set(ind, _position->get(i, j));
      ind += _maxsize / 2;
      state->set(ind, _vorticity->get(i, j));
    }
}

void __wake::getIdealRates(vec *rate) { rate->set_zero(); }

void __wake::translate(const size_t &n, const double &x) {
  for (size_t i = 0; i < _numParticles; ++i) _position->set(i, n, x + _position->get(i, n));
}

void __wake::translate(const double *x) {
  for (size_t i = 0; i < _numParticles; ++i)
    for (size_t j = 0; j < 3; ++j) _position->set(i, j, x[j] + _position->get(i, j));
}

// The second updated coordinate uses the first one already rotated, exactly
// as wake.cpp:916-937 does, so that drivers get the same wakes.
void __wake::rotate(const size_t &n, const double &a) {
  const double c = std::cos(a), s = std::sin(a);
  for (size_t i = 0; i < _numParticles; ++i) {
    if (n == 0) {
      _position->set(i, 1, c * _position->get(i, 1) - s * _position->get(i, 2));
      _position->set(i, 2, s * _position->get(i, 1) + c * _position->get(i, 2));
    } else if (n == 1) {
      _position->set(i, 0, c * _position->get(i, 0) + s * _position->get(i, 2));
      _position->set(i, 2, -s * _position->get(i, 0) + c * _position->get(i, 2));
    } else if (n == 2) {
      _position->set(i, 1, c * _position->get(i, 0) - s * _position->get(i, 1));
      _position->set(i, 2, s * _position->get(i, 0) + c * _position->get(i, 1));
    }
  }
}

// ---------------------------------------------------------------------------
// __interaction (src/interaction/interaction.cpp)
// ---------------------------------------------------------------------------

__interaction::__interaction() {
  _nu = 1.56e-5;  // interaction.cpp:15
  _nWake = 0;
  _totalVorticity = new vec(3);
  _linearImpulse = new vec(3);
  _angularImpulse = new vec(3);
  _enstrophy = _helicity = _kineticEnergy = _enstrophyF = _kineticEnergyF = 0.0;
  _precision = PW_FP64;
}

__interaction::__interaction(__wake *W) : __interaction() { addWake(W); }

__interaction::__interaction(__wake *W1, __wake *W2) : __interaction() {
  addWake(W1);
  addWake(W2);
}

__interaction::~__interaction() {
  if (_ctx) pw_destroy(_ctx);
  delete _totalVorticity;
  delete _linearImpulse;
  delete _angularImpulse;
}

void __interaction::fail(const char *what) const {
  throw std::runtime_error(std::string(what) + ": " + pw_last_error());
}

void __interaction::addWake(__wake *W) {
  _W.push_back(W);
  _size = W->_size;  // interaction.cpp:47
  _totalmaxsize += W->_maxsize;
  _nWake++;
  if (_ctx) {  // the device copy is rebuilt with the new wake on next use
    pw_destroy(_ctx);
    _ctx = nullptr;
  }
}

// interaction.cpp:367-373.  The wakes change on the host; a wake that grew
// needs a new device copy (as after addWake), otherwise positions,
// strengths and volumes are sent to the existing one.
void __interaction::relax(size_t &stepnum) {
  std::vector<size_t> n0;
  for (__wake *w : _W) n0.push_back(w->_numParticles);
  for (__wake *w : _W) w->relax(stepnum);
  bool grew = false;
  for (size_t k = 0; k < _W.size(); ++k) grew = grew || _W[k]->_numParticles != n0[k];
  if (!_ctx) return;
  if (grew) {
    pw_destroy(_ctx);
    _ctx = nullptr;
    _diagValid = false;
    return;
  }
  for (size_t k = 0; k < _W.size(); ++k)
    if (pw_set_wake_state(_ctx, static_cast<int>(k), _W[k]->_position->ptr(), _W[k]->_vorticity->ptr()) != PW_OK ||
        pw_set_wake_volume(_ctx, static_cast<int>(k), _W[k]->_volume->ptr()) != PW_OK)
      fail("__interaction::relax");
  _deviceStale = false;
  _diagValid = false;
}

void __interaction::ensureDevice() {
  if (_ctx) return;
  if (pw_create(&_ctx, 0) != PW_OK) fail("pw_create");
#if defined(HIGHORDER) && HIGHORDER
  // the reference's compile-time kernel choice (interaction_utils.h:22-26)
  if (pw_set_kernel(_ctx, PW_KERNEL_HIGHORDER) != PW_OK) fail("pw_set_kernel");
#endif
  for (__wake *w : _W) {
    if (pw_add_wake(_ctx, static_cast<int64_t>(w->_numParticles),
                    static_cast<int64_t>(w->_maxnumParticles), w->_position->ptr(),
                    w->_vorticity->ptr(), w->_radius->ptr(), w->_volume->ptr(), nullptr) != PW_OK)
      fail("pw_add_wake");
  }
  _deviceStale = false;
}

pw_ctx *__interaction::context() {
  pushState();
  return _ctx;
}

void __interaction::pushState() {
  ensureDevice();
  if (pw_set_nu(_ctx, _nu) != PW_OK) fail("pw_set_nu");
  if (!_deviceStale) return;
  for (size_t k = 0; k < _W.size(); ++k)
    if (pw_set_wake_state(_ctx, static_cast<int>(k), _W[k]->_position->ptr(),
                          _W[k]->_vorticity->ptr()) != PW_OK)
      fail("pw_set_wake_state");
  _deviceStale = false;
}

void __interaction::pullState() {
  for (size_t k = 0; k < _W.size(); ++k)
    if (pw_get_wake_state(_ctx, static_cast<int>(k), _W[k]->_position->ptr(),
                          _W[k]->_vorticity->ptr()) != PW_OK)
      fail("pw_get_wake_state");
}

void __interaction::pullRates() {
  for (size_t k = 0; k < _W.size(); ++k)
    if (pw_get_wake_rates(_ctx, static_cast<int>(k), _W[k]->_velocity->ptr(),
                          _W[k]->_retvorcity->ptr()) != PW_OK)
      fail("pw_get_wake_rates");
}

void __interaction::addRates() {
  for (size_t k = 0; k < _W.size(); ++k) {
    __wake *w = _W[k];
    const size_t n3 = 3 * w->_numParticles;
    std::vector<double> u(n3), r(n3);
    if (pw_get_wake_rates(_ctx, static_cast<int>(k), u.data(), r.data()) != PW_OK) fail("pw_get_wake_rates");
    double *hu = w->_velocity->ptr(), *hr = w->_retvorcity->ptr();
    for (size_t i = 0; i < n3; ++i) {
      hu[i] += u[i];
      hr[i] += r[i];
    }
  }
}

int __interaction::wakeIndex(const __wake *W) const {
  for (size_t k = 0; k < _W.size(); ++k)
    if (_W[k] == W) return static_cast<int>(k);
  throw std::invalid_argument("wake not held by this interaction");
}

// interaction.cpp:196-210.  The base hooks queue their blocks; the queue
// then runs on the device and its rates are added to the wakes' rows, which
// an overriding hook may already have filled.
void __interaction::interact() {
  for (__wake *w : _W) {
    w->_velocity->set_zero();
    w->_retvorcity->set_zero();
  }
  _pending.clear();
  for (__wake *w : _W) interact(w);
  for (size_t i = 0; i < _nWake; ++i)
    for (size_t j = i + 1; j < _nWake; ++j) interact(_W[i], _W[j]);
  const size_t blocks = _nWake * (_nWake + 1) / 2;
  _interactHooks = _pending.size() / 2 == blocks ? 0 : 1;
  if (_pending.empty()) return;
  pushState();
  const int rc = _interactHooks == 0 ? pw_solve(_ctx, _precision)
                                     : pw_interact_wakes(_ctx, _precision, _pending.data(),
                                                         static_cast<int>(_pending.size() / 2));
  if (rc != PW_OK) fail("pw_interact_wakes");
  addRates();
}

void __interaction::queueBlock(__wake *W1, __wake *W2) {
  const int a = wakeIndex(W1), b = wakeIndex(W2);
  _pending.push_back(std::min(a, b));
  _pending.push_back(std::max(a, b));
}

void __interaction::interact(__wake *W) { queueBlock(W, W); }
void __interaction::interact(__wake *W1, __wake *W2) { queueBlock(W1, W2); }

// interaction.cpp:212-226, the field blocks likewise
void __interaction::influence() {
  for (__wake *w : _W) w->_vorticityfield->set_zero();
  _pending.clear();
  for (__wake *w : _W) influence(w);
  for (size_t i = 0; i < _nWake; ++i)
    for (size_t j = i + 1; j < _nWake; ++j) influence(_W[i], _W[j]);
  _influenceHooks = _pending.size() / 2 == _nWake * (_nWake + 1) / 2 ? 0 : 1;
  if (_pending.empty()) return;
  pushState();
  size_t n = 0;
  for (__wake *w : _W) n += w->_numParticles;
  std::vector<double> field(3 * n);
  const size_t blocks = _nWake * (_nWake + 1) / 2;
  const bool all = _pending.size() / 2 == blocks;
  if (pw_vorticity_field(_ctx, all ? nullptr : _pending.data(), all ? 0 : static_cast<int>(_pending.size() / 2),
                         field.data(), static_cast<int64_t>(field.size())) != PW_OK)
    fail("pw_vorticity_field");
  size_t off = 0;
  for (__wake *w : _W) {
    double *k = w->_vorticityfield->ptr();
    for (size_t i = 0; i < 3 * w->_numParticles; ++i) k[i] += field[off + i];
    off += 3 * w->_numParticles;
  }
}

void __interaction::influence(__wake *W) { queueBlock(W, W); }
void __interaction::influence(__wake *W1, __wake *W2) { queueBlock(W1, W2); }

void __interaction::resolve() { influence(); }

void __interaction::solve() { interact(); }

// All stages of a step on the device, when every interaction block is the
// base class's; a subclass that replaces a block is stepped through the
// reference's setStates / solve / getRates sequence instead (returns false).
bool __interaction::stepOnDevice(int scheme, const double &dt, vec *state) {
  if (_interactHooks < 0) {
    setStates(state);
    interact();  // one right-hand side, to learn which blocks are overridden
  }
  if (_interactHooks != 0) return false;
  setStates(state);
  pushState();
  if (pw_step(_ctx, scheme, dt, _precision, 1) != PW_OK) fail("pw_step");
  pullState();
  pullRates();
  getStates(state);
  return true;
}

// The diagnostics of the block (W1, W2) (W1 == W2: one wake), computed on
// the device for every block at the first request of a diagnose() call.
const double *__interaction::blockDiag(__wake *W1, __wake *W2) {
  int a = wakeIndex(W1), b = wakeIndex(W2);
  if (a > b) std::swap(a, b);
  const size_t nw = _W.size();
  if (!_diagValid) {
    pushState();
    _diagBlocks.assign(16 * nw * nw, 0.0);
    for (size_t i = 0; i < nw; ++i)
      for (size_t j = i; j < nw; ++j)
        if (pw_diagnose_wakes(_ctx, static_cast<int>(i), static_cast<int>(j), &_diagBlocks[16 * (i * nw + j)],
                              16) != PW_OK)
          fail("pw_diagnose_wakes");
    _diagValid = true;
  }
  return &_diagBlocks[16 * (a * nw + b)];
}

void __interaction::calculateTotalVorticity(__wake *W, vec *O) {
  const double *d = blockDiag(W, W);
  for (int k = 0; k < 3; ++k) O->set(k, d[k]);
}
void __interaction::calculateLinearImpulse(__wake *W, vec *I) {
  const double *d = blockDiag(W, W);
  for (int k = 0; k < 3; ++k) I->set(k, d[3 + k]);
}
void __interaction::calculateAngularImpulse(__wake *W, vec *A) {
  const double *d = blockDiag(W, W);
  for (int k = 0; k < 3; ++k) A->set(k, d[6 + k]);
}
double __interaction::calculateHelicity(__wake *W) { return blockDiag(W, W)[9]; }
double __interaction::calculateHelicity(__wake *W1, __wake *W2) { return blockDiag(W1, W2)[9]; }
double __interaction::calculateEnstrophy(__wake *W) { return blockDiag(W, W)[10]; }
double __interaction::calculateEnstrophy(__wake *W1, __wake *W2) { return blockDiag(W1, W2)[10]; }
double __interaction::calculateEnstrophyF(__wake *W) { return blockDiag(W, W)[11]; }
double __interaction::calculateEnstrophyF(__wake *W1, __wake *W2) { return blockDiag(W1, W2)[11]; }
double __interaction::calculateKineticEnergy(__wake *W) { return blockDiag(W, W)[12]; }
double __interaction::calculateKineticEnergy(__wake *W1, __wake *W2) { return blockDiag(W1, W2)[12]; }
double __interaction::calculateKineticEnergyF(__wake *W) { return blockDiag(W, W)[13]; }
double __interaction::calculateKineticEnergyF(__wake *W1, __wake *W2) { return blockDiag(W1, W2)[13]; }

// interaction.cpp:64-157: the reference's loops over the calculate*
// virtuals; Zc from the last wake with a non-zero denominator
void __interaction::diagnose() {
  _diagValid = false;
  _totalVorticity->set_zero();
  _linearImpulse->set_zero();
  _angularImpulse->set_zero();
  vec O(3), I(3), A(3);
  for (__wake *w : _W) {
    calculateTotalVorticity(w, &O);
    calculateLinearImpulse(w, &I);
    calculateAngularImpulse(w, &A);
    for (int k = 0; k < 3; ++k) {
      _totalVorticity->data[k] += O.data[k];
      _linearImpulse->data[k] += I.data[k];
      _angularImpulse->data[k] += A.data[k];
    }
  }
  _helicity = 0.0;
  for (size_t i = 0; i < _nWake; ++i) {
    _helicity += calculateHelicity(_W[i]);
    for (size_t j = i + 1; j < _nWake; ++j) _helicity += calculateHelicity(_W[i], _W[j]);
  }
  _enstrophy = 0.0;
  for (size_t i = 0; i < _nWake; ++i) {
    _enstrophy += calculateEnstrophy(_W[i]);
    for (size_t j = i + 1; j < _nWake; ++j) _enstrophy += calculateEnstrophy(_W[i], _W[j]);
  }
  _enstrophyF = 0.0;
  for (size_t i = 0; i < _nWake; ++i) {
    _enstrophyF += calculateEnstrophyF(_W[i]);
    for (size_t j = i + 1; j < _nWake; ++j) _enstrophyF += calculateEnstrophyF(_W[i], _W[j]);
  }
  _kineticEnergy = 0.0;
  for (size_t i = 0; i < _nWake; ++i) {
    _kineticEnergy += calculateKineticEnergy(_W[i]);
    for (size_t j = i + 1; j < _nWake; ++j) _kineticEnergy += calculateKineticEnergy(_W[i], _W[j]);
  }
  _kineticEnergyF = 0.0;
  for (size_t i = 0; i < _nWake; ++i) {
    _kineticEnergyF += calculateKineticEnergyF(_W[i]);
    for (size_t j = i + 1; j < _nWake; ++j) _kineticEnergyF += calculateKineticEnergyF(_W[i], _W[j]);
  }
  for (__wake *w : _W) {
    const double *d = blockDiag(w, w);
    if (d[15] != 0.0) _Zc = d[14] / d[15];
  }
  _diagValid = false;
}

void __interaction::write(FILE *f) {
  std::fwrite(&_nWake, sizeof(size_t), 1, f);
  // interaction.cpp:308-318 writes the int maxnumparticles with
  // sizeof(size_t); readers take the first 4 bytes and skip 4 (readWake.py
  // :94-96).  Written here as a zero-extended 8-byte value.
  if (!_W.empty()) {
    const int64_t maxn = static_cast<int64_t>(_W[0]->_maxnumParticles);
    std::fwrite(&maxn, sizeof(int64_t), 1, f);
  }
  for (__wake *w : _W) w->write(f);
}

void __interaction::writenu(FILE *fdiag) { std::fwrite(&_nu, sizeof(double), 1, fdiag); }

void __interaction::writediagnosis(FILE *fdiag) {
  std::fwrite(_totalVorticity->ptr(), sizeof(double), 3, fdiag);
  std::fwrite(_linearImpulse->ptr(), sizeof(double), 3, fdiag);
  std::fwrite(_angularImpulse->ptr(), sizeof(double), 3, fdiag);
  std::fwrite(&_helicity, sizeof(double), 1, fdiag);
  std::fwrite(&_enstrophy, sizeof(double), 1, fdiag);
  std::fwrite(&_enstrophyF, sizeof(double), 1, fdiag);
  std::fwrite(&_kineticEnergy, sizeof(double), 1, fdiag);
  std::fwrite(&_kineticEnergyF, sizeof(double), 1, fdiag);
  std::fwrite(&_Zc, sizeof(double), 1, fdiag);
}

// state-vector views per wake (interaction.cpp:337-365)
static vec subvector(const vec *v, size_t off, size_t len) {
  vec s(len);
  for (size_t i = 0; i < len; ++i) s.data[i] = v->data[off + i];
  return s;
}

void __interaction::setStates(const vec *state) {
  size_t offset = 0;
  for (__wake *w : _W) {
    vec st = subvector(state, offset, w->_maxsize);
    w->setStates(&st);
    offset += w->_maxsize;
  }
  _deviceStale = true;
}

void __interaction::getRates(vec *rate) {
  size_t offset = 0;
  for (__wake *w : _W) {
    vec rt = subvector(rate, offset, w->_maxsize);
    w->getRates(&rt);
    for (size_t i = 0; i < w->_maxsize; ++i) rate->data[offset + i] = rt.data[i];
    offset += w->_maxsize;
  }
}

void __interaction::getStates(vec *state) {
  size_t offset = 0;
  for (__wake *w : _W) {
    vec st = subvector(state, offset, w->_maxsize);
    w->getStates(&st);
    for (size_t i = 0; i < w->_maxsize; ++i) state->data[offset + i] = st.data[i];
    offset += w->_maxsize;
  }
}

// interaction.cpp:480-487: each wake's ideal rates in its block
void __interaction::getIdealRates(vec *rate) {
  size_t offset = 0;
  for (__wake *w : _W) {
    vec rt = subvector(rate, offset, w->_maxsize);
    w->getIdealRates(&rt);
    for (size_t i = 0; i < w->_maxsize; ++i) rate->data[offset + i] = rt.data[i];
    offset += w->_maxsize;
  }
}

void __interaction::getVi(const vec *r, vec *vi, const size_t & /*n*/) {
  pushState();
  double u[3];
  if (pw_velocity_at(_ctx, r->ptr(), 1, u) != PW_OK) fail("pw_velocity_at");
  for (int k = 0; k < 3; ++k) vi->data[k] += u[k];
}

// ---------------------------------------------------------------------------
// __io (src/io/io.cpp)
// ---------------------------------------------------------------------------

static std::string data_root() {
  const char *r = std::getenv("PAWAN_DATA_DIR");
  std::string s = r ? r : ".";
  if (!s.empty() && s.back() != '/') s += '/';
  return s;
}

__io::__io() : _root(data_root()), _name("temp") {}
__io::__io(std::string name) : _root(data_root()), _name(name + "Wake") {}
void __io::print() { std::printf("%s\n%s\n", _root.c_str(), _name.c_str()); }
std::string __io::getFile() { return _root + _name; }

FILE *__io::create_binary_file(std::string suffix) {
  std::string fn = _root + _name + suffix;
  FILE *f = std::fopen(fn.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot create " + fn);
  return f;
}

FILE *__io::open_binary_file(std::string suffix) {
  std::string fn = _root + _name + suffix;
  FILE *f = std::fopen(fn.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot open " + fn);
  return f;
}

// ---------------------------------------------------------------------------
// __integration, __rk4, __rk3 (src/integration/)
// ---------------------------------------------------------------------------

__integration::__integration(const double &t, const size_t &n) : _dt(t / n), _t(t), _n(n) {}
__integration::__integration() : _dt(0.0), _t(0.0), _n(0) {}

void __integration::integrate(__system *S, __io *IO, bool diagnose) {
  vec states(S->_totalmaxsize);
  FILE *f = IO->create_binary_file(".wake");
  FILE *fdiag = IO->create_binary_file(".diagnosis");
  double t = 0.0;
  std::fwrite(&t, sizeof(double), 1, f);
  S->write(f);
  if (diagnose) {
    S->writenu(fdiag);
    std::fwrite(&t, sizeof(double), 1, fdiag);
    S->diagnose();
    S->writediagnosis(fdiag);
  }
  const auto t0 = std::chrono::steady_clock::now();
  for (size_t i = 1; i <= _n; ++i) {
    S->getStates(&states);
    t = i * _dt;
    step(_dt, S, &states);
    std::fwrite(&t, sizeof(double), 1, f);
    S->write(f);
    if (diagnose) {
      S->diagnose();
      std::fwrite(&t, sizeof(double), 1, fdiag);
      S->writediagnosis(fdiag);
    }
  }
  std::fclose(f);
  std::fclose(fdiag);
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("Total Time (s) %g\n", secs);
}

// integration.cpp:126-136
void __integration::step(const double &dt, __system *S, vec *states) {
  if (S->stepOnDevice(PW_EULER, dt, states)) return;
  vec rates(S->_totalmaxsize);
  S->solve();
  S->getRates(&rates);
  for (size_t i = 0; i < rates.size; ++i) rates.data[i] *= dt;
  for (size_t i = 0; i < states->size; ++i) states->data[i] += rates.data[i];
  S->setStates(states);
}

__rk4::__rk4(const double &t, const size_t &n) : __integration(t, n) {}

// rk4.cpp:12-77
void __rk4::step(const double &dt, __system *S, vec *states) {
  if (S->stepOnDevice(PW_RK4, dt, states)) return;
  const size_t L = states->size;
  vec x1(L), k1(L), k2(L), k3(L), k4(L);
  double *y = states->ptr();
  S->setStates(states);
  S->solve();
  S->getRates(&k1);
  for (size_t i = 0; i < L; ++i) x1.data[i] = k1.data[i] * (0.5 * dt) + y[i];
  S->setStates(&x1);
  S->solve();
  S->getRates(&k2);
  for (size_t i = 0; i < L; ++i) x1.data[i] = k2.data[i] * (0.5 * dt) + y[i];
  S->setStates(&x1);
  S->solve();
  S->getRates(&k3);
  for (size_t i = 0; i < L; ++i) x1.data[i] = k3.data[i] * dt + y[i];
  S->setStates(&x1);
  S->solve();
  S->getRates(&k4);
  const double c6 = dt / 6., c3 = dt / 3.;
  for (size_t i = 0; i < L; ++i)
    y[i] = y[i] + ((k1.data[i] + k4.data[i]) * c6 + (k2.data[i] + k3.data[i]) * c3);
  S->setStates(states);
}

__rk3::__rk3(const double &t, const size_t &n) : __integration(t, n) {}

// rk3.cpp:12-65
void __rk3::step(const double &dt, __system *S, vec *states) {
  if (S->stepOnDevice(PW_RK3, dt, states)) return;
  const size_t L = states->size;
  vec x1(L), k1(L), k2(L), k3(L);
  double *y = states->ptr();
  S->setStates(states);
  S->solve();
  S->getRates(&k1);
  for (size_t i = 0; i < L; ++i) x1.data[i] = k1.data[i] * (0.5 * dt) + y[i];
  S->setStates(&x1);
  S->solve();
  S->getRates(&k2);
  for (size_t i = 0; i < L; ++i) {
    const double x2 = k2.data[i] * (2. * dt) + y[i];
    k1.data[i] = k1.data[i] * dt;
    x1.data[i] = x2 - k1.data[i];
  }
  S->setStates(&x1);
  S->solve();
  S->getRates(&k3);
  const double c6 = dt / 6.0, s6 = 1.0 / 6.0, c23 = dt * 2.0 / 3.0;
  for (size_t i = 0; i < L; ++i)
    y[i] = y[i] + ((k1.data[i] * s6 + k3.data[i] * c6) + k2.data[i] * c23);
  S->setStates(states);
}

// ---------------------------------------------------------------------------
// __vring, __resolve
// ---------------------------------------------------------------------------
__vring::__vring(const double &radius, const double &core, const int &nLayers, const int &nAzi,
                 const double &sig)
    : __wake(nAzi * (2 * nLayers + 1) * (2 * nLayers + 1)) {
  _R = radius;
  _sigmaR = core;
  _nLayers = static_cast<size_t>(nLayers);
  _nAzi = static_cast<size_t>(nAzi);
  _dPsi = 2.0 * M_PI / static_cast<double>(nAzi);
  const size_t m = (2 * _nLayers + 1) * (2 * _nLayers + 1);
  _size = m;  // the state is the cell strengths (vring.cpp:62, 13)
  _maxsize = m;
  _strength = new vec(m);
  if (pw_vring_particles(radius, core, nLayers, nAzi, sig, _position->ptr(), _radius->ptr(),
                         _volume->ptr()) != PW_OK)
    throw std::runtime_error(std::string("__vring: ") + pw_last_error());
  for (size_t i = 0; i < _numParticles; ++i) _birthstrength->set(i, 1.0);  // vring.cpp:81, 98
  set_vorticity();
}

__vring::~__vring() { delete _strength; }

void __vring::set_vorticity() {
  const size_t m = _strength->size;
  size_t p = 0;
  for (size_t a = 0; a < _nAzi; ++a) {
    const double psi = static_cast<double>(a) * _dPsi, sp = std::sin(psi), cp = std::cos(psi);
    for (size_t i = 0; i < m; ++i, ++p) {
      _vorticity->set(p, 0, -sp * _strength->get(i));
      _vorticity->set(p, 1, cp * _strength->get(i));
      _vorticity->set(p, 2, 0.0);
    }
  }
}

void __vring::setStates(const vec *state) {
  for (size_t i = 0; i < _strength->size; ++i) _strength->set(i, state->get(i));
  set_vorticity();
}

void __vring::getRates(vec *rate) {
  for (size_t i = 0; i < _size; ++i) {
    const double ax = _vorticityfield->get(i, 0), ay = _vorticityfield->get(i, 1);
    rate->set(i, std::sqrt(ax * ax + ay * ay));
  }
}

void __vring::getStates(vec *state) {
  for (size_t i = 0; i < _strength->size; ++i) state->set(i, _strength->get(i));
}

void __vring::getIdealRates(vec *rate) {
  const double mag0 = 0.5 / (M_PI * _sigmaR * _sigmaR);
  const double r0 = 0.35 / (2 * static_cast<double>(_nLayers) + 1);  // vring.cpp:227
  size_t idx = 0;
  rate->set(idx++, mag0);
  for (size_t j = 1; j <= _nLayers; ++j) {
    const double jj = static_cast<double>(j);
    const double r = r0 * (12. * jj * jj + 1.0) / (6. * jj);
    const size_t ncore = 8 * j;
    const double dth = 2. * M_PI / static_cast<double>(ncore);
    const double decay = std::exp(-0.5 * r * r / (_sigmaR * _sigmaR));
    for (size_t k = 0; k < ncore; ++k)
      rate->set(idx++, mag0 * (1.0 + r * std::cos(static_cast<double>(k) * dth) / _R) * decay);
  }
}

// A x = b by LU with partial pivoting, the row of largest |a_ij| at or below
// the diagonal (first on ties), as gsl_linalg_LU_decomp / LU_solve; A is
// overwritten.  Returns false if a pivot is zero or not finite.
static bool lu_solve(size_t n, std::vector<double> &A, const std::vector<double> &b, std::vector<double> &x) {
  std::vector<size_t> perm(n);
  for (size_t i = 0; i < n; ++i) perm[i] = i;
  bool ok = true;
  for (size_t j = 0; j < n; ++j) {
    size_t ip = j;
    double mx = std::fabs(A[j * n + j]);
    for (size_t i = j + 1; i < n; ++i)
      if (std::fabs(A[i * n + j]) > mx) {
        mx = std::fabs(A[i * n + j]);
        ip = i;
      }
    if (ip != j) {
      for (size_t c = 0; c < n; ++c) std::swap(A[j * n + c], A[ip * n + c]);
      std::swap(perm[j], perm[ip]);
    }
    const double piv = A[j * n + j];
    if (!(std::fabs(piv) > 0.0) || !std::isfinite(piv)) {
      ok = false;
      continue;
    }
    for (size_t i = j + 1; i < n; ++i) {
      const double l = A[i * n + j] / piv;
      A[i * n + j] = l;
      for (size_t c = j + 1; c < n; ++c) A[i * n + c] -= l * A[j * n + c];
    }
  }
  x.resize(n);
  for (size_t i = 0; i < n; ++i) x[i] = b[perm[i]];
  for (size_t i = 0; i < n; ++i)
    for (size_t c = 0; c < i; ++c) x[i] -= A[i * n + c] * x[c];
  for (size_t i = n; i-- > 0;) {
    for (size_t c = i + 1; c < n; ++c) x[i] -= A[i * n + c] * x[c];
    x[i] /= A[i * n + i];
  }
  return ok;
}

// resolve.cpp:12-63: column i of the influence matrix is the rate vector of
// the unit state e_i after resolve(), every unit state is written to the
// .wakeinfluence file after the initial one, and the ideal rates are solved
// for.  A single __vring whose influence blocks are the base class's takes
// the matrix (and the solve) from the device in one call
// (pw_vring_resolve); the file records are written from the host state.
void __resolve::rebuild(__interaction *IN, __io *IO) {
  const size_t m = IN->_size;  // resolve.cpp:13-14
  vec states(IN->_totalmaxsize);
  std::vector<double> M(m * m);
  FILE *f = IO ? IO->create_binary_file(".wakeinfluence") : nullptr;
  double t = 0.0;
  if (f) {
    std::fwrite(&t, sizeof(double), 1, f);
    IN->write(f);
  }
  IN->getStates(&states);
  __vring *ring = IN->_W.size() == 1 ? dynamic_cast<__vring *>(IN->_W[0]) : nullptr;
  bool device = false;
  vec col(IN->_totalmaxsize);
  for (size_t i = 0; i < m; ++i) {
    states.set(i, 1.0);
    IN->setStates(&states);
    if (!device) {
      IN->resolve();
      IN->getRates(&col);
      for (size_t p = 0; p < m; ++p) M[p * m + i] = col.get(p);
      // the first column shows whether the influence blocks are the base
      // class's; then the device builds the whole matrix
      if (i == 0 && ring && IN->_influenceHooks == 0) device = true;
    }
    states.set(i, 0.0);
    t += 1.0;
    if (f) {
      std::fwrite(&t, sizeof(double), 1, f);
      IN->write(f);
    }
  }
  if (f) std::fclose(f);
  vec ideal(IN->_totalmaxsize);
  IN->getIdealRates(&ideal);
  std::vector<double> s(m);
  if (device) {
    pw_ctx *ctx = IN->context();
    if (pw_vring_resolve(ctx, 0, static_cast<int>(ring->_nLayers), static_cast<int>(ring->_nAzi), ring->_R,
                         ring->_sigmaR, s.data(), M.data()) != PW_OK)
      IN->fail("pw_vring_resolve");
  } else {
    std::vector<double> b(ideal.data.begin(), ideal.data.begin() + static_cast<std::ptrdiff_t>(m)), x;
    if (!lu_solve(m, M, b, x)) throw std::runtime_error("__resolve::rebuild: singular influence matrix");
    s = x;
  }
  vec result(IN->_totalmaxsize);
  for (size_t i = 0; i < m; ++i) result.set(i, s[i]);
  IN->setStates(&result);  // resolve.cpp:58
}

}  // namespace pawan

