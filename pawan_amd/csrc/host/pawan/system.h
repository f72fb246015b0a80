// system.h -- pawan::__system, the abstract solver the integrators drive
// (reference: src/system/system.h:20-90).
#pragma once

#include <cstddef>
#include <cstdio>

#include "pw_vector.h"

namespace pawan {

class __system {
 public:
  size_t _size = 0;          /*!< size of the state vector               */
  size_t _totalmaxsize = 0;  /*!< sum over wakes of 6 maxN               */

  __system() = default;
  virtual ~__system() = default;

  virtual void solve() {}
  virtual void write(FILE *f) {}
  virtual void writediagnosis(FILE *fdiag) {}
  virtual void writenu(FILE *fdiag) {}
  virtual void setStates(const vec *state) {}
  virtual void getRates(vec *rate) {}
  virtual void getStates(vec *state) {}
  virtual void diagnose() {}
  //! particle management of the wakes (system.h:75); nothing by default
  virtual void relax(size_t &stepnum) {}

  //! Advance the system by one step of `scheme` (1 Euler, 3 RK3, 4 RK4)
  //! without leaving the device; `state` is the integrator's state vector,
  //! read before and written after the step.  Returns false when the
  //! system cannot do this (the integrator then falls back to the
  //! setStates/solve/getRates sequence of the reference).  No counterpart
  //! in the reference: this is the hook the GPU system uses to keep all
  //! stages of a step device-resident.
  virtual bool stepOnDevice(int scheme, const double &dt, vec *state) { return false; }
};

}  // namespace pawan
