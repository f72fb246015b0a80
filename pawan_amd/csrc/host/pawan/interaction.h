// interaction.h -- pawan::__interaction and pawan::__parallel, the
// multi-wake vortex-particle system (reference:
// src/interaction/interaction.h:22-261, src/interaction/parallel.h:20-155).
//
// Same class names, constructors and virtual shapes as the reference; the
// pair loops and diagnostics run on the GPU through the C-ABI of
// include/pawan_gpu.h.  The wakes stay owned by the caller (non-owning
// pointers, interaction.cpp:44-52); their host arrays are kept in step with
// the device so that drivers written against the reference read the same
// state after every call.
#pragma once

#include <vector>

#include "system.h"
#include "wake.h"

struct pw_ctx;

namespace pawan {

class __resolve;

class __interaction : public __system {
  friend class __resolve;

 protected:
  double _nu;                       /*!< kinematic viscosity             */
  size_t _nWake;                    /*!< number of wakes                 */
  vec *_totalVorticity;             /*!< diagnostics (interaction.h:27-35) */
  vec *_linearImpulse;
  vec *_angularImpulse;
  double _enstrophy;
  double _helicity;
  double _kineticEnergy;
  double _enstrophyF;
  double _kineticEnergyF;
  double _Zc = 0.0;
  int _precision;                   /*!< PW_FP64 (default), PW_FP32C, ... */

  //! Diagnostics of one wake and of one pair of wakes (interaction.h:
  //! 41-111, interaction.cpp:489-718).  Computed on the device
  //! (pw_diagnose_wakes); diagnose() calls them per wake and per pair in the
  //! reference's order, so a subclass may replace any of them.
  virtual double calculateKineticEnergy(__wake *W);
  virtual double calculateKineticEnergyF(__wake *W);
  virtual double calculateKineticEnergy(__wake *W1, __wake *W2);
  virtual double calculateKineticEnergyF(__wake *W1, __wake *W2);
  virtual double calculateEnstrophy(__wake *W);
  virtual double calculateEnstrophyF(__wake *W);
  virtual double calculateEnstrophy(__wake *W1, __wake *W2);
  virtual double calculateEnstrophyF(__wake *W1, __wake *W2);
  virtual double calculateHelicity(__wake *W);
  virtual double calculateHelicity(__wake *W1, __wake *W2);
  virtual void calculateAngularImpulse(__wake *W, vec *A);
  virtual void calculateLinearImpulse(__wake *W, vec *I);
  virtual void calculateTotalVorticity(__wake *W, vec *O);

  //! Zero the rates, then interact(W) for every wake and interact(W1, W2)
  //! for every pair (interaction.cpp:196-210).  The blocks left to the base
  //! class run on the device, in one launch when no block is overridden.
  virtual void interact();
  //! Zero the vorticity field, then influence(W) and influence(W1, W2)
  //! likewise (interaction.cpp:212-226).
  virtual void influence();

 private:
  std::vector<__wake *> _W;
  pw_ctx *_ctx = nullptr;
  bool _deviceStale = true;   // host wake state newer than the device's
  std::vector<int> _pending;  // wake blocks the base hooks left to the device
  int _interactHooks = -1;    // -1 not yet seen, 0 all base, 1 some overridden
  int _influenceHooks = -1;   // the same for the influence blocks
  // per-block diagnostics of the current diagnose() call
  std::vector<double> _diagBlocks;
  bool _diagValid = false;
  void ensureDevice();
  void pushState();           // host wakes -> device
  void pullState();           // device -> host wakes (positions, strengths)
  void pullRates();           // device rates -> host wakes
  void addRates();            // host wake rates += device rates
  int wakeIndex(const __wake *W) const;
  void queueBlock(__wake *W1, __wake *W2);  // a block left to the device
  const double *blockDiag(__wake *W1, __wake *W2);
  [[noreturn]] void fail(const char *what) const;

  //! The pairs within one wake (interaction.cpp:262-281) and between two
  //! wakes (interaction.cpp:283-302), added to the wakes' velocity and
  //! vorticity-rate rows.  Overridable, as __parallel does in the reference
  //! (parallel.h:48-56); the base versions hand the block to the device.
  virtual void interact(__wake *W);
  virtual void interact(__wake *W1, __wake *W2);
  //! The smoothed vorticity field of one wake / between two wakes, added to
  //! the wakes' _vorticityfield rows (interaction.cpp:228-260).
  virtual void influence(__wake *W);
  virtual void influence(__wake *W1, __wake *W2);

 public:
  __interaction();
  explicit __interaction(__wake *W);
  __interaction(__wake *W1, __wake *W2);
  virtual ~__interaction();

  virtual void diagnose() override;
  //! relax every wake (interaction.cpp:367-373); the device copy follows
  virtual void relax(size_t &stepnum) override;
  virtual void addWake(__wake *W);
  virtual void solve() override;
  //! influence() (interaction.cpp:58-61)
  virtual void resolve();
  virtual void write(FILE *f) override;
  virtual void writediagnosis(FILE *fdiag) override;
  virtual void writenu(FILE *fdiag) override;
  virtual void setStates(const vec *state) override;
  virtual void getRates(vec *rate) override;
  virtual void getStates(vec *state) override;
  virtual void getIdealRates(vec *rate);
  //! Velocity induced at point r (3 doubles) by every particle of every
  //! wake, each with its own radius, added to vi (interaction.cpp:453-476);
  //! n is unused, as in the reference.  FP64 on the device (pw_velocity_at).
  virtual void getVi(const vec *r, vec *vi, const size_t &n);
  virtual bool stepOnDevice(int scheme, const double &dt, vec *state) override;

  //! The reference hard-codes nu = 1.56e-5 (interaction.cpp:15); settable.
  void setNu(double nu) { _nu = nu; }
  double getNu() const { return _nu; }
  //! Pair arithmetic of solve() and the device steps (PW_FP64 default).
  void setPrecision(int precision) { _precision = precision; }
  int getPrecision() const { return _precision; }
  //! The device context (created on first use and loaded with the current
  //! host state), for the pw_* calls without a class-API counterpart, such
  //! as the tree path's training and parameters.
  pw_ctx *context();

  // diagnostics of the last diagnose() call
  const vec *totalVorticity() const { return _totalVorticity; }
  const vec *linearImpulse() const { return _linearImpulse; }
  const vec *angularImpulse() const { return _angularImpulse; }
  double helicity() const { return _helicity; }
  double enstrophy() const { return _enstrophy; }
  double enstrophyF() const { return _enstrophyF; }
  double kineticEnergy() const { return _kineticEnergy; }
  double kineticEnergyF() const { return _kineticEnergyF; }
  double centroid() const { return _Zc; }
};

//! The reference's OpenMP system (parallel.cpp); here the same GPU path.
class __parallel : public __interaction {
 public:
  explicit __parallel(__wake *W) : __interaction(W) {}
  __parallel(__wake *W1, __wake *W2) : __interaction(W1, W2) {}
};

}  // namespace pawan
