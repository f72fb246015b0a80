// wake.h -- pawan::__wake, the particle container
// (reference: src/wake/wake.h:26-172, src/wake/wake.cpp), and the vortex
// ring pawan::__vring (src/wake/vring.h, src/wake/vring.cpp).
//
// Same public members, state-vector layout and constructors as the
// reference, with the gsl containers replaced by pawan::vec / pawan::mat.
// split / merge / relax run on the device (pw_wake_split, pw_wake_merge).
// Dymore coupling members (addParticles(PawanRecvData), bound-vortex
// effects) are not part of this library.
#pragma once

#include <cstdio>

#include "pw_vector.h"

#define MAXNUMPARTICLES 100000  // max particles of a copied wake (wake.h:22)

namespace pawan {

class __wake {
 protected:
  // wake.cpp:10-25: arrays of exactly n particles
  virtual void create_particles(const int &n);
  // wake.cpp:27-44: arrays of MAXNUMPARTICLES rows, no particles yet
  virtual void initialise_memory();

 public:
  size_t _size;             /*!< state entries of the particles, 6 N     */
  size_t _maxsize;          /*!< state-vector block of the wake, 6 maxN  */
  size_t _numParticles;     /*!< particles N                            */
  size_t _maxnumParticles;  /*!< rows allocated, maxN                   */
  //! Row of the next particle split off (wake.h:44).  The reference leaves
  //! it 0 (initialise_memory) or unset outside the Dymore coupling, whose
  //! particle emission keeps it at _numParticles; here it follows
  //! _numParticles, so split appends after the last particle.
  size_t _npidx = 0;
  bool _maxmem = false;     /*!< rows exhausted (wake.h:45; not reached here) */
  mat *_position;           /*!< N x 3                                  */
  mat *_velocity;
  mat *_vorticity;
  mat *_retvorcity;
  vec *_radius;
  vec *_volume;
  vec *_birthstrength;
  mat *_vorticityfield;
  vec *_active;

  //! One particle at the origin, radius and volume 1 (wake.cpp:46-59)
  __wake();
  //! n particles, zero positions and strengths (create_particles)
  explicit __wake(const int &n);
  //! n particles from row-major arrays (pos, alpha: n x 3; radius,
  //! volume: n).  Not in the reference, whose wakes are built by its
  //! initial-condition subclasses; this is the generic way in.
  __wake(size_t n, const double *pos, const double *alpha, const double *radius,
         const double *volume);
  //! Copy into MAXNUMPARTICLES-row arrays (wake.cpp:558-562)
  explicit __wake(__wake *W);
  //! Copy the particles of W into this wake (wake.cpp:542-556)
  virtual void addParticles(__wake *W);
  virtual ~__wake();

  //! Split particles older than 45 steps whose strength has doubled since
  //! birth (wake.cpp:61-110), on the device; throws std::runtime_error when
  //! the new particles do not fit in _maxnumParticles rows (the reference
  //! wraps and overwrites its first particles)
  virtual void split(size_t &stepnum);
  //! Merge close, misaligned particles (wake.cpp:112-159), on the device
  virtual void merge(size_t &stepnum);
  //! split, then merge (wake.cpp:161-165)
  virtual void relax(size_t &stepnum);

  virtual void print();
  //! [N][pos][alpha][radius][volume][birthstrength] (wake.cpp:821-830)
  virtual void save(FILE *f);
  //! [N][pos][alpha][radius][active][volume][birthstrength]
  //! (wake.cpp:832-841), the layout scripts/readWake.py reads
  virtual void write(FILE *f);

  //! state = [pos (maxN x 3) ; alpha (maxN x 3)] (wake.cpp:856-867)
  virtual void setStates(const vec *state);
  //! rate = [vel ; retvor] for the first N rows (wake.cpp:869-880)
  virtual void getRates(vec *rate);
  //! state = [pos ; alpha] for the first N rows (wake.cpp:882-893)
  virtual void getStates(vec *state);
  //! zero rates (wake.cpp:895-899)
  virtual void getIdealRates(vec *rate);

  //! add x to coordinate n of every particle (wake.cpp:901-906)
  virtual void translate(const size_t &n, const double &x);
  //! add the 3-vector x to every particle (wake.cpp:908-914)
  virtual void translate(const double *x);
  //! elemental rotation by angle a about axis n (wake.cpp:916-937)
  virtual void rotate(const size_t &n, const double &a);
};

//! The vortex-ring initial condition (src/wake/vring.h, vring.cpp:50-100):
//! nAzi azimuthal stations of (2 nLayers + 1)^2 cross-section cells, a
//! centre particle and layers of 8j particles, with zero vorticity until
//! __resolve::rebuild gives the cells their strengths.  As in the
//! reference, its state vector is the m = (2 nLayers + 1)^2 cell strengths
//! (_size = _maxsize = m, vring.cpp:10-13, 59-62): setStates sets them and
//! the particles' vorticity from them, getRates gives the in-plane
//! magnitude of the smoothed vorticity field at the first station's m
//! particles (vring.cpp:201-221), getIdealRates the Gaussian core profile
//! (vring.cpp:223-242).  Copy it into a plain __wake to integrate it
//! (src/pawan.cpp:76-77).
class __vring : public __wake {
 private:
  //! vorticity of every particle from its cell's strength (vring.cpp:25-48)
  virtual void set_vorticity();

 public:
  double _R;             /*!< ring radius                                  */
  double _sigmaR;        /*!< core radius of the ideal profile             */
  size_t _nLayers;       /*!< layers around the centre particle            */
  size_t _nAzi;          /*!< azimuthal stations                           */
  double _dPsi;          /*!< azimuthal step                               */
  vec *_strength;        /*!< (2 nLayers + 1)^2 cell strengths             */

  __vring(const double &radius, const double &core, const int &nLayers, const int &nAzi,
          const double &sig);
  virtual ~__vring();

  virtual void setStates(const vec *state) override;
  virtual void getRates(vec *rate) override;
  virtual void getStates(vec *state) override;
  virtual void getIdealRates(vec *rate) override;
};

}  // namespace pawan
