// integration.h -- pawan::__integration (explicit Euler and the driver
// loop), pawan::__rk4, pawan::__rk3 and pawan::__io (reference:
// src/integration/integration.h:23-70, rk4.h:14-41, rk3.h:14-41,
// src/io/io.h:17-60).
//
// step() is the reference's: when the system can advance on the device
// (__interaction can), the whole step runs there with the stage arithmetic
// of the reference; otherwise the setStates / solve / getRates sequence of
// the reference is run on the host.
#pragma once

#include <string>

#include "system.h"

namespace pawan {

class __io {
 private:
  std::string _root;
  std::string _name;

 public:
  //! Root directory from $PAWAN_DATA_DIR (default: the working directory);
  //! the reference hard-codes a user path (io.cpp:14).
  __io();
  //! file stem `name` + "Wake" (io.cpp:18-23)
  explicit __io(std::string name);
  virtual ~__io() = default;
  virtual void print();
  virtual std::string getFile();
  virtual FILE *create_binary_file(std::string suffix);
  virtual FILE *open_binary_file(std::string suffix);
};

class __integration {
 protected:
  double _dt;  /*!< time step   */
  double _t;   /*!< total time  */
  size_t _n;   /*!< step count  */
  virtual void step(const double &dt, __system *S, vec *state);

 public:
  //! dt = t / n (integration.cpp:11-16)
  __integration(const double &t, const size_t &n);
  __integration();
  virtual ~__integration() = default;
  //! The uncoupled driver loop (integration.cpp:24-60): writes
  //! <file>.wake every step and, when diagnose is set, <file>.diagnosis.
  void integrate(__system *S, __io *IO, bool diagnose = false);
};

class __rk4 : public __integration {
 protected:
  virtual void step(const double &dt, __system *S, vec *state) override;

 public:
  __rk4(const double &t, const size_t &n);
};

class __rk3 : public __integration {
 protected:
  virtual void step(const double &dt, __system *S, vec *state) override;

 public:
  __rk3(const double &t, const size_t &n);
};

}  // namespace pawan
