// resolve.h -- pawan::__resolve (reference: src/resolve/resolve.h:14-27,
// src/resolve/resolve.cpp:12-63).
#pragma once

#include "integration.h"
#include "interaction.h"

namespace pawan {

//! rebuild(IN, IO) (resolve.cpp:12-63): the influence matrix of IN's state
//! vector, column i the rates after resolve() of the unit state e_i; the
//! state solving it for the ideal rates (LU with partial pivoting) is set
//! into IN.  With IO non-null, the initial and every unit state are written
//! to <file>.wakeinfluence as [t][IN->write] records (t = 0, 1, ..., m), as
//! the reference does.  For an interaction holding one __vring (the
//! reference's use, src/pawan.cpp:57-70) whose influence blocks are the
//! base class's, the matrix and the solve come from the device in one call
//! (pw_vring_resolve); otherwise every column goes through the virtual
//! resolve() / getRates(), and the solve runs on the host.
class __resolve {
 public:
  __resolve() = default;
  ~__resolve() = default;
  void rebuild(__interaction *IN, __io *IO);
};

}  // namespace pawan
