// resolve.h -- pawan::__resolve (reference: src/resolve/resolve.h:14-27,
// src/resolve/resolve.cpp:12-63).
#pragma once

#include "integration.h"
#include "interaction.h"

namespace pawan {

//! rebuild(IN, IO) gives the ring held by IN the cell strengths whose
//! smoothed vorticity field at its first station matches the ring's ideal
//! core profile (resolve.cpp:12-63), computed on the device
//! (pw_vring_resolve: influence matrix and pivoted LU solve), and loads the
//! resolved vorticity back into the host wake.  IN must hold exactly one
//! wake, a __vring, as in the reference's driver (src/pawan.cpp:57-70).
//! The .wakeinfluence file the reference writes through IO is not written.
class __resolve {
 public:
  __resolve() = default;
  ~__resolve() = default;
  void rebuild(__interaction *IN, __io *IO);
};

}  // namespace pawan
