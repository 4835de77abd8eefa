// lib.hpp -- single include of the drop-in C++ API (reference: lib/include/lib.hpp).
//
// Link with -lvectorizedadjoint_amd -lamdhip64.  The reference's examples
// build against it after two edits: include this header instead of
// lib/include/lib.hpp, and drop the Boost include and its
// `using namespace boost::numeric::odeint` (the stepper names live in
// vectorizedadjoint::odeint).
#pragma once

#include "odeint.hpp"
#include "systems.hpp"
#include "Driver.hpp"
#include "runge_kutta.hpp"
#include "backpropagation.hpp"
#include "adjoint_function.hpp"
