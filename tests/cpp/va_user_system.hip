// The Brusselator (brusselator.hpp) registered as a device system of the
// caller's with VA_DEVICE_SYSTEM, as a caller would: built by
// vectorizedadjoint_amd/build.py into tests/cpp/libva_user_system.so, whose
// va_brusselator_kernels the tests and the drop-in example pass to
// va_create_with_kernels.
#include "vectorizedadjoint/device_system.hpp"

#include "brusselator.hpp"

VA_DEVICE_SYSTEM(Brusselator, va_brusselator_kernels)
