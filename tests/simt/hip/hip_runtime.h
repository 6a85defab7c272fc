// Host build of the library for CPU tests: the HIP names it uses, from the
// fiber-based execution model in tests/simt/simt.hpp.
#pragma once
#include "../simt.hpp"
