// roctx for the host build of the library: the named ranges around its
// public calls do nothing here (the library proper links ROCm's roctx).
#pragma once

inline int roctxRangePushA(const char*) { return 0; }
inline int roctxRangePop() { return 0; }
