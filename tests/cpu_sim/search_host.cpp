// Host build of the device search helpers of csrc/srd_device.h (the real
// code, compiled with g++ through the host stand-in header) for CPU tests.
#include <hip/hip_runtime.h>

#include "srd_device.h"

extern "C" uint32_t host_bound_la(const uint32_t* a, uint32_t lo, uint32_t hi, uint32_t v,
                                  int upper) {
  return upper ? srd::upper_bound_la(a, lo, hi, v) : srd::lower_bound_la(a, lo, hi, v);
}

extern "C" uint32_t host_bound(const uint32_t* a, uint32_t lo, uint32_t hi, uint32_t v, int upper) {
  return upper ? srd::upper_bound(a, lo, hi, v) : srd::lower_bound(a, lo, hi, v);
}

extern "C" uint32_t host_find_slot(const uint64_t* start, uint32_t lo, uint32_t n, uint64_t x,
                                   int from) {
  return from ? srd::find_slot_from(start, lo, n, x) : srd::find_slot(start, n, x);
}
