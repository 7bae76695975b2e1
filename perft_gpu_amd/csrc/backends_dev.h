// backends_dev.h -- builds a slider backend object from the device table blob.
#pragma once
#include "sliders.h"

namespace pf {

template <int B>
struct SlOf;
template <>
struct SlOf<kKS> {
    using T = KSBackend;
    __device__ static T make(const SliderTables*) { return T{}; }
};
template <>
struct SlOf<kPlain> {
    using T = PlainBackend;
    __device__ static T make(const SliderTables* t) { return T{t->plain_b, t->plain_r, t->plain_btab, t->plain_rtab}; }
};
template <>
struct SlOf<kFancy> {
    using T = FancyBackend;
    __device__ static T make(const SliderTables* t) { return T{t->fancy_b, t->fancy_r, t->fancy_tab}; }
};
template <>
struct SlOf<kByte> {
    using T = ByteBackend;
    __device__ static T make(const SliderTables* t) { return T{t->fancy_b, t->fancy_r, t->byte_tab, t->byte_r, t->byte_b}; }
};

}  // namespace pf
