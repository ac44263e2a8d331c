// TEST INFRASTRUCTURE: a C entry point over the reference's own pcg32
// (rainbow-cpu/shared/random_generator.cpp:91-139), compiled from the
// reference tree by oracle/Makefile into oracle/_ref/libref_pcg.so.
#include "/root/reference/rainbow-cpu/shared/random_generator.hpp"

extern "C" void ref_pcg32_reals(unsigned long long seed, unsigned n, float* out) {
    rainbow::cpus::shared::pcg32 g(seed);
    for (unsigned i = 0; i < n; i++) out[i] = g.uniform_real();
}

extern "C" void ref_pcg32_u32(unsigned long long seed, unsigned n, unsigned* out) {
    rainbow::cpus::shared::pcg32 g(seed);
    for (unsigned i = 0; i < n; i++) out[i] = g.uniform_uint32();
}

extern "C" void ref_pcg32_range(unsigned long long seed, unsigned n, const unsigned* lo, const unsigned* hi,
                                unsigned* out) {
    rainbow::cpus::shared::pcg32 g(seed);
    for (unsigned i = 0; i < n; i++) out[i] = g.uint32(lo[i], hi[i]);
}
