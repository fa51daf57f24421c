/* Accessors around the reference's xoshiro128starstar.c (compiled from the
 * reference tree by oracle/Makefile; the source itself is never copied). */
#include REF_SRC
void ref_xoshiro128ss_seed(const uint32_t seed[4]) { s[0] = seed[0]; s[1] = seed[1]; s[2] = seed[2]; s[3] = seed[3]; }
uint32_t ref_xoshiro128ss_next(void) { return next(); }
