/* Accessors around the reference's splitmix64.c (compiled from the reference
 * tree by oracle/Makefile; the source itself is never copied). */
#include REF_SRC
void ref_splitmix64_seed(uint64_t seed) { x = seed; }
uint64_t ref_splitmix64_next(void) { return next(); }
