// Fast-family (FMA contraction, reciprocals) instantiations.
#define FVM_FAST 1
#include "fvm_k_visc.hip"
