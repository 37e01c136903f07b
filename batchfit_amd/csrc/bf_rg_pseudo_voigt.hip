// bf_rg_pseudo_voigt.hip — ragged-batch kernels for model id 2.
#define BF_MODEL 2
#include "bf_ragged_instantiate.hpp"
