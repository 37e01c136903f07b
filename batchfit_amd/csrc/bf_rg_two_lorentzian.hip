// bf_rg_two_lorentzian.hip — ragged-batch kernels for model id 3.
#define BF_MODEL 3
#include "bf_ragged_instantiate.hpp"
