// bf_rg_lorentzian.hip — ragged-batch kernels for model id 1.
#define BF_MODEL 1
#include "bf_ragged_instantiate.hpp"
