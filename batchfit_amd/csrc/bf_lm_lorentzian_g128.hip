// bf_lm_lorentzian_g128.hip — LM kernels for model id 1, a workgroup of 128 lanes per fit.
#define BF_MODEL 1
#define BF_GROUP 128
#include "bf_instantiate.hpp"
