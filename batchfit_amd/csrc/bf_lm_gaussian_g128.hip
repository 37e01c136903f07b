// bf_lm_gaussian_g128.hip — LM kernels for model id 0, a workgroup of 128 lanes per fit.
#define BF_MODEL 0
#define BF_GROUP 128
#include "bf_instantiate.hpp"
