// bf_lm_pseudo_voigt_g128.hip — LM kernels for model id 2, a workgroup of 128 lanes per fit.
#define BF_MODEL 2
#define BF_GROUP 128
#include "bf_instantiate.hpp"
