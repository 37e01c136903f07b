// bf_lm_pseudo_voigt_g64.hip — LM kernels for model id 2, 64 lane(s) per fit.
#define BF_MODEL 2
#define BF_GROUP 64
#include "bf_instantiate.hpp"
