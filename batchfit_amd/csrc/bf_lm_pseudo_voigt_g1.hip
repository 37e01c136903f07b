// bf_lm_pseudo_voigt_g1.hip — LM kernels for model id 2, 1 lane(s) per fit.
#define BF_MODEL 2
#define BF_GROUP 1
#include "bf_instantiate.hpp"
