// bf_lm_pseudo_voigt_g32.hip — LM kernels for model id 2, 32 lane(s) per fit.
#define BF_MODEL 2
#define BF_GROUP 32
#include "bf_instantiate.hpp"
