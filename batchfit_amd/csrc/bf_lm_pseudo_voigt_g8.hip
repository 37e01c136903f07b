// bf_lm_pseudo_voigt_g8.hip — LM kernels for model id 2, 8 lane(s) per fit.
#define BF_MODEL 2
#define BF_GROUP 8
#include "bf_instantiate.hpp"
