// bf_lm_pseudo_voigt_g4.hip — LM kernels for model id 2, 4 lane(s) per fit.
#define BF_MODEL 2
#define BF_GROUP 4
#include "bf_instantiate.hpp"
