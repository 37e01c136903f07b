// bf_lm_pseudo_voigt_g16.hip — LM kernels for model id 2, 16 lane(s) per fit.
#define BF_MODEL 2
#define BF_GROUP 16
#include "bf_instantiate.hpp"
