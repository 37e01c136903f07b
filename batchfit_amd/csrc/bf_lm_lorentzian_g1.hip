// bf_lm_lorentzian_g1.hip — LM kernels for model id 1, 1 lane(s) per fit.
#define BF_MODEL 1
#define BF_GROUP 1
#include "bf_instantiate.hpp"
