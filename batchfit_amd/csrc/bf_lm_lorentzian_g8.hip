// bf_lm_lorentzian_g8.hip — LM kernels for model id 1, 8 lane(s) per fit.
#define BF_MODEL 1
#define BF_GROUP 8
#include "bf_instantiate.hpp"
