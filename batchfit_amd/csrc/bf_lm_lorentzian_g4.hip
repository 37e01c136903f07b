// bf_lm_lorentzian_g4.hip — LM kernels for model id 1, 4 lane(s) per fit.
#define BF_MODEL 1
#define BF_GROUP 4
#include "bf_instantiate.hpp"
