// bf_lm_two_lorentzian_g64.hip — LM kernels for model id 3, 64 lane(s) per fit.
#define BF_MODEL 3
#define BF_GROUP 64
#include "bf_instantiate.hpp"
