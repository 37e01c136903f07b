// bf_lm_lorentzian_g64.hip — LM kernels for model id 1, 64 lane(s) per fit.
#define BF_MODEL 1
#define BF_GROUP 64
#include "bf_instantiate.hpp"
