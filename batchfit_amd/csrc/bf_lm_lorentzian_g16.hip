// bf_lm_lorentzian_g16.hip — LM kernels for model id 1, 16 lane(s) per fit.
#define BF_MODEL 1
#define BF_GROUP 16
#include "bf_instantiate.hpp"
