// bf_model_lorentzian.hip — kernel instantiations and launchers for model id 1.
#define BF_MODEL 1
#include "bf_instantiate.hpp"
