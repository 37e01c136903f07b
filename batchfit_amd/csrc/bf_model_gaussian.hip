// bf_model_gaussian.hip — kernel instantiations and launchers for model id 0.
#define BF_MODEL 0
#include "bf_instantiate.hpp"
