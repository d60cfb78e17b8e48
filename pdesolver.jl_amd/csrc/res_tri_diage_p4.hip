// res_tri_diage_p4.hip -- residual kernel instantiations for: 2,27,6,6,true
#include "residual_launch_impl.hpp"
PDES_INSTANTIATE_SHAPE(2,27,6,6,true)
