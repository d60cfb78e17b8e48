// res_tri_diage_p3.hip -- residual kernel instantiations for: 2,18,5,5,true
#include "residual_launch_impl.hpp"
PDES_INSTANTIATE_SHAPE(2,18,5,5,true)
