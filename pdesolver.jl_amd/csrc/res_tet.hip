// res_tet.hip -- residual kernel instantiations for: 3,4,3,4,false;3,10,6,10,false
#include "residual_launch_impl.hpp"
PDES_INSTANTIATE_SHAPE(3,4,3,4,false)
PDES_INSTANTIATE_SHAPE(3,10,6,10,false)
