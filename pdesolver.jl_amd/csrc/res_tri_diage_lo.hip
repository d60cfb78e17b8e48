// res_tri_diage_lo.hip -- residual kernel instantiations for: 2,7,3,3,true;2,12,4,4,true
#include "residual_launch_impl.hpp"
PDES_INSTANTIATE_SHAPE(2,7,3,3,true)
PDES_INSTANTIATE_SHAPE(2,12,4,4,true)
