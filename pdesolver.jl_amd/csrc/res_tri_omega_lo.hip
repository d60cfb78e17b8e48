// res_tri_omega_lo.hip -- residual kernel instantiations for: 2,3,2,3,false;2,6,3,6,false
#include "residual_launch_impl.hpp"
PDES_INSTANTIATE_SHAPE(2,3,2,3,false)
PDES_INSTANTIATE_SHAPE(2,6,3,6,false)
