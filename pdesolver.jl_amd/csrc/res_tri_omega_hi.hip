// res_tri_omega_hi.hip -- residual kernel instantiations for: 2,10,4,10,false;2,15,5,15,false
#include "residual_launch_impl.hpp"
PDES_INSTANTIATE_SHAPE(2,10,4,10,false)
PDES_INSTANTIATE_SHAPE(2,15,5,15,false)
