// placeholder (replaced by jacobian.hip / solver.hip)
#include "ctx.hpp"
namespace pdes {
int jacobian_free(pdes_ctx*) { return PDES_OK; }
int precond_free(pdes_ctx*) { return PDES_OK; }
}
