// shapes.hpp -- operator shapes compiled for gfx950: triangle Omega p=1..4,
// triangle diagonal-E p=1..4, tetrahedron Omega p=1..2.
//   X(dim, Np, Nf, stencil, diagonal-E)
#pragma once
#define PDES_FOR_EACH_SHAPE(X)          \
  X(2, 3, 2, 3, false)                  \
  X(2, 6, 3, 6, false)                  \
  X(2, 10, 4, 10, false)                \
  X(2, 15, 5, 15, false)                \
  X(2, 7, 3, 3, true)                   \
  X(2, 12, 4, 4, true)                  \
  X(2, 18, 5, 5, true)                  \
  X(2, 27, 6, 6, true)                  \
  X(3, 4, 3, 4, false)                  \
  X(3, 10, 6, 10, false)
