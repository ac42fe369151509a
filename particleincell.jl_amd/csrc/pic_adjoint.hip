// pic_adjoint.hip -- adjoint (reverse-time) sweep of the particle-mesh step.
#include "pic_internal.h"
