// CabanaBondedDEM_Integrator.hpp -- BondedDEMIntegrator with the
// reference's API (src/CabanaBondedDEM_Integrator.hpp:15-106).  The
// reference's class is a copy of the plain integrator (stage1/stage3 half
// kick with 1/m and 1/I per use, stage2 drift); here it is the same class.
#pragma once

#include "CabanaDEM_Integrator.hpp"

namespace CabanaBondedDEM {

template <class ExecutionSpace>
using BondedDEMIntegrator = CabanaDEM::Integrator<ExecutionSpace>;

} // namespace CabanaBondedDEM
