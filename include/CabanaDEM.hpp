// CabanaDEM.hpp -- umbrella header of the reference-shaped C++ API
// (mirrors src/CabanaDEM.hpp:1-16).  Link with libcabanadem_hip.so.
#ifndef CABANADEM_HPP
#define CABANADEM_HPP

#include "CabanaDEM_Core.hpp"
#include "CabanaDEM_Input.hpp"
#include "CabanaDEM_Particles.hpp"
#include "CabanaDEM_Wall.hpp"
#include "CabanaDEM_Neighbor.hpp"
#include "CabanaDEM_Force.hpp"
#include "CabanaDEM_Integrator.hpp"
#include "CabanaDEM_Solver.hpp"

#include "CabanaBondedDEM_Particles.hpp"
#include "CabanaBondedDEM_Force.hpp"
#include "CabanaBondedDEM_Solver.hpp"
#include "CabanaBondedDEM_Integrator.hpp"

#endif
