/* simlib/sim_lib.h -- drop-in for the reference's simlib/sim_lib.h: a caller
 * that includes "simlib/sim_lib.h" (sim.cpp:66, main.cpp:102) gets
 * SimLib::EkgSim, SimLib::ActionPotential, WohlfartPlus, ColumnFile and
 * SimLib::exportVectors from this library's C++ facade, and links
 * libekgsim.so. */
#include "../../ekgsim_amd/csrc/sim_lib.hpp"
