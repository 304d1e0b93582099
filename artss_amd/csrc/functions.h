// Initial conditions of the reference (Functions.cpp, dispatched by
// SolverController::set_up_fields, SolverController.cpp:94-205), evaluated
// on the host with libm in the reference's expression order, so the fields a
// run starts from are bit-identical to the reference's.  Cell midpoints are
// X1 + (i - 0.5) dx (GlobalMacrosTypes.h:76).
#pragma once

#include <map>
#include <string>
#include <vector>

#include "domain.h"

namespace artss {

struct InitialFields {
    // full level-0 arrays to upload, by field name ("u", "v", "w", "p", "T", "rhs")
    std::map<std::string, std::vector<double>> arrays;
    // fields set to a constant on every cell afterwards (Field::set_value)
    std::map<std::string, double> constants;
};

// the initial condition the XML names (usr_fct and its parameters); throws
// config_error for a function the reference cannot run either (it has no
// parsing implementation for it) or one this library does not carry
InitialFields initial_fields(const Domain &d);

}  // namespace artss
