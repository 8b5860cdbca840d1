// host_internal.h -- helpers shared by the host and HIP translation units.
#pragma once
#include <string>

#include "aerofx.h"

namespace aerofx {
int set_error(int code, const std::string &msg);
void freestream_state(const aerofx_params &p, double U[5]);
void face_state(const double Unode[5], double Ub[5]);
}  // namespace aerofx
