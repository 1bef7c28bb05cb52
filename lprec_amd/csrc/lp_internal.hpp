// Declarations shared by the host translation units of liblprec_amd.
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lprec_amd.h"

namespace lp {

int padded_size(int N0, double cor);
int os_angles(int N, int Nproj);
int filter_code(const char* name);
std::vector<float> take_filter(int N, int code);

}  // namespace lp
