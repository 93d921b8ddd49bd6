// pcs_common.h -- host helpers shared by the C-ABI translation units.
#pragma once

#include <vector>

#include "pcs_internal.h"

namespace pcs {

int set_error(int code, const char* fmt, ...);
int hip_error(hipError_t e, const char* what);
int validate_system(const pcs_system* s);
int validate_proposal(const pcs_system& s, const pcs_proposal& p);
void build_sat(const uint32_t* H, uint32_t n_k, uint32_t n_de, std::vector<uint32_t>& S, uint32_t& maxrow);
DevSys make_devsys(const pcs_system& s, uint32_t maxrow, const uint32_t* dsat);
uint32_t record_capacity(const pcs_system& s, const pcs_proposal& p, const uint32_t* frames,
                         const uint32_t* particles, double factor);

}  // namespace pcs
