// Helpers shared by the C-ABI translation units (not part of the ABI).
#pragma once
#include <hip/hip_runtime_api.h>

#include <string>
#include <vector>

#include "solar_kernels.hpp"
#include "solar_tables.hpp"

namespace solar {

// Records msg as this thread's last error (solar_last_error) and returns code.
int set_error(int code, const std::string& msg);
void clear_error();

// Device copies of the time-invariant tables; every allocation is appended to
// *allocs and its size added to *bytes.
int upload_tables(const Tables& t, DevTables* d, std::vector<void*>* allocs, size_t* bytes);

// The static device scratch of solar_compute_rhs_noalloc (solar_noalloc.hip):
// its size, and its address on `device` (the current device), looked up once
// per device.
size_t scratch_bytes();
hipError_t scratch_address(int device, void** addr);

}  // namespace solar
