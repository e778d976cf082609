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

}  // namespace solar
