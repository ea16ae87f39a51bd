// Error reporting for the C-ABI: every entry point returns an nsem_status and
// leaves a human-readable message in a thread-local buffer.
#pragma once
#include <string>

#include "nebulasem_amd.h"

int nsem_fail(int code, const std::string& msg);
