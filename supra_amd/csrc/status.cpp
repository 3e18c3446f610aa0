#include "status.hpp"
#include "supra_bf.h"

namespace supra_amd {

char* error_buffer() {
    static thread_local char buf[512] = {0};
    return buf;
}

}  // namespace supra_amd

extern "C" const char* supra_last_error(void) { return supra_amd::error_buffer(); }

extern "C" const char* supra_version(void) { return "supra_amd 0.1 (gfx950)"; }
