// The calling thread's last error message (ckf_last_error), shared by every
// C-ABI entry point of the library.
#include <string>

#include "traccc_amd/ckf.h"

namespace tamd {
std::string& last_error() {
    thread_local std::string s;
    return s;
}
}  // namespace tamd

extern "C" const char* ckf_last_error(void) { return tamd::last_error().c_str(); }
