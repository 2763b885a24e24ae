// Per-thread error state of the C ABI: the message of the last failed call (phx_last_error).
// Host-only, so the builder and the scene-file code link without the HIP runtime.
#include <cstdarg>
#include <cstdio>
#include <string>

#include "phx_internal.hpp"

namespace phx {

namespace {
thread_local std::string tLastError;
}

phx_status fail(phx_status status, const char* fmt, ...)
{
	char buf[1024];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);
	tLastError = buf;
	return status;
}

phx_status ok()
{
	tLastError.clear();
	return PHX_OK;
}

}  // namespace phx

extern "C" const char* phx_last_error(void) { return phx::tLastError.c_str(); }
