// Serving engine (continuous batching): under construction.
#include "fn2_backend.h"

extern "C" {
FASTNOISE_API void* fn2ServeCreate( size_t, int, int ) { fn2::SetError( "serving engine not built" ); return nullptr; }
FASTNOISE_API void fn2ServeDestroy( void* ) {}
FASTNOISE_API int fn2ServeAddTree( void*, const char* ) { return -1; }
FASTNOISE_API int64_t fn2ServeSubmit( void*, int, const fn2GenSpec*, float* ) { return -1; }
FASTNOISE_API int fn2ServeWait( void*, int64_t, float*, double* ) { return 1; }
FASTNOISE_API int fn2ServeDrain( void* ) { return 1; }
FASTNOISE_API int fn2ServeStats( void*, int64_t*, int64_t*, int64_t* ) { return 1; }
}
