#ifndef FASTNOISE_EXPORT_H
#define FASTNOISE_EXPORT_H
// Symbol visibility for the shared library (same macro as FastNoise2's Utility/Export.h).
#if defined( _WIN32 ) && !defined( FASTNOISE_STATIC_LIB )
#   if defined( FASTNOISE_EXPORT )
#       define FASTNOISE_API __declspec( dllexport )
#   else
#       define FASTNOISE_API __declspec( dllimport )
#   endif
#else
#   define FASTNOISE_API __attribute__( ( visibility( "default" ) ) )
#endif
#endif
